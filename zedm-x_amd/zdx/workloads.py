"""Synthetic workloads of the Route-B contact step.

granular_bed():   100k polyhedra from 8 templates in a 0.2 m box with a
                  12-triangle box wall (the step throughput workload)
uneven_pairs():   4096 independent pairs, icospheres of 80-5120 triangles and
                  concave stars (the batched per-pair workload)
drop_on_plane():  64 convex hulls dropped onto a plane wall, colliding with it
                  and with each other (the differentiable drop)
concave_packing(): 1M concave stars of 320-2560 triangles in a 0.5 m box (the
                  large single-GPU configuration)

Particles carry the zdem_cpu fields (template, scale, pos, quat, vel, omega,
material); write_reference_config() writes any subset as a zdem_cpu config
plus the template STLs, so the reference program can run the same inputs.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import synth


@dataclass
class Workload:
    name: str
    templates: list                    # raw template triangle soups (T, 3, 3)
    template_names: list
    tpl: np.ndarray                    # per particle template index
    scale: np.ndarray                  # per particle scale (multiple of 1e-6)
    init: np.ndarray                   # per particle pos(3) quat(4) vel(3) omega(3)
    material: np.ndarray               # per particle density young poisson mu restitution
    walls: list = field(default_factory=list)   # (tris, pos, quat, mu, restitution, name)
    gravity: tuple = (0.0, 0.0, -9.81)
    dt: float = 1e-5
    steps: int = 1000
    split: bool = True
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.tpl)


def _centroid(tris):
    """Volume centroid by the divergence theorem (src/main.cpp:382-428)."""
    a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
    det = np.einsum("ij,ij->i", a, np.cross(b, c))
    return (((a + b + c) * det[:, None]).sum(axis=0)) / (4.0 * det.sum())


def _centred_radius(tris):
    """Bounding radius of a template after build_mesh's centring (src/main.cpp:458-482)."""
    c = _centroid(tris)
    return float(np.max(np.linalg.norm(tris.reshape(-1, 3) - c, axis=1)))


def bed_templates(seed: int = 3):
    names, tris = [], []
    for k, n in enumerate((32, 64, 128, 256, 512)):
        names.append(f"hull{n}")
        tris.append(synth.sphere_hull(n, seed * 100 + k))
    for k, lvl in enumerate((2, 3, 3)):
        names.append(f"star{lvl}_{k}")
        tris.append(synth.star_mesh(seed * 100 + 10 + k, level=lvl))
    return names, tris


def granular_bed(n: int = 100_000, seed: int = 3, box: float = 0.2, r_min: float = 0.5e-3, r_max: float = 2.0e-3,
                 speed: float = 0.5) -> Workload:
    """Granular bed: n particles from 8 templates (convex hulls of 32-512 sphere
    points, radial-noise stars of 320-1280 triangles), radii log-uniform in
    [r_min, r_max], density 2500, young 1e7, in a box of edge `box` bounded by
    one 12-triangle box wall, gravity -9.81 z, dt 1e-5.

    Seeding: a cubic lattice of spacing 2 r_max + 0.1 mm filled from the
    bottom; each particle is jittered inside its own lattice cell so that no
    two bounding spheres overlap (non-overlapping seeding).  Velocities are
    Gaussian with `speed` m/s per component, so the bed is agitated and
    collisions start within ~100 steps.
    """
    names, tris = bed_templates(seed)
    rad_t = np.array([_centred_radius(t) for t in tris])
    g = np.random.Generator(np.random.PCG64(seed))
    tpl = g.integers(0, len(tris), n)
    r = np.exp(g.uniform(math.log(r_min), math.log(r_max), n))
    scale = np.round(r / rad_t[tpl], 6)           # reference mesh cache key resolution
    scale = np.maximum(scale, 1e-6)
    rad = scale * rad_t[tpl]
    s = 2.0 * float(rad.max()) + 1e-4
    m = int(math.floor(box / s))
    if m ** 3 < n:
        raise ValueError(f"{n} particles do not fit a lattice of spacing {s} in a box of {box}")
    k = np.arange(n)
    ix, iy, iz = k % m, (k // m) % m, k // (m * m)
    centre = (np.stack([ix, iy, iz], axis=1) + 0.5) * s + (box - m * s) / 2.0
    slack = np.maximum(s / 2.0 - rad - 2e-5, 0.0)
    pos = centre + g.uniform(-1.0, 1.0, (n, 3)) * slack[:, None]
    q = g.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    vel = g.normal(size=(n, 3)) * speed
    init = np.concatenate([pos, q, vel, np.zeros((n, 3))], axis=1)
    mat = np.tile(np.array([2500.0, 1e7, 0.25, 0.5, 0.5]), (n, 1))
    wall = (synth.box_wall(box, box, box), np.array([box / 2] * 3), np.array([1.0, 0.0, 0.0, 0.0]), 0.5, 0.5, "box")
    return Workload("granular_bed", tris, names, tpl, scale, init, mat, [wall], (0.0, 0.0, -9.81), 1e-5, 1000, True,
                    dict(seed=seed, box=box, lattice=s, r_min=r_min, r_max=r_max, speed=speed))


def _rotmat(q):
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _ray_extent(tris, d):
    """Distance from the origin to the farthest crossing of the ray t d (t > 0)
    with the triangle soup `tris` (Moller-Trumbore, all triangles at once)."""
    v0, e1, e2 = tris[:, 0], tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]
    h = np.cross(d, e2)
    a = np.einsum("ij,ij->i", e1, h)
    ok = np.abs(a) > 1e-300
    f = np.where(ok, 1.0 / np.where(ok, a, 1.0), 0.0)
    s = -v0
    u = f * np.einsum("ij,ij->i", s, h)
    q = np.cross(s, e1)
    v = f * (q @ d)
    t = f * np.einsum("ij,ij->i", e2, q)
    hit = ok & (u >= 0.0) & (v >= 0.0) & (u + v <= 1.0) & (t > 0.0)
    return float(t[hit].max())


def uneven_pairs(n_pairs: int = 4096, seed: int = 2):
    """Independent pairs: half icosphere pairs (levels 1-4 each, 80-5120
    triangles), half concave star pairs (level-3 icosphere, 1280 triangles,
    radius x (1 + 0.3 noise)); random quaternions; overlap depth uniform 1-10 %
    of the smaller bounding radius along a random direction d.  The depth is
    measured along the line of centres: A's surface is crossed by the ray
    from its centroid along d, B's by the ray along -d, and the two crossings
    overlap by the depth, so every pair interpenetrates, the stars included
    (their surface lies well inside the bounding sphere in most directions,
    and each is star-shaped about its centroid).  Positions refer to the
    volume centroid, which is where mesh registration with centring puts the
    body origin.  Returns (meshes, names, pairs) with pairs = list of
    (mesh a, mesh b, posA, quatA, posB, quatB)."""
    names = ["ico1", "ico2", "ico3", "ico4", "starA", "starB"]
    meshes = [synth.icosphere(1), synth.icosphere(2), synth.icosphere(3), synth.icosphere(4),
              synth.star_mesh(seed * 100 + 1, 3), synth.star_mesh(seed * 100 + 2, 3)]
    rad = [_centred_radius(t) for t in meshes]
    cen = [t - _centroid(t) for t in meshes]
    g = np.random.Generator(np.random.PCG64(seed))
    pairs = []
    for p in range(n_pairs):
        if p < n_pairs // 2:
            a, b = int(g.integers(0, 4)), int(g.integers(0, 4))
        else:
            a, b = 4 + int(g.integers(0, 2)), 4 + int(g.integers(0, 2))
        d = g.normal(size=3)
        d /= np.linalg.norm(d)
        qa, qb = g.normal(size=4), g.normal(size=4)
        qa /= np.linalg.norm(qa)
        qb /= np.linalg.norm(qb)
        depth = g.uniform(0.01, 0.10) * min(rad[a], rad[b])
        # the rays in the body frames: R^T d
        ha = _ray_extent(cen[a], _rotmat(qa).T @ d)
        hb = _ray_extent(cen[b], _rotmat(qb).T @ -d)
        dist = ha + hb - depth
        pairs.append((a, b, np.zeros(3), qa, d * dist, qb))
    return meshes, names, pairs


def drop_on_plane(n: int = 64, seed: int = 4, radius: float = 5e-3, height: float = 0.2e-3, gap: float = 0.2e-3,
                  speed: float = 0.1):
    """n convex hulls of 128 points (radius 5 mm) on a square grid just above
    a 2-triangle plane wall at z = 0, gravity -9.81 z, dt 1e-5, 500 steps.
    The grid spacing leaves `gap` between neighbouring bounding spheres and
    the bodies start with Gaussian horizontal velocities of `speed` m/s per
    component and -0.05 m/s downwards, so within the 500 steps they hit the
    wall and each other."""
    hull = synth.sphere_hull(128, seed)
    rt = _centred_radius(hull)
    scale = round(radius / rt, 6)
    g = np.random.Generator(np.random.PCG64(seed))
    side = int(math.ceil(math.sqrt(n)))
    k = np.arange(n)
    s = 2.0 * radius + gap
    pos = np.stack([(k % side) * s, (k // side) * s, np.full(n, radius + height)], axis=1)
    pos[:, 2] += g.uniform(0, 0.5e-3, n)
    q = g.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    vel = np.zeros((n, 3))
    vel[:, 0:2] = g.normal(size=(n, 2)) * speed
    vel[:, 2] = -0.05
    init = np.concatenate([pos, q, vel, np.zeros((n, 3))], axis=1)
    mat = np.tile(np.array([2500.0, 1e7, 0.25, 0.5, 0.5]), (n, 1))
    wall = (synth.plane_wall(1.0, 1.0), np.array([(side - 1) * s / 2, (side - 1) * s / 2, 0.0]),
            np.array([1.0, 0.0, 0.0, 0.0]), 0.5, 0.5, "plane")
    return Workload("drop_on_plane", [hull], ["hull128"], np.zeros(n, dtype=np.int64), np.full(n, scale), init, mat,
                    [wall], (0.0, 0.0, -9.81), 1e-5, 500, True, dict(seed=seed))


def concave_packing(n: int = 1_000_000, seed: int = 5, box: float = 0.5, r_min: float = 1e-3, r_max: float = 4e-3,
                    n_templates: int = 16, radius_levels: int = 256, speed: float = 0.2,
                    overlap: float = 0.12) -> Workload:
    """Large concave packing: n radial-noise stars from n_templates templates
    of 320-2560 triangles (hulls of 162-1282 seeded sphere points, each vertex
    scaled by 1 + 0.3 noise), radii log-uniform in [r_min, r_max] on
    radius_levels log-spaced values (so at most n_templates x radius_levels
    distinct meshes, the reference's path|scale cache), density 2500, young
    1e7, in a box of edge `box` bounded by one 12-triangle box wall, gravity
    -9.81 z, dt 5e-6.

    Seeding: the particles sorted by radius (largest first) fill horizontal
    layers from the bottom; a layer is a hexagonal lattice whose spacing is
    (1 - overlap) x twice the largest radius in it, so the bounding spheres of
    neighbours in a layer overlap by up to 12 % of their diameter: every
    neighbour pair is a candidate pair from the first step and the stars'
    lobes meet in about a third of them, often in several contact loops.  Gaussian velocities of
    `speed` m/s per component, random orientations."""
    names, tris = [], []
    for k, m in enumerate(np.linspace(162, 1282, n_templates).round().astype(int)):
        names.append(f"star{2 * m - 4}")
        tris.append(synth.star_hull_mesh(seed * 1000 + 17 * k, int(m)))
    rad_t = np.array([_centred_radius(t) for t in tris])
    g = np.random.Generator(np.random.PCG64(seed))
    tpl = g.integers(0, n_templates, n)
    lv = np.exp(np.linspace(math.log(r_min), math.log(r_max), radius_levels))
    r = lv[g.integers(0, radius_levels, n)]
    scale = np.maximum(np.round(r / rad_t[tpl], 6), 1e-6)
    rad = scale * rad_t[tpl]
    order = np.argsort(-rad, kind="stable")
    pos = np.zeros((n, 3))
    i, z = 0, 0.0
    while i < n:
        s = 2.0 * rad[order[i]] * (1.0 - overlap)
        dy = s * math.sqrt(3.0) / 2.0
        nx = int(math.floor((box - s) / s)) + 1
        ny = int(math.floor((box - s) / dy)) + 1
        if z == 0.0:
            z = s / 2.0
        if z + s / 2.0 > box:
            raise ValueError(f"{n} particles do not fit the box of {box} (stopped at {i})")
        row, col = np.divmod(np.arange(nx * ny), nx)
        x = s / 2.0 + col * s + np.where(row % 2 == 1, s / 2.0, 0.0)
        keep = x <= box - s / 2.0           # odd rows hold one site fewer
        x, y = x[keep], s / 2.0 + row[keep] * dy
        cnt = min(n - i, len(x))
        pos[order[i:i + cnt]] = np.stack([x[:cnt], y[:cnt], np.full(cnt, z)], axis=1)
        i += cnt
        z += s
    q = g.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    vel = g.normal(size=(n, 3)) * speed
    init = np.concatenate([pos, q, vel, np.zeros((n, 3))], axis=1)
    mat = np.tile(np.array([2500.0, 1e7, 0.25, 0.5, 0.5]), (n, 1))
    wall = (synth.box_wall(box, box, box), np.array([box / 2] * 3), np.array([1.0, 0.0, 0.0, 0.0]), 0.5, 0.5, "box")
    return Workload("concave_packing", tris, names, tpl, scale, init, mat, [wall], (0.0, 0.0, -9.81), 5e-6, 200, True,
                    dict(seed=seed, box=box, r_min=r_min, r_max=r_max, height=z, overlap=overlap))


def build_sim(ctx, wl: Workload, center: bool = True):
    """Registers the workload's meshes (one per template and scale, as the
    reference's path|scale mesh cache does) and creates the simulation."""
    from . import Sim

    cache = {}
    mesh_id = np.zeros(wl.n, dtype=np.int32)
    for i in range(wl.n):
        key = (int(wl.tpl[i]), float(wl.scale[i]))
        if key not in cache:
            cache[key] = ctx.register_mesh(wl.templates[key[0]] * key[1], center)
        mesh_id[i] = cache[key]
    walls = []
    for (tris, pos, quat, mu, rest, _name) in wl.walls:
        walls.append((ctx.register_mesh(tris, center), pos, quat, mu, rest))
    sim = Sim.create(ctx, mesh_id, wl.init, wl.material, wl.gravity, wl.dt, wl.split, walls)
    sim.mesh_id = mesh_id
    return sim, len(cache)


def write_reference_config(wl: Workload, idx, out_dir: str, steps: int, output_interval: int, state=None,
                           prefix: str = "zdx", vtk_prefix: str | None = None) -> str:
    """Writes the particles `idx` (with `state` rows pos quat vel omega if
    given) as a zdem_cpu config `<prefix>.txt` with the template STLs next to
    it; the program's output files are named by vtk_prefix (default: prefix)."""
    os.makedirs(out_dir, exist_ok=True)
    for t, name in zip(wl.templates, wl.template_names):
        p = os.path.join(out_dir, name + ".stl")
        if not os.path.exists(p):
            synth.write_stl_ascii(p, t, name)
    walls = []
    for (tris, pos, quat, mu, rest, name) in wl.walls:
        p = os.path.join(out_dir, f"wall_{name}.stl")
        if not os.path.exists(p):
            synth.write_stl_ascii(p, tris, name)
        walls.append(dict(stl=p, pos=list(map(float, pos)), quat=list(map(float, quat)), scale=1.0, mu=float(mu),
                          restitution=float(rest)))
    parts = []
    for i in idx:
        row = wl.init[i] if state is None else state[i]
        parts.append(dict(stl=os.path.join(out_dir, wl.template_names[wl.tpl[i]] + ".stl"),
                          pos=list(map(float, row[0:3])), vel=list(map(float, row[7:10])),
                          quat=list(map(float, row[3:7])), omega=list(map(float, row[10:13])),
                          scale=float(wl.scale[i]), density=float(wl.material[i, 0]), young=float(wl.material[i, 1]),
                          poisson=float(wl.material[i, 2]), mu=float(wl.material[i, 3]),
                          restitution=float(wl.material[i, 4])))
    g = dict(steps=steps, dt=wl.dt, split_contacts=int(wl.split), contact_debug=0, gravity=list(wl.gravity),
             center_mesh=1, vtk_prefix=prefix if vtk_prefix is None else vtk_prefix, output_dir=os.path.join(out_dir, "out"), output_interval=output_interval)
    path = os.path.join(out_dir, prefix + ".txt")
    with open(path, "w") as f:
        f.write(synth.format_config(g, parts, walls))
    return path
