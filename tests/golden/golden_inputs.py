"""Input definitions shared by make_golden.py (which records the reference's
outputs on them) and the tests (which recompute them and compare).

Inputs are reference data files (ref_inputs/, copied verbatim from the
reference's geometry/ and config/ folders) and seeded synthetic meshes; each
is fingerprinted in the fixture so a change of inputs is caught.
"""
from __future__ import annotations

import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "zedm-x_amd"))
from zdx import synth  # noqa: E402


def read_stl(path: str) -> np.ndarray:
    """Plain STL reader for test inputs (binary if 84 + 50 n == size, else ASCII)."""
    data = open(path, "rb").read()
    if len(data) >= 84:
        n = struct.unpack("<I", data[80:84])[0]
        if 84 + 50 * n == len(data):
            out = np.zeros((n, 3, 3))
            for i in range(n):
                v = struct.unpack("<12f", data[84 + 50 * i:84 + 50 * i + 48])
                out[i] = np.asarray(v[3:], dtype=np.float64).reshape(3, 3)
            return out
    tris, vb = [], []
    for line in data.decode("utf-8", "ignore").splitlines():
        s = line.lower()
        if "vertex" in s:
            p = s.split()
            try:
                vb.append([float(p[1]), float(p[2]), float(p[3])])
            except (IndexError, ValueError):
                pass
        if "endloop" in s and len(vb) >= 3:
            tris.append(vb[-3:])
    return np.asarray(tris, dtype=np.float64)


def synth_meshes() -> dict:
    return {
        "ellipsoid200": synth.ellipsoid_hull(200, (1.0, 0.6, 0.4), 1),
        "ico1": synth.icosphere(1),
        "ico2": synth.icosphere(2),
        "star2a": synth.star_mesh(7, level=2),
        "star2b": synth.star_mesh(8, level=2),
        "hull32": synth.sphere_hull(32, 3),
        "hull128": synth.sphere_hull(128, 4),
    }


def mesh_inputs(inputs_dir: str):
    out = []
    for f in ["low-poly-banbana.stl", "convex-particle.stl", "wall_plane.stl", "wall_box.stl", "table.stl"]:
        out.append((f, read_stl(os.path.join(inputs_dir, "geometry", f))))
    for k, v in synth_meshes().items():
        out.append((k, v))
    return out


def tri_pairs(seed: int, n: int):
    """Triangle pairs: mostly overlapping random ones, plus coplanar, shared-vertex
    and degenerate (collinear) cases."""
    rng = synth.Rng(seed)
    u = rng.uniform(n * 18).reshape(n, 2, 3, 3) * 2.0 - 1.0
    A = u[:, 0].copy()
    B = u[:, 1] * 0.8 + A.mean(axis=1, keepdims=True) * 0.5
    k = n // 10
    A[:k, :, 2] = 0.25
    B[:k, :, 2] = 0.25                                     # coplanar
    B[k:2 * k, 0] = A[k:2 * k, 1]                          # shared vertex
    A[2 * k:3 * k, 2] = 0.5 * (A[2 * k:3 * k, 0] + A[2 * k:3 * k, 1])   # degenerate A
    return np.ascontiguousarray(A), np.ascontiguousarray(B)


def _pose(rng, rA: float, rB: float, depth_frac: float, rotate: bool):
    d = rng.unit_vectors(1)[0]
    qA = rng.quaternions(1)[0] if rotate else np.array([1.0, 0.0, 0.0, 0.0])
    qB = rng.quaternions(1)[0] if rotate else np.array([1.0, 0.0, 0.0, 0.0])
    dist = (rA + rB) * (1.0 - depth_frac)
    return [np.zeros(3), qA, d * dist, qB]


def pair_cases(inputs_dir: str):
    """(name, raw tris A, raw tris B, [posA, quatA, posB, quatB], split)."""
    m = synth_meshes()
    banana = read_stl(os.path.join(inputs_dir, "geometry", "low-poly-banbana.stl"))
    rng = synth.Rng(21)
    I = np.array([1.0, 0.0, 0.0, 0.0])
    cases = [("cfg1_contact", m["ellipsoid200"], m["ellipsoid200"],
              [np.zeros(3), I, np.array([1.95, 0.0, 0.0]), I], True)]
    # (name, mesh A, mesh B, pose seed): pose = random unit direction, random
    # quaternions and an overlap depth of 5-35 % of the bounding radii; the
    # seeds were picked to cover multi-component, multi-loop and filtered-loop
    # cases.
    specs = [("ico1_ico2", "ico1", "ico2", 201), ("star_star_2loops", "star2a", "star2b", 203),
             ("star_star_2loops_b", "star2a", "star2b", 212), ("star_star_2comp_1loop", "star2a", "star2b", 226),
             ("star_star_c", "star2a", "star2b", 233), ("star_star_d", "star2a", "star2b", 253),
             ("ell_star_filtered", "ellipsoid200", "star2a", 257), ("ell_star", "ellipsoid200", "star2a", 205),
             ("ico2_star_2comp", "ico2", "star2b", 206), ("ico2_star_2loops", "ico2", "star2b", 215),
             ("hull32_hull128", "hull32", "hull128", 207), ("hull128_ico1", "hull128", "ico1", 209)]
    for name, a, b, seed in specs:
        rA = float(np.max(np.linalg.norm(m[a].reshape(-1, 3), axis=1)))
        rB = float(np.max(np.linalg.norm(m[b].reshape(-1, 3), axis=1)))
        r = synth.Rng(seed)
        dep = 0.05 + 0.3 * float(r.uniform(1)[0])
        cases.append((name, m[a], m[b], _pose(r, rA, rB, dep, True), True))
    cases.append(("star_star_nosplit", m["star2a"], m["star2b"], _pose(synth.Rng(212), 1.2, 1.2, 0.2, True), False))
    cases.append(("banana_banana", banana, banana, _pose(rng, 103.4, 103.4, 0.7, True), True))
    return cases


def _particle(stl, pos, vel, quat, omega, scale=1.0, density=2500.0, young=1e7, poisson=0.25, mu=0.5, rest=0.5):
    return dict(stl=stl, pos=list(map(float, pos)), vel=list(map(float, vel)), quat=list(map(float, quat)),
                omega=list(map(float, omega)), scale=float(scale), density=float(density), young=float(young),
                poisson=float(poisson), mu=float(mu), restitution=float(rest))


def sim_cases(inputs_dir: str):
    """(name, config text, {stl file name: tris}) run through zdem_cpu's main loop."""
    m = synth_meshes()
    cases = []
    g0 = dict(steps=1000, dt=1e-4, split_contacts=1, contact_debug=0, gravity=[0.0, 0.0, 0.0], center_mesh=1,
              vtk_prefix="c1", output_dir="out", output_interval=100)
    parts = [_particle("ell.stl", (0, 0, 0), (1, 0, 0), (1, 0, 0, 0), (0, 0, 0)),
             _particle("ell.stl", (1.95, 0, 0), (-1, 0, 0), (1, 0, 0, 0), (0, 0, 0))]
    cases.append(("cfg1_contact", synth.format_config(g0, parts), {"ell.stl": m["ellipsoid200"]}))

    ex = open(os.path.join(inputs_dir, "config", "example_sim.txt")).read()
    ex = ex.replace("output_interval = 100", "output_interval = 1000").replace("output_dir = output", "output_dir = out")
    cases.append(("example_sim", ex, {}))

    wt = open(os.path.join(inputs_dir, "config", "wall_test.txt")).read()
    wt = (wt.replace("steps = 200000", "steps = 4000").replace("output_interval = 1000", "output_interval = 250")
          .replace("output_dir = output", "output_dir = out").replace("pos = 0 -0.5 0.2", "pos = 0 -0.5 0.0565")
          .replace("pos = 0 0.5 0.2", "pos = 0 0.5 0.0562"))
    cases.append(("wall_drop", wt, {}))

    # cluster of mixed polyhedra inside a 12-triangle box wall, gravity on, with
    # initial overlaps and several partners per particle
    rng = synth.Rng(31)
    names = ["ico1.stl", "hull32.stl", "star2a.stl", "ell.stl", "hull128.stl"]
    files = {"ico1.stl": m["ico1"], "hull32.stl": m["hull32"], "star2a.stl": m["star2a"],
             "ell.stl": m["ellipsoid200"], "hull128.stl": m["hull128"], "box.stl": synth.box_wall(0.1, 0.1, 0.1)}
    parts = []
    k = 0
    for ix in range(3):
        for iy in range(2):
            for iz in range(2):
                pos = np.array([-0.012 + 0.012 * ix, -0.006 + 0.012 * iy, -0.043 + 0.0115 * iz])
                pos = pos + (rng.uniform(3) - 0.5) * 0.002
                vel = (rng.uniform(3) - 0.5) * 0.4
                om = (rng.uniform(3) - 0.5) * 20.0
                q = rng.quaternions(1)[0]
                parts.append(_particle(names[k % len(names)], pos, vel, q, om, scale=0.0065,
                                       rest=0.3 + 0.1 * (k % 4), mu=0.3 + 0.05 * (k % 5)))
                k += 1
    g = dict(steps=400, dt=1e-5, split_contacts=1, gravity=[0.0, 0.0, -9.81], center_mesh=1,
             vtk_prefix="cl", output_dir="out", output_interval=50)
    wall = [dict(stl="box.stl", pos=[0.0, 0.0, 0.0], quat=[1.0, 0.0, 0.0, 0.0], scale=1.0, mu=0.4, restitution=0.3)]
    cases.append(("cluster_box", synth.format_config(g, parts, wall), files))

    g2 = dict(steps=300, dt=2e-5, split_contacts=0, gravity=[0.0, 0.0, 0.0], center_mesh=1,
              vtk_prefix="ns", output_dir="out", output_interval=50)
    parts = [_particle("star2a.stl", (0, 0, 0), (0.5, 0, 0), (1, 0, 0, 0), (0, 0, 0), scale=0.01),
             _particle("star2a.stl", (0.017, 0.003, 0.001), (-0.5, 0, 0), (0.8, 0.6, 0, 0), (0, 0, 5.0), scale=0.01)]
    cases.append(("star_nosplit", synth.format_config(g2, parts), {"star2a.stl": m["star2a"]}))
    return cases


def box_pair_case(depth, off, pa, k):
    """Two boxes at rest, no gravity, minimal-energy force with stiffness k:
    A (1 x 1.2 x 0.8) centred at (0, pa), B (2 x 2 x 2) whose -x face sits at
    x = 0.5 - depth inside A, shifted by (off) in y, z.  The contact loop is
    A's 1.2 x 0.8 cross-section at x = 0.5 - depth.  Returns (config text,
    {stl name: tris}, pA, pB, rectangle centre)."""
    pA = np.array([0.0, pa[0], pa[1]])
    pB = pA + np.array([1.5 - depth, off[0], off[1]])
    g = dict(steps=1, dt=1e-6, split_contacts=1, gravity=[0.0, 0.0, 0.0], center_mesh=1, vtk_prefix="bx",
             output_dir="out", output_interval=1, force_model="minimal", minimal_k=float(k))
    parts = [_particle("boxA.stl", pA, (0, 0, 0), (1, 0, 0, 0), (0, 0, 0)),
             _particle("boxB.stl", pB, (0, 0, 0), (1, 0, 0, 0), (0, 0, 0))]
    files = {"boxA.stl": synth.box_wall(1.0, 1.2, 0.8), "boxB.stl": synth.box_wall(2.0, 2.0, 2.0)}
    centre = np.array([0.5 - depth, pa[0], pa[1]])
    return synth.format_config(g, parts), files, pA, pB, centre

