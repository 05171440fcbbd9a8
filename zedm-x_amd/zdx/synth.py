"""Synthetic polyhedra and seeded workloads for the Route-B contact step.

Meshes are FP64 triangle soups of shape (T, 3, 3), outward-oriented, written to
ASCII STL with ``%.17g`` so that the reference loader (src/main.cpp:345-370,
``iss >> x``) reads back the identical doubles.

Random numbers come from a project PRNG (splitmix64 seeding a xoshiro256**
stream) so every workload is reproducible bit for bit on any host.
"""
from __future__ import annotations

import math
import os
from typing import Iterable

import numpy as np

_M64 = (1 << 64) - 1


class Rng:
    """xoshiro256** seeded by splitmix64 (Blackman & Vigna, public domain algorithm)."""

    def __init__(self, seed: int):
        x = seed & _M64
        s = []
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & _M64
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
            s.append(z ^ (z >> 31))
        self.s = s

    def next_u64(self) -> int:
        s = self.s
        r = (((s[1] * 5) & _M64) << 7 | ((s[1] * 5) & _M64) >> 57) & _M64
        r = (r * 9) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = ((s[3] << 45) | (s[3] >> 19)) & _M64
        return r

    def uniform(self, n: int = 1) -> np.ndarray:
        """Doubles in [0, 1) from the top 53 bits."""
        return np.array([(self.next_u64() >> 11) * (1.0 / (1 << 53)) for _ in range(n)], dtype=np.float64)

    def normal(self, n: int) -> np.ndarray:
        """Box-Muller normals (pairs)."""
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        u1 = np.where(u1 <= 0.0, 1e-300, u1)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])
        return z[:n]

    def unit_vectors(self, n: int) -> np.ndarray:
        v = self.normal(3 * n).reshape(n, 3)
        return v / np.linalg.norm(v, axis=1, keepdims=True)

    def quaternions(self, n: int) -> np.ndarray:
        """Uniform random unit quaternions (w, x, y, z)."""
        q = self.normal(4 * n).reshape(n, 4)
        return q / np.linalg.norm(q, axis=1, keepdims=True)


def _orient_outward(pts: np.ndarray, faces: np.ndarray) -> np.ndarray:
    c = pts.mean(axis=0)
    tris = pts[faces]
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    flip = np.einsum("ij,ij->i", n, tris.mean(axis=1) - c) < 0
    f = faces.copy()
    f[flip] = f[flip][:, [0, 2, 1]]
    return f


def convex_hull_mesh(pts: np.ndarray) -> np.ndarray:
    """Outward-oriented convex hull triangles of a point cloud (scipy qhull)."""
    from scipy.spatial import ConvexHull

    hull = ConvexHull(pts)
    faces = _orient_outward(pts, np.asarray(hull.simplices, dtype=np.int64))
    return np.ascontiguousarray(pts[faces], dtype=np.float64)


def ellipsoid_hull(n_points: int = 200, axes=(1.0, 0.6, 0.4), seed: int = 1) -> np.ndarray:
    """Hull of n seeded points uniform in direction on an ellipsoid surface."""
    d = Rng(seed).unit_vectors(n_points)
    return convex_hull_mesh(d * np.asarray(axes, dtype=np.float64))


def sphere_hull(n_points: int, seed: int) -> np.ndarray:
    return convex_hull_mesh(Rng(seed).unit_vectors(n_points))


def icosphere(level: int) -> np.ndarray:
    """Unit icosphere with 20 * 4**level triangles."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.asarray(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache: dict = {}
        nf = []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    P = np.asarray(verts)
    F = _orient_outward(P, np.asarray(faces, dtype=np.int64))
    return np.ascontiguousarray(P[F])


def _radial_noise(seed: int, u: np.ndarray) -> np.ndarray:
    rng = Rng(seed)
    d = rng.unit_vectors(6)
    ph = rng.uniform(6) * 2.0 * math.pi
    return np.mean(np.cos(4.0 * math.pi * (u @ d.T) + ph[None, :]), axis=1)


def star_hull_mesh(seed: int, n_points: int, amp: float = 0.3) -> np.ndarray:
    """Concave radial-noise star on the triangulation of the convex hull of
    n_points seeded unit vectors (2 n_points - 4 triangles): each vertex u is
    scaled by 1 + amp*noise(u), noise as in star_mesh (seeded by seed + 1)."""
    tris = convex_hull_mesh(Rng(seed).unit_vectors(n_points))
    flat = tris.reshape(-1, 3)
    u = flat / np.linalg.norm(flat, axis=1, keepdims=True)
    noise = _radial_noise(seed + 1, u)
    return np.ascontiguousarray((u * (1.0 + amp * noise)[:, None]).reshape(-1, 3, 3))


def star_mesh(seed: int, level: int = 3, amp: float = 0.3) -> np.ndarray:
    """Concave radial-noise star: icosphere vertices scaled by 1 + amp*noise(u).

    noise(u) = mean over 6 seeded terms of cos(4 * pi * (u . d_k) + phi_k), a
    smooth field of angular degree ~4 whose peaks make non-convex lobes that
    produce several contact loops per pair.
    """
    rng = Rng(seed)
    d = rng.unit_vectors(6)
    ph = rng.uniform(6) * 2.0 * math.pi
    tris = icosphere(level)
    flat = tris.reshape(-1, 3)
    u = flat / np.linalg.norm(flat, axis=1, keepdims=True)
    noise = np.mean(np.cos(4.0 * math.pi * (u @ d.T) + ph[None, :]), axis=1)
    return np.ascontiguousarray((u * (1.0 + amp * noise)[:, None]).reshape(-1, 3, 3))


def box_wall(sx: float, sy: float, sz: float) -> np.ndarray:
    """12-triangle closed box centred at the origin (the 12-tri wall mesh)."""
    hx, hy, hz = sx / 2, sy / 2, sz / 2
    V = np.array([(-hx, -hy, -hz), (hx, -hy, -hz), (hx, hy, -hz), (-hx, hy, -hz),
                  (-hx, -hy, hz), (hx, -hy, hz), (hx, hy, hz), (-hx, hy, hz)], dtype=np.float64)
    F = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4), (2, 3, 7), (2, 7, 6),
         (0, 4, 7), (0, 7, 3), (1, 2, 6), (1, 6, 5)]
    return np.ascontiguousarray(V[np.asarray(F)])


def plane_wall(sx: float, sy: float) -> np.ndarray:
    hx, hy = sx / 2, sy / 2
    V = np.array([(-hx, -hy, 0.0), (hx, -hy, 0.0), (hx, hy, 0.0), (-hx, hy, 0.0)], dtype=np.float64)
    return np.ascontiguousarray(V[np.asarray([(0, 1, 2), (0, 2, 3)])])


def write_stl_ascii(path: str, tris: np.ndarray, name: str = "zdx") -> None:
    tris = np.asarray(tris, dtype=np.float64)
    lines = [f"solid {name}"]
    for t in tris:
        n = np.cross(t[1] - t[0], t[2] - t[0])
        ln = np.linalg.norm(n)
        n = n / ln if ln > 0 else n
        lines.append(f"  facet normal {n[0]:.9g} {n[1]:.9g} {n[2]:.9g}")
        lines.append("    outer loop")
        for v in t:
            lines.append(f"      vertex {v[0]:.17g} {v[1]:.17g} {v[2]:.17g}")
        lines.append("    endloop")
        lines.append("  endfacet")
    lines.append(f"endsolid {name}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_stl_binary(path: str, tris: np.ndarray) -> None:
    import struct

    tris = np.asarray(tris, dtype=np.float64)
    with open(path, "wb") as f:
        f.write(b"zdx binary stl".ljust(80, b"\0"))
        f.write(struct.pack("<I", len(tris)))
        for t in tris:
            n = np.cross(t[1] - t[0], t[2] - t[0])
            ln = np.linalg.norm(n)
            n = n / ln if ln > 0 else n
            f.write(struct.pack("<3f", *n))
            for v in t:
                f.write(struct.pack("<3f", *v))
            f.write(struct.pack("<H", 0))


def mesh_radius(tris: np.ndarray) -> float:
    return float(np.max(np.linalg.norm(np.asarray(tris).reshape(-1, 3), axis=1)))


def format_config(globals_: dict, particles: Iterable[dict], walls: Iterable[dict] = ()) -> str:
    """zdem_cpu config text (src/main.cpp:1531-1689) with full-precision numbers."""

    def fmt(v):
        if isinstance(v, (list, tuple, np.ndarray)):
            return " ".join(fmt(x) for x in v)
        if isinstance(v, float) or isinstance(v, np.floating):
            return f"{float(v):.17g}"
        return str(v)

    out = [f"{k} = {fmt(v)}" for k, v in globals_.items()]
    for w in walls:
        out.append("wall")
        out += [f"{k} = {fmt(v)}" for k, v in w.items()]
        out.append("end_wall")
    for p in particles:
        out.append("particle")
        out += [f"{k} = {fmt(v)}" for k, v in p.items()]
        out.append("end_particle")
    return "\n".join(out) + "\n"


def ensure_dir(p: str) -> str:
    os.makedirs(p, exist_ok=True)
    return p
