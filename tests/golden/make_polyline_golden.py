"""Writes tests/golden/polyline/: input/output vectors of the polyline.py
functions that only polyline.py itself pins -- its transform (numpy's
norm of q, polyline.py:36-52), its per-loop contact numbers (numpy dots,
polyline.py:193-218, 655-701) and the bytes of its two VTK writers
(polyline.py:544-649).  The geometry pipeline is pinned by the C++ reference
instead.  Run in a checkout that has the reference next to it:

    python tests/golden/make_polyline_golden.py /root/reference

The fixtures are data (hex floats and the files the writers wrote); the
reference itself is not kept."""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "polyline")


def hexs(a):
    return [float(x).hex() for x in np.asarray(a, dtype=float).ravel()]


def tri_soups():
    rng = np.random.default_rng(5)
    tet = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 0, 0], [0, 0, 1], [1, 0, 0]],
                    [[0, 0, 0], [0, 1, 0], [0, 0, 1]], [[1, 0, 0], [0, 0, 1], [0, 1, 0]]], dtype=float) * 0.37
    soup = rng.normal(size=(12, 3, 3)) * 10.0 ** rng.uniform(-3, 2, size=(12, 1, 1))
    # near-duplicates: within and just beyond the writer's 1e-9 merge radius
    near = np.array([[[0.5, 0.5, 0.5], [0.5 + 4e-10, 0.5, 0.5], [0.5, 0.5 + 2e-9, 0.5]],
                     [[0.5, 0.5, 0.5 + 9.9e-10], [1e-17, 1e16, -3e-5], [123456.789, -0.0, 2.5e-4]]])
    return {"tet": tet, "soup": soup, "near": near}


POSES = [([0.0, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0]),
         ([0.1, -0.2, 0.3], [0.9238795325112867, 0.0, 0.3826834323650898, 0.0]),
         ([1.5, 2.25, -0.75], [2.1, -0.4, 1.3, 0.77]),
         ([1e-3, 0.0, 7.0], [1e-31, 0.0, 0.0, 0.0])]


def loops_cases():
    rng = np.random.default_rng(9)
    sq = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 1.0, 0.0]]
    out = {"unit_square": [sq]}
    ring = []
    for s in range(3):
        m = int(rng.integers(3, 12))
        ang = np.sort(rng.uniform(0, 2 * np.pi, size=m))
        r = rng.uniform(0.5, 2.0, size=m)
        pts = np.stack([r * np.cos(ang), r * np.sin(ang), rng.normal(scale=0.05, size=m)], axis=1)
        ring.append((pts * 10.0 ** (-3 + s) + rng.normal(size=3)).tolist())
    out["rings"] = ring
    out["degenerate"] = [[[0.0, 0.0, 0.0], [1e-8, 0.0, 0.0], [0.0, 1e-8, 0.0]]]
    return out


def writer_cases():
    sq = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 1.0, 0.0]]
    sq2 = [[1.0, 0.0, 0.0], [2.0, 0.0, 0.0], [2.0, 1.0, 0.0], [1.0 + 5e-10, 1.0, 0.0]]   # shares points with sq
    tiny = [[3.0, 3.0, 3.0]]                                                              # < 2 ids: no cell
    c1, c2 = [-1 / 3, -1 / 3, 0.0], [0.1, 1e-20, 5e16]
    return {
        "two_loops": ([sq, sq2], [c1, c2], [0, 1]),
        "skip_short": ([sq, tiny, sq2], [c1, c2, None], [4, 5, 6]),
        "fewer_points": ([sq, sq2], [c1], [7, 8]),
        "none_point": ([sq, sq2], [None, c2], [0, 1]),
    }


def main(ref):
    sys.path.insert(0, ref)
    import polyline as PL

    os.makedirs(OUT, exist_ok=True)
    fx = {"transform": [], "contact": [], "tris_files": [], "loop_files": []}
    for name, tris in tri_soups().items():
        for pi, (t, q) in enumerate(POSES):
            res = PL.transform_points(tris.reshape(-1, 3), t, q)
            fx["transform"].append({"mesh": name, "t": hexs(t), "q": hexs(q), "in": hexs(tris), "out": hexs(res)})
            fn = f"tris_{name}_{pi}.vtk"
            PL.write_vtk_polydata_tris(os.path.join(OUT, fn), res.reshape(-1, 3, 3), comment=f"Mesh {name} {pi}")
            fx["tris_files"].append({"file": fn, "comment": f"Mesh {name} {pi}", "tris": hexs(res)})
    for name, loops in loops_cases().items():
        for k in (1.0, 2.5e6):
            man = PL.build_contacts_from_loops(0, 1, loops, list(range(len(loops))), k_stiff=k)
            for lp, c in zip(loops, man["contacts"]):
                fx["contact"].append({"case": name, "k": float(k).hex(), "pts": hexs(lp), "Sn": hexs(c["Sn"]),
                                      "Gn": hexs(c["Gn"]), "area": float(c["area"]).hex(), "nA": hexs(c["nA"]),
                                      "xc0": None if c["xc0"] is None else hexs(c["xc0"]),
                                      "Fn": float(c["Fn"]).hex(), "F_A": hexs(c["F_A"])})
    for name, (loops, cps, ids) in writer_cases().items():
        fn = f"loops_{name}.vtk"
        PL.write_vtk_polydata_polylines(os.path.join(OUT, fn), loops, cps, ids)
        fx["loop_files"].append({"file": fn, "loops": [hexs(lp) for lp in loops],
                                 "points": [None if c is None else hexs(c) for c in cps], "ids": ids})
    with open(os.path.join(OUT, "polyline_vectors.json"), "w") as f:
        json.dump(fx, f, indent=0)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
