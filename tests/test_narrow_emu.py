"""CPU: the narrow kernels (zedm-x_amd/csrc/zdx_narrow.h, the cull and sweep
of src/main.cpp:2123-2145 with float boxes) compiled for the host and run
under the SIMT emulation of tests/csrc/simt_emu.h -- each workgroup's threads
as fibers with real barriers, wave operations that need all 64 lanes at the
same call -- with AddressSanitizer and UndefinedBehaviorSanitizer: the
one-wave-per-pair kernel with the workgroup kernel over the pairs it leaves,
as the library runs them, and the workgroup kernel alone over every pair.

The hit lists of every compiled workgroup shape, in the time step's form
(shuffled work items from the queue) and the per-pair call's (strided pairs,
with and without provenance), must equal a plain restatement: the triangles
of each body whose outward-rounded float box meets the other body's float
AABB, and the pairs of those, in (ta, tb) order, whose float boxes overlap.
The pairs are chosen so that every path of the kernel runs: bodies above
2048 vertices (the general cull), meshes that take several cull rounds at
small workgroups, kept triangles of both sides in LDS, of B only, of
neither, vertex boxes in LDS or not, kept boxes made by the thread that
tested the triangle or in a parallel pass, pairs with no kept triangle."""
import os
import re
import shutil
import struct
import subprocess

import numpy as np
import pytest

from zdx import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def indexed(tris):
    """Unique corners (bitwise) and the triangles' indices into them."""
    v, inv = np.unique(np.asarray(tris, float).reshape(-1, 3), axis=0, return_inverse=True)
    return v, inv.reshape(-1, 3).astype(np.int32)


def rot(rng):
    q = rng.normal(size=4)
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def write_input(path):
    meshes = {"box": synth.box_wall(1.6, 1.6, 1.6), "ico1": synth.icosphere(1), "ico2": synth.icosphere(2),
              "ico3": synth.icosphere(3), "ico4": synth.icosphere(4),
              "star": synth.star_hull_mesh(3, 642)}
    names = list(meshes)
    idx = {k: indexed(meshes[k]) for k in names}
    # (mesh A, mesh B, centre distance as a fraction of the radii's sum)
    cases = [("box", "box", 0.8), ("ico1", "ico2", 0.9), ("ico1", "ico2", 0.5), ("ico2", "ico2", 0.95),
             ("ico2", "ico2", 0.6), ("ico2", "ico2", 0.2), ("ico3", "ico3", 0.97), ("ico3", "ico3", 0.7),
             ("ico3", "ico3", 0.0), ("ico3", "ico1", 0.5), ("star", "ico3", 0.8), ("ico2", "star", 0.6),
             ("ico4", "ico2", 0.9), ("ico2", "ico4", 0.9), ("ico4", "ico4", 0.95), ("ico4", "ico3", 0.4),
             ("ico2", "ico1", 1.6), ("box", "ico3", 0.3),
             # small pairs, which one wave takes
             ("ico1", "ico1", 0.9), ("box", "ico1", 0.7), ("ico2", "ico1", 0.8), ("ico2", "ico2", 0.99),
             ("box", "box", 1.3), ("ico1", "box", 0.4), ("ico2", "ico2", 0.85), ("star", "ico1", 0.9),
             ("ico1", "star", 0.95), ("ico2", "box", 0.75)]
    rng = np.random.default_rng(7)
    bodies, pairs = [], []
    for a, b, f in cases:
        ra = np.max(np.linalg.norm(idx[a][0], axis=1))
        rb = np.max(np.linalg.norm(idx[b][0], axis=1))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        base = rng.uniform(-5, 5, 3)
        for m, c in ((a, base), (b, base + d * (ra + rb) * f)):
            bodies.append((names.index(m), len(idx[m][1]) - 1, idx[m][0] @ rot(rng).T + c))
        pairs.append((len(bodies) - 2, len(bodies) - 1))
    with open(path, "wb") as fh:
        fh.write(struct.pack("<i", len(names)))
        for k in names:
            v, t = idx[k]
            fh.write(struct.pack("<ii", len(v), len(t)))
            fh.write(np.ascontiguousarray(t, "<i4").tobytes())
        fh.write(struct.pack("<i", len(bodies)))
        for m, nd, wv in bodies:
            fh.write(struct.pack("<ii", m, nd))
            fh.write(np.ascontiguousarray(wv, "<f8").tobytes())
        fh.write(struct.pack("<i", len(pairs)))
        for i, j in pairs:
            fh.write(struct.pack("<ii", i, j))
    return len(pairs)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_narrow_kernel_under_simt_emulation(tmp_path):
    exe = str(tmp_path / "narrow_emu")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-w", "-o", exe,
           os.path.join(ROOT, "tests", "csrc", "narrow_emu.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    inp = str(tmp_path / "pairs.bin")
    n = write_input(inp)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, inp], capture_output=True, text=True, env=env, timeout=900)
    print(r.stdout)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-4000:])
    assert "narrow_emu ok" in r.stdout
    assert len(re.findall(r"^ok workgroup shape", r.stdout, re.M)) == 11
    assert len(re.findall(r"^ok wave\+heavy shape", r.stdout, re.M)) == 4
    assert all(f"{n} pairs" in l for l in r.stdout.splitlines() if l.startswith("ok "))
    # the wave kernel takes most pairs and leaves some to the workgroup kernel
    nwave = int(re.search(r"wave kernel takes (\d+) of", r.stdout).group(1))
    assert 0 < nwave < n
    assert all(f"({n - nwave} to the workgroup kernel)" in l for l in r.stdout.splitlines()
               if l.startswith("ok wave+heavy"))
    # every path of the kernel ran at some workgroup size
    paths = {int(m.group(1)): list(map(int, m.groups()[1:]))
             for m in re.finditer(r"paths at (\d+) threads: general (\d+), several cull rounds (\d+), kept in LDS: "
                                  r"both (\d+), B only (\d+), neither (\d+); vertex boxes in LDS (\d+), "
                                  r"boxes made where tested (\d+)", r.stdout)}
    assert set(paths) == {64, 128, 256, 512}
    tot = np.array(list(paths.values())).sum(0)
    assert all(tot[:5] > 0), paths
    assert paths[512][5] < n - paths[512][0], paths          # some pairs' vertex boxes do not fit the LDS
    assert paths[64][1] > 0 and paths[512][2] > 0, paths
    # kept boxes made by the testing thread, and in the parallel pass
    fast = {k: n - v[0] for k, v in paths.items()}
    assert all(v[6] > 0 for v in paths.values()) and all(paths[k][6] < fast[k] for k in paths), paths
