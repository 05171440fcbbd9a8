"""The per-pair tool of polyline.py (polyline.py:707-834) as zdx-pair.

CPU: the host pieces it is made of (zedm-x_amd/csrc/zdx_output.cpp, host
build) against vectors and files polyline.py's own functions wrote
(tests/golden/polyline/, made by tests/golden/make_polyline_golden.py):
Python's float text, the transform, the per-loop contact numbers and the two
VTK writers, byte for byte; plus closed forms that catch a dropped term.
GPU: zdx-pair end to end on box-on-box cases with closed-form loops."""
import ctypes as C
import json
import math
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from zdx import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "zedm-x_amd", "csrc")
GOLD = os.path.join(ROOT, "tests", "golden", "polyline")
PAIR = os.path.join(ROOT, "zedm-x_amd", "bin", "zdx-pair")
dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


@pytest.fixture(scope="module")
def W(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("pltest") / "libpolylinetest.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", out,
                    os.path.join(ROOT, "tests", "csrc", "outtest.cpp"), os.path.join(CSRC, "zdx_output.cpp"),
                    os.path.join(CSRC, "zdx_host.cpp")], check=True)
    L = C.CDLL(out)
    L.ot_py_repr.argtypes = [C.c_double, C.c_char_p, C.c_int]
    L.ot_transform.argtypes = [dp, C.c_longlong, dp, dp, dp]
    L.ot_contact.argtypes = [dp, C.c_longlong, C.c_double, dp]
    L.ot_write_tris.argtypes = [C.c_char_p, dp, C.c_longlong, C.c_char_p]
    L.ot_write_loops.argtypes = [C.c_char_p, C.c_int, ip, dp, C.c_int, dp, ip, ip, C.c_int]
    return L


@pytest.fixture(scope="module")
def FX():
    with open(os.path.join(GOLD, "polyline_vectors.json")) as f:
        return json.load(f)


def unhex(v):
    return np.array([float.fromhex(x) for x in v], dtype=np.float64)


def repr_of(W, x):
    buf = C.create_string_buffer(64)
    n = W.ot_py_repr(x, buf, 64)
    assert n > 0
    return buf.value.decode()


def test_float_text_is_python_repr(W):
    """py_float_repr against Python's own repr (the text of f"{x}")."""
    rng = np.random.default_rng(3)
    vals = [0.0, -0.0, 1.0, -1.0, 0.1, 0.2 + 0.1, 1 / 3, 2 / 3, 100.0, 1e15, 1e16, 1e17, 9999999999999998.0,
            1.2345678901234567e16, 0.0001, 0.00012, 1e-5, 9.99e-5, 5e-324, 2.2250738585072014e-308,
            1.7976931348623157e308, 123456789.0, 2.0 ** 53, 2.0 ** 60, -3.5e-7, 1e22, 1e23, math.pi, math.e,
            float("inf"), float("-inf"), float("nan")]
    vals += list(rng.normal(size=500) * 10.0 ** rng.integers(-30, 30, size=500))
    vals += list(rng.integers(-10 ** 12, 10 ** 12, size=100).astype(float))
    vals += [struct.unpack("<d", rng.bytes(8))[0] for _ in range(500)]
    for x in vals:
        assert repr_of(W, float(x)) == repr(float(x)), x


def test_transform_matches_polyline(W, FX):
    for c in FX["transform"]:
        pts = unhex(c["in"])
        out = np.zeros_like(pts)
        W.ot_transform(pts, len(pts) // 3, unhex(c["t"]), unhex(c["q"]), out)
        assert np.array_equal(out, unhex(c["out"])), (c["mesh"], c["q"])


def test_transform_is_a_rotation(W):
    """Closed form: a unit quaternion rotates (orthonormal map, det +1), the
    q = (cos a/2, sin a/2 z) case turns x into (cos a, sin a, 0)."""
    a = 0.7
    q = np.array([math.cos(a / 2), 0.0, 0.0, math.sin(a / 2)]) * 3.0    # not normalised
    pts = np.eye(3).ravel().copy()
    out = np.zeros(9)
    W.ot_transform(pts, 3, np.zeros(3), q, out)
    R = out.reshape(3, 3).T
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-15) and abs(np.linalg.det(R) - 1.0) < 1e-15
    assert np.allclose(out[0:3], [math.cos(a), math.sin(a), 0.0], atol=1e-15)


def _contact(W, pts, k):
    o = np.zeros(18)
    W.ot_contact(np.ascontiguousarray(pts, dtype=np.float64).ravel(), len(pts), k, o)
    return o


def test_contact_matches_polyline(W, FX):
    for c in FX["contact"]:
        pts = unhex(c["pts"]).reshape(-1, 3)
        o = _contact(W, pts, float.fromhex(c["k"]))
        assert np.array_equal(o[0:3], unhex(c["Sn"])) and np.array_equal(o[3:6], unhex(c["Gn"])), c["case"]
        assert o[6] == float.fromhex(c["area"]) and np.array_equal(o[7:10], unhex(c["nA"])), c["case"]
        if c["xc0"] is None:
            assert o[17] == 0.0
        else:
            assert o[17] == 1.0 and np.array_equal(o[10:13], unhex(c["xc0"])), c["case"]
        assert o[13] == float.fromhex(c["Fn"]) and np.array_equal(o[14:17], unhex(c["F_A"])), c["case"]


def test_contact_closed_forms(W):
    """Planar loop with area A, centroid c and unit normal m: Sn = A m,
    xc0 = -2/3 (c - m (m.c)), nA = -m, Fn = k A / 2, F_A = Fn nA."""
    rng = np.random.default_rng(4)
    for _ in range(20):
        m = int(rng.integers(3, 9))
        ang = np.sort(rng.uniform(0, 2 * np.pi, size=m))
        r = rng.uniform(0.5, 2.0)
        loc = np.stack([r * np.cos(ang), r * np.sin(ang), np.zeros(m)], axis=1)
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        if np.linalg.det(Q) < 0:
            Q[:, 0] *= -1
        off = rng.normal(size=3)
        pts = loc @ Q.T + off
        x, y = loc[:, 0], loc[:, 1]
        x1, y1 = np.roll(x, -1), np.roll(y, -1)
        cr = x * y1 - x1 * y
        A = 0.5 * cr.sum()
        cl = np.array([((x + x1) * cr).sum(), ((y + y1) * cr).sum(), 0.0]) / (6 * A)
        c = Q @ cl + off
        mvec = Q[:, 2]
        k = 3.0
        o = _contact(W, pts, k)
        assert np.allclose(o[0:3], A * mvec, atol=1e-12) and abs(o[6] - A) < 1e-12
        assert np.allclose(o[7:10], -mvec, atol=1e-12)
        assert np.allclose(o[10:13], -2.0 / 3.0 * (c - mvec * (mvec @ c)), atol=1e-12)
        assert abs(o[13] - 0.5 * k * A) < 1e-12 and np.allclose(o[14:17], -0.5 * k * A * mvec, atol=1e-12)
    o = _contact(W, [[0, 0, 0], [1e-8, 0, 0], [0, 1e-8, 0]], 1.0)
    assert o[17] == 0.0 and o[6] == 0.0 and o[13] == 0.0


def test_loop_filter(W):
    """polyline.py:798-806: fewer than 3 points, or a bounding-box diagonal
    below 2 tol, drops the loop; a diagonal of exactly 2 tol keeps it."""
    W.ot_keeps_loop.argtypes = [dp, C.c_longlong, C.c_double]
    f = lambda pts, tol: W.ot_keeps_loop(np.ascontiguousarray(pts, dtype=np.float64).ravel(), len(pts), tol)
    assert not f([[0, 0, 0], [1, 0, 0]], 1e-6)
    assert f([[0, 0, 0], [1, 0, 0], [0, 1, 0]], 1e-6)
    assert f([[0, 0, 0], [1, 0, 0], [0, 0, 0]], 0.5)                 # diagonal 1.0 == 2 * 0.5
    assert not f([[0, 0, 0], [1, 0, 0], [0, 0, 0]], 0.5000000000000001)
    assert not f([[0, 0, 0], [1e-6, 0, 0], [0, 1e-6, 0]], 1e-6)       # sqrt(2) e-6 < 2e-6
    assert f([[5, 5, 5], [5 + 3e-6, 5, 5], [5, 5 + 3e-6, 5]], 1e-6)


def test_tris_writer_bytes(W, FX, tmp_path):
    for c in FX["tris_files"]:
        tris = unhex(c["tris"])
        p = tmp_path / c["file"]
        assert W.ot_write_tris(str(p).encode(), tris, len(tris) // 9, c["comment"].encode()) == 0
        assert p.read_bytes() == open(os.path.join(GOLD, c["file"]), "rb").read(), c["file"]


def _write_loops(W, path, loops, points, ids):
    off = np.cumsum([0] + [len(lp) // 3 for lp in loops]).astype(np.int32)
    pts = np.concatenate(loops) if loops else np.zeros(0)
    cp = np.array([p if p is not None else np.zeros(3) for p in points], dtype=np.float64).reshape(-1)
    has = np.array([p is not None for p in points], dtype=np.int32)
    ida = np.array(ids, dtype=np.int32)
    return W.ot_write_loops(str(path).encode(), len(loops), off, np.ascontiguousarray(pts), len(points),
                            np.ascontiguousarray(cp) if len(cp) else np.zeros(1), has if len(has) else
                            np.zeros(1, np.int32), ida, len(ida))


def test_loops_writer_bytes(W, FX, tmp_path):
    for c in FX["loop_files"]:
        p = tmp_path / c["file"]
        rc = _write_loops(W, p, [unhex(lp) for lp in c["loops"]],
                          [None if q is None else unhex(q) for q in c["points"]], c["ids"])
        assert rc == 0
        assert p.read_bytes() == open(os.path.join(GOLD, c["file"]), "rb").read(), c["file"]


def test_zdx_pair_arguments(tmp_path):
    """argparse's behaviour for the command line: usage and exit status 2
    on missing or unknown arguments, exit status 1 on unreadable STL files
    (no GPU is touched before the meshes are read)."""
    if not os.path.exists(PAIR):
        pytest.skip("zdx-pair is not built (make -C zedm-x_amd)")
    for args in ([], ["a.stl"], ["a.stl", "b.stl", "--bogus"], ["a.stl", "b.stl", "--tA", "1", "2"]):
        r = subprocess.run([PAIR] + args, cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and r.stderr.startswith("usage: zdx-pair stlA stlB"), args
    r = subprocess.run([PAIR, "missing_a.stl", "missing_b.stl"], cwd=str(tmp_path), capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "failed to read" in r.stderr


# ------------------------------------------------------------------ GPU
def _run_pair(cwd, args):
    r = subprocess.run([PAIR] + args, cwd=str(cwd), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def _vec(s):
    return np.array([float(x) for x in s.strip("[]").split(",")])


@pytest.mark.gpu
@pytest.mark.parametrize("depth,off,pa,qa,split", [(0.2, (0.0, 0.0), (0.0, 0.0), None, False),
                                                    (0.3, (0.2, 0.2), (0.3, -0.7), None, True),
                                                    (0.25, (0.1, -0.3), (0.0, 0.4), (0.0, 1.0, 0.0, 0.0), True)])
def test_zdx_pair_box_on_box(W, tmp_path, depth, off, pa, qa, split):
    """Box A (1 x 1.2 x 0.8) with its +x end inside box B (2 x 2 x 2): one
    contact, the outline of A's cross-section at B's -x face, so |Sn| =
    0.96, nA = +-x, xc0 = -2/3 of the rectangle centre's in-plane part,
    Fn = k/2 |Sn|; the loop VTK holds one LINES and one VERTICES cell, both
    contact 0, the four rectangle corners among its points; the transformed
    meshes are the host writer's bytes.  qa = (0, 1, 0, 0) turns A by pi
    about x (the box maps onto itself, the triangles do not)."""
    import zdx
    if not os.path.exists(PAIR):
        pytest.fail("zdx-pair is not built (make -C zedm-x_amd)")
    A, B = synth.box_wall(1.0, 1.2, 0.8), synth.box_wall(2.0, 2.0, 2.0)
    synth.write_stl_ascii(str(tmp_path / "A.stl"), A)
    synth.write_stl_ascii(str(tmp_path / "B.stl"), B)
    tA = np.array([0.0, pa[0], pa[1]])
    tB = tA + np.array([1.5 - depth, off[0], off[1]])
    k = 5.0
    args = ["A.stl", "B.stl", "--tA"] + [repr(float(v)) for v in tA] + ["--tB"] + [repr(float(v)) for v in tB]
    if qa is not None:
        args += ["--qA"] + [repr(float(v)) for v in qa]
    args += ["--k", repr(k)] + (["--split-contacts"] if split else [])
    out = _run_pair(tmp_path, args)
    print(out)
    assert "Transformed meshes written: A_transformed.vtk, B_transformed.vtk" in out
    assert ("Coarse components: 1 (no split)" in out) != split
    if split:
        assert "Coarse components: 1 (tol=1e-06)" in out
    assert "idA=0 idB=1 nContact=1" in out
    m = re.search(r"\[cid=0\] nVert=(\d+) area=(\S+) Fn=(\S+)\n    nA=(\[.*?\]) xc0=(\[.*?\]) F_A=(\[.*?\])", out)
    assert m, out
    nv, area, Fn = int(m.group(1)), float(m.group(2)), float(m.group(3))
    nA, xc0, FA = _vec(m.group(4)), _vec(m.group(5)), _vec(m.group(6))
    c = np.array([0.5 - depth, pa[0], pa[1]])
    assert nv >= 4 and abs(area - 0.96) < 1e-12 and abs(Fn - 0.5 * k * 0.96) < 1e-12
    assert np.allclose(np.abs(nA), [1.0, 0.0, 0.0], atol=1e-12)
    assert np.allclose(xc0, -2.0 / 3.0 * np.array([0.0, c[1], c[2]]), atol=1e-12)
    assert np.allclose(FA, Fn * nA, atol=1e-12)
    # the per-contact lines print the same numbers
    m2 = re.search(r"\[Contact 0\] nVert=(\d+) \|Sn\|=(\S+)\n    nA=(\[.*?\])\n    xc\(0\)=(\[.*?\])", out)
    assert m2 and int(m2.group(1)) == nv and float(m2.group(2)) == area
    assert np.array_equal(_vec(m2.group(3)), nA) and np.array_equal(_vec(m2.group(4)), xc0)
    # the loops file
    txt = (tmp_path / "routeB_intersection.vtk").read_text()
    lines = txt.split("\n")
    assert lines[:4] == ["# vtk DataFile Version 3.0", "reconstructed_contact_loops", "ASCII", "DATASET POLYDATA"]
    npts = int(lines[4].split()[1])
    P = np.array([[float(x) for x in ln.split()] for ln in lines[5:5 + npts]])
    assert lines[5 + npts] == f"LINES 1 {nv + 2}"
    ids = [int(x) for x in lines[6 + npts].split()]
    assert ids[0] == nv + 1 and ids[1] == ids[-1] and len(ids) == nv + 2
    assert lines[7 + npts] == "VERTICES 1 2"
    vid = int(lines[8 + npts].split()[1])
    assert np.array_equal(P[vid], xc0)
    assert lines[9 + npts:15 + npts] == ["", "CELL_DATA 2", "SCALARS contact_id int 1", "LOOKUP_TABLE default", "0", "0"]
    for sy in (-0.6, 0.6):
        for sz in (-0.4, 0.4):
            corner = c + np.array([0.0, sy, sz])
            assert np.min(np.linalg.norm(P[:npts] - corner, axis=1)) < 1e-12, corner
    assert np.allclose(P[np.arange(npts) != vid, 0], c[0], atol=1e-12)
    # the transformed meshes: the host writer's bytes on polyline's transform
    for name, tris, t, q, label in (("A", A, tA, qa, "Mesh A (transformed)"), ("B", B, tB, None, "Mesh B (transformed)")):
        src = zdx.load_stl(str(tmp_path / f"{name}.stl")).reshape(-1)
        res = np.zeros_like(src)
        W.ot_transform(np.ascontiguousarray(src), len(src) // 3, np.ascontiguousarray(t, dtype=np.float64),
                       np.array(q if q is not None else (1.0, 0.0, 0.0, 0.0), dtype=np.float64), res)
        ref = tmp_path / f"{name}_host.vtk"
        assert W.ot_write_tris(str(ref).encode(), res, len(res) // 9, label.encode()) == 0
        assert (tmp_path / f"{name}_transformed.vtk").read_bytes() == ref.read_bytes()


@pytest.mark.gpu
def test_zdx_pair_no_contact(tmp_path):
    if not os.path.exists(PAIR):
        pytest.fail("zdx-pair is not built (make -C zedm-x_amd)")
    synth.write_stl_ascii(str(tmp_path / "A.stl"), synth.box_wall(1.0, 1.0, 1.0))
    out = _run_pair(tmp_path, ["A.stl", "A.stl", "--tB", "3.0", "0.0", "0.0"])
    assert out.strip().split("\n")[-1] == "No intersection segments found."
    assert not (tmp_path / "routeB_intersection.vtk").exists()
