"""The product's per-pair geometry (zedm-x_amd/csrc/zdx_geom.h — the code the
GPU kernels run) compiled for the host and compared bit-for-bit with the
oracle: tri-tri, the libstdc++-equivalent introsort, and the whole per-pair
topology pipeline."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import golden_inputs as GI
import oracle as O
from zdx import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "csrc", "hosttest.cpp")
HDRS = [os.path.join(ROOT, "zedm-x_amd", "csrc", f) for f in ("zdx_geom.h", "zdx_sort.h")]


@pytest.fixture(scope="module")
def H(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hosttest") / "libhosttest.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", out, SRC],
                   check=True)
    L = C.CDLL(out)
    ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    L.ht_start_order.argtypes = [ip, C.c_int, ip]
    L.ht_start_order_keyed.argtypes = [ip, C.c_int, ip]
    L.ht_sort_keys.argtypes = [ip, C.c_int, ip]
    L.ht_tri_tri.argtypes = [dp, dp, dp]
    L.ht_pair.argtypes = [dp, C.c_int, dp, C.c_int, dp, dp, dp, dp, C.c_double, C.c_double, C.c_int,
                          dp, C.c_int, dp, C.c_int, dp, C.c_int, np.ctypeslib.ndpointer(dtype=np.int64)]
    L.ht_snap_clusters.argtypes = [dp, C.c_int, C.c_double, C.c_int, ip]
    return L


def _degree_patterns(rng):
    for n in list(range(0, 40)) + [63, 64, 65, 100, 257, 1000, 4096]:
        yield rng.integers(0, 4, n)                         # few distinct degrees
        yield rng.integers(0, 9, n)
        yield np.sort(rng.integers(0, 9, n))
        yield np.sort(rng.integers(0, 9, n))[::-1]
        yield np.full(n, 2)
        yield np.where(rng.random(n) < 0.9, 2, 3)
        yield 1000 - np.arange(n) % 97                     # many distinct keys, periodic


def test_introsort_matches_std_sort(H):
    """Loop start order (src/main.cpp:977-982) must be std::sort's permutation,
    including its instability for more than 16 vertices."""
    rng = np.random.default_rng(5)
    L = O.lib()
    n_unstable = 0
    for deg in _degree_patterns(rng):
        deg = np.ascontiguousarray(deg, dtype=np.int32)
        n = len(deg)
        a = np.zeros(n, dtype=np.int32)
        b = np.zeros(n, dtype=np.int32)
        L.orc_loop_start_order(deg, n, a)
        H.ht_start_order(deg, n, b)
        assert np.array_equal(a, b), (n, deg[:20])
        c = np.zeros(n, dtype=np.int32)
        H.ht_start_order_keyed(deg, n, c)
        assert np.array_equal(a, c), (n, deg[:20])
        stable = np.array(sorted(range(n), key=lambda i: (0 if deg[i] == 2 else 1, -deg[i])), dtype=np.int32)
        n_unstable += int(not np.array_equal(a, stable))
    assert n_unstable > 5     # the cases do exercise the unstable permutation


def test_tri_tri_host_matches_oracle(H, golden_geometry):
    g = golden_geometry["tri_tri"]
    A, B = GI.tri_pairs(g["seed"], g["n"])
    L = O.lib()
    for i in range(len(A)):
        a = np.ascontiguousarray(A[i].reshape(-1))
        b = np.ascontiguousarray(B[i].reshape(-1))
        o1, o2 = np.zeros(6), np.zeros(6)
        f1 = L.orc_tri_tri(a, b, o1)
        f2 = H.ht_tri_tri(a, b, o2)
        assert f1 == f2, i
        assert np.array_equal(o1, o2), i


def host_pair(H, ta, tb, pose, me, bd, split):
    seg = np.zeros((20000, 6))
    rec = np.zeros((512, 28))
    vtx = np.zeros((20000, 3))
    cnt = np.zeros(7, dtype=np.int64)
    f = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    H.ht_pair(f(ta).reshape(-1), len(ta), f(tb).reshape(-1), len(tb), f(pose[0]), f(pose[1]), f(pose[2]),
              f(pose[3]), me, bd, int(split), seg.reshape(-1), 20000, rec.reshape(-1), 512, vtx.reshape(-1),
              20000, cnt)
    return cnt, seg[:cnt[1]], rec[:cnt[4]], vtx[:cnt[5]]


def _check_pair(H, ta, tb, pose, pa, split, name):
    r = O.pair_contact(ta, tb, *pose, pa[5], pa[6], split=split)
    cnt, seg, rec, vtx = host_pair(H, ta, tb, pose, pa[5], pa[6], split)
    assert cnt[6] == 0, name
    assert cnt[0] == r["tri_checks"], name
    assert np.array_equal(seg, r["segments"]), name
    assert cnt[2] == r["comps"] and cnt[3] == r["loops_raw"], name
    assert np.array_equal(rec[:, :13], r["loops"][:, :13]), name
    assert np.array_equal(rec[:, 23:26], r["loops"][:, 23:26]), name
    assert np.array_equal(vtx, r["loop_vtx"]), name
    return len(seg), len(rec)


def test_pair_pipeline_host_matches_oracle(H, ref_inputs):
    tot_loops = 0
    for name, mA, mB, pose, split in GI.pair_cases(ref_inputs):
        ta, pa, _ = O.build_mesh(mA)
        tb, _, _ = O.build_mesh(mB)
        _, nl = _check_pair(H, ta, tb, pose, pa, split, name)
        tot_loops += nl
    assert tot_loops >= 15


def test_pair_pipeline_host_random(H):
    """Seeded random poses over concave stars and hulls (many multi-loop pairs)."""
    m = GI.synth_meshes()
    built = {k: O.build_mesh(v) for k, v in m.items()}
    rng = synth.Rng(77)
    names = list(m)
    n_multi = 0
    for it in range(60):
        a = names[int(rng.uniform(1)[0] * len(names))]
        b = names[int(rng.uniform(1)[0] * len(names))]
        if it % 2 == 0:
            a, b = "star2a", "star2b"      # concave pairs: several loops per pair
        ta, pa, _ = built[a]
        tb, pb, _ = built[b]
        dep = 0.05 + 0.35 * float(rng.uniform(1)[0])
        pose = GI._pose(rng, pa[4], pb[4], dep, True)
        split = bool(it % 5)
        ns, nl = _check_pair(H, ta, tb, pose, pa, split, f"{a}-{b}-{it}")
        n_multi += nl > 1
    assert n_multi >= 2


def test_endpoint_jacobian_matches_finite_differences(H):
    """d(endpoint)/d(rigid motions of the two bodies) from the edge-plane
    provenance (zdx_geom.h endpoint_jacobian) against central differences of
    the product's own tri-tri on rotated/translated triangles."""
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    H.ht_endpoint.argtypes = [dp, dp, dp, dp, dp, dp, np.ctypeslib.ndpointer(dtype=np.int32), dp, dp]
    rng = np.random.default_rng(17)
    checked = 0
    for _ in range(400):
        a9 = rng.normal(size=9)
        b9 = rng.normal(size=9) * 0.8 + 0.1
        cA = a9.reshape(3, 3).mean(0) + rng.normal(size=3) * 0.3
        cB = b9.reshape(3, 3).mean(0) + rng.normal(size=3) * 0.3
        p = np.zeros(6)
        codes = np.zeros(2, dtype=np.int32)
        J0, J1 = np.zeros(36), np.zeros(36)
        xi0 = np.zeros(12)
        if H.ht_endpoint(a9, b9, cA, cB, xi0, p, codes, J0, J1) != 1:
            continue
        if codes[0] & 24 or codes[1] & 24:
            continue
        h = 1e-6
        num = np.zeros((6, 12))
        ok = True
        for k in range(12):
            for sgn, store in ((1, 0), (-1, 1)):
                xi = np.zeros(12)
                xi[k] = sgn * h
                q = np.zeros(6)
                c2 = np.zeros(2, dtype=np.int32)
                if H.ht_endpoint(a9, b9, cA, cB, xi, q, c2, np.zeros(36), np.zeros(36)) != 1 or not np.array_equal(c2, codes):
                    ok = False
                    break
                if sgn == 1:
                    qp = q
                else:
                    num[:, k] = (qp - q) / (2 * h)
            if not ok:
                break
        if not ok:
            continue
        ana = np.vstack([J0.reshape(3, 12), J1.reshape(3, 12)])
        scale = np.max(np.abs(ana)) + 1e-12
        assert np.max(np.abs(ana - num)) <= 1e-6 * scale, (codes, np.max(np.abs(ana - num)), scale)
        checked += 1
    assert checked >= 50


def test_loop_sensitivities_match_finite_differences(H):
    """dSn/dxi and dGn/dxi of every loop (vertex Jacobians = mean of the
    snapped endpoints' Jacobians, then the polyline sums) against central
    differences of the whole per-pair pipeline under rigid motions."""
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    H.ht_pair_jac.argtypes = [dp, C.c_int, dp, C.c_int, dp, dp, dp, dp, dp, C.c_double, dp, C.c_int, dp]
    m = GI.synth_meshes()
    cases = [("ellipsoid200", "ellipsoid200", [1.95, 0.0, 0.0], [1, 0, 0, 0]),
             ("hull128", "star2a", [1.5, 0.3, -0.1], [0.9, 0.1, 0.3, -0.2]),
             ("star2a", "star2b", [1.45, 0.1, -0.05], [0.8, -0.3, 0.2, 0.4])]
    done = 0
    for na, nb, pB, qB in cases:
        ta, pa, _ = O.build_mesh(m[na])
        tb, _, _ = O.build_mesh(m[nb])
        args = lambda xi, rec, lj: H.ht_pair_jac(np.ascontiguousarray(ta.reshape(-1)), len(ta),
                                                 np.ascontiguousarray(tb.reshape(-1)), len(tb), np.zeros(3),
                                                 np.array([1.0, 0, 0, 0]), np.asarray(pB, float),
                                                 np.asarray(qB, float) / np.linalg.norm(qB), xi, pa[5], rec, 64, lj)
        rec0, lj0 = np.zeros((64, 28)), np.zeros((64, 72))
        n0 = args(np.zeros(12), rec0, lj0)
        assert n0 >= 1, na
        h = 1e-7
        num = np.zeros((n0, 72))
        for k in range(12):
            out = []
            for sgn in (1, -1):
                xi = np.zeros(12)
                xi[k] = sgn * h
                rec, lj = np.zeros((64, 28)), np.zeros((64, 72))
                assert args(xi, rec, lj) == n0
                out.append(rec[:n0])
            d = (out[0] - out[1]) / (2 * h)
            for l in range(n0):
                num[l, 0 + k:36:12] = d[l, 0:3]
                num[l, 36 + k:72:12] = d[l, 3:6]
        ana = lj0[:n0]
        for l in range(n0):
            for part in (slice(0, 36), slice(36, 72)):
                scale = np.max(np.abs(ana[l, part])) + 1e-300
                assert np.max(np.abs(ana[l, part] - num[l, part])) <= 2e-5 * scale, (na, nb, l, part)
        done += 1
    assert done == len(cases)


def test_snap_clusters_found_beforehand_give_the_same_walk(H):
    """Walk B with snap clusters found over all of a pair's snap hits (each
    point's cluster named by its smallest point, as the GPU's cluster kernel
    finds them for pairs of 40 or more segments) gives bitwise the
    components, vertices, cells, edges and T-junction word counts of the
    per-component union-find in the walk (main.cpp:680-746): when no cluster
    spans two components no snap hit joins two components, and otherwise the
    walk unites per component itself.  Seeded concave and convex pairs, split
    into components, at the pipeline's tolerance and at four times it (more
    snaps, larger clusters)."""
    m = GI.synth_meshes()
    built = {k: O.build_mesh(v) for k, v in m.items()}
    rng = synth.Rng(91)
    names = list(m)
    n_multi = n_cases = n_big = 0
    for it in range(80):
        a, b = ("star2a", "star2b") if it % 2 == 0 else (names[int(rng.uniform(1)[0] * len(names))],
                                                           names[int(rng.uniform(1)[0] * len(names))])
        ta, pa, _ = built[a]
        tb, pb, _ = built[b]
        pose = GI._pose(rng, pa[4], pb[4], 0.05 + 0.35 * float(rng.uniform(1)[0]), True)
        cnt, seg, rec, vtx = host_pair(H, ta, tb, pose, pa[5], pa[6], True)
        S = len(seg)
        if S == 0:
            continue
        h0 = pa[5] * 0.1 if pa[5] > 0 else pa[6] * 1e-2
        for h in (h0, 4.0 * h0):
            out = np.zeros(4, dtype=np.int32)
            H.ht_snap_clusters(np.ascontiguousarray(seg.reshape(-1)), S, h, 1, out)
            assert out[0] == 1, (a, b, it, h, list(out))
            n_cases += 1
            n_multi += out[1] > 1
            n_big += S >= 40
    assert n_cases >= 40 and n_multi >= 5 and n_big >= 5, (n_cases, n_multi, n_big)
