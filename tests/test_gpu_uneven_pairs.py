"""The uneven-pair batch on the device: 4096 independent pairs (half
icosphere pairs of levels 1-4, 80-5120 triangles per body; half concave star
pairs of 1280 triangles) in ONE batched zdx_contact_pairs call, the per-pair
body of the reference step (src/main.cpp:2061-2219, narrow loop 2123-2145)
with the minimal-energy force of polyline.py:655-701.

* Every pair with an ico4 body (2562 world vertices: more than the cull's
  per-side outcode stage holds, so the cull takes its staged path) and a
  seeded sample of the others is compared bit-exactly with the oracle, inside
  the full 4096-pair batch: segments, counters, loop records and loop
  vertices.
* All 4096: plain mode equals batched mode bitwise.
* All 4096, size-independent properties of the contact geometry: a common
  translation of both bodies leaves the area |Sn| and the A->B normal n
  unchanged, and, for pairs of one mesh with itself, swapping A and B keeps
  |Sn| and xc0 and reverses n (xc0 = (nA x Gn)/|Sn| does not depend on the
  loop's walk direction).  The reference's snapping and loop walk depend on
  segment order and on rounding, so a complex concave contact can be routed
  differently after a translation or a swap; at most 2 % of the pairs may
  differ, every other pair to 1e-9.
"""
import numpy as np
import pytest

import oracle as O
from zdx import workloads as W

zdx = pytest.importorskip("zdx")
pytestmark = pytest.mark.gpu

N_PAIRS = 4096


@pytest.fixture(scope="module")
def batch():
    ctx = zdx.Context(0)
    meshes, names, pairs = W.uneven_pairs(N_PAIRS)
    ids = [ctx.register_mesh(m) for m in meshes]
    a = np.array([p[0] for p in pairs])
    b = np.array([p[1] for p in pairs])
    P = dict(mesh_a=[ids[i] for i in a], mesh_b=[ids[i] for i in b], pos_a=np.array([p[2] for p in pairs]),
             quat_a=np.array([p[3] for p in pairs]), pos_b=np.array([p[4] for p in pairs]),
             quat_b=np.array([p[5] for p in pairs]))
    print(f"[uneven] {N_PAIRS} pairs, meshes {names}", flush=True)
    res = ctx.contact_pairs(**P, split=True, mode=zdx.MODE_BATCHED, k=2.0, segments=True)
    yield ctx, meshes, names, a, b, P, res
    ctx.close()


def _loops(res, p):
    return res["loops"][res["loop_offset"][p]:res["loop_offset"][p + 1]]


def test_uneven_batch_has_contacts(batch):
    ctx, meshes, names, a, b, P, res = batch
    nl = np.diff(res["loop_offset"])
    # every pair interpenetrates along its line of centres
    assert (res["n_segments"] > 0).all()
    assert (nl > 0).mean() > 0.98
    assert res["tri_checks"].sum() > 5e9


def test_uneven_subset_matches_oracle(batch):
    ctx, meshes, names, a, b, P, res = batch
    built = [O.build_mesh(m) for m in meshes]
    ico4 = names.index("ico4")
    sel = set(np.nonzero((a == ico4) | (b == ico4))[0].tolist())
    rng = np.random.default_rng(7)
    sel |= set(rng.choice(np.nonzero((a != ico4) & (b != ico4))[0], 160, replace=False).tolist())
    assert sum(1 for p in sel if a[p] == ico4 and b[p] == ico4) > 100
    for p in sorted(sel):
        ta, pa, _ = built[a[p]]
        tb, _, _ = built[b[p]]
        ref = O.pair_contact(ta, tb, P["pos_a"][p], P["quat_a"][p], P["pos_b"][p], P["quat_b"][p], pa[5], pa[6],
                             split=True, k=2.0)
        name = (p, names[a[p]], names[b[p]])
        assert res["tri_checks"][p] == ref["tri_checks"], name
        so, se = res["seg_offset"][p], res["seg_offset"][p + 1]
        assert np.array_equal(res["segments"][so:se], ref["segments"]), name
        assert res["n_comps"][p] == ref["comps"], name
        assert res["n_loops_raw"][p] == ref["loops_raw"], name
        got = _loops(res, p)
        assert len(got) == len(ref["loops"]), name
        assert np.array_equal(got[:, :26], ref["loops"][:, :26]), name
        lo, hi = res["loop_offset"][p], res["loop_offset"][p + 1]
        v0, v1 = res["vtx_offset"][lo], res["vtx_offset"][hi]
        assert np.array_equal(res["loop_vtx"][v0:v1], ref["loop_vtx"]), name


def test_uneven_plain_equals_batched(batch):
    ctx, meshes, names, a, b, P, res = batch
    r0 = ctx.contact_pairs(**P, split=True, mode=zdx.MODE_PLAIN, k=2.0, segments=True)
    for key in ("tri_checks", "n_segments", "n_comps", "n_loops_raw", "loop_offset", "seg_offset", "loops",
                "vtx_offset", "loop_vtx", "segments"):
        assert np.array_equal(r0[key], res[key]), key


def _match(l1, l2, kind):
    """Worst relative difference of two loop sets of one pair, matched by
    area: area and n (kind 'shift'), or area, -n and xc0 (kind 'swap')."""
    if len(l1) != len(l2):
        return np.inf
    w = 0.0
    for x, y in zip(sorted(l1, key=lambda r: -r[12]), sorted(l2, key=lambda r: -r[12])):
        w = max(w, abs(x[12] - y[12]) / x[12])
        if kind == "shift":
            w = max(w, np.abs(x[6:9] - y[6:9]).max())
        else:
            w = max(w, np.abs(x[6:9] + y[6:9]).max(), np.abs(x[9:12] - y[9:12]).max() / max(np.abs(x[9:12]).max(), 1.0))
    return w


def test_uneven_translation_invariance(batch):
    ctx, meshes, names, a, b, P, res = batch
    s = np.array([0.37, -1.21, 2.53])
    Q = dict(P, pos_a=P["pos_a"] + s, pos_b=P["pos_b"] + s)
    r = ctx.contact_pairs(**Q, split=True, mode=zdx.MODE_BATCHED, k=2.0)
    w = np.array([_match(_loops(res, p), _loops(r, p), "shift") for p in range(N_PAIRS)])
    bad = np.nonzero(w > 1e-9)[0]
    print(f"[uneven] translation: {len(bad)} of {N_PAIRS} pairs re-routed", flush=True)
    assert len(bad) <= 0.02 * N_PAIRS, bad[:20]


def test_uneven_swap_reverses_normal(batch):
    ctx, meshes, names, a, b, P, res = batch
    same = np.nonzero(a == b)[0]
    assert len(same) > 1000
    Q = dict(mesh_a=[P["mesh_b"][p] for p in same], mesh_b=[P["mesh_a"][p] for p in same], pos_a=P["pos_b"][same],
             quat_a=P["quat_b"][same], pos_b=P["pos_a"][same], quat_b=P["quat_a"][same])
    r = ctx.contact_pairs(**Q, split=True, mode=zdx.MODE_BATCHED, k=2.0)
    w = np.array([_match(_loops(res, p), _loops(r, i), "swap") for i, p in enumerate(same)])
    bad = np.nonzero(w > 1e-9)[0]
    print(f"[uneven] swap: {len(bad)} of {len(same)} pairs re-routed", flush=True)
    assert len(bad) <= 0.02 * len(same), same[bad[:20]]
