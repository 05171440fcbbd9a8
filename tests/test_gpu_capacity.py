"""Grow-and-retry paths: with ZDX_START_CAPACITY=tiny every growable device
buffer of a context or simulation starts at a few entries (pairs, hits,
segment staging, loops per pair, topology rows, snap and T-junction words,
topology work units, wall section points), so the first steps and calls
overflow each of them, grow and run again from the same state.  The results
must be bitwise those of a normally sized run; and a caller's output arrays
that are too small give ZDX_CAPACITY with the needed totals filled in.
"""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

import golden_inputs as GI
from zdx import workloads as W

zdx = pytest.importorskip("zdx")
pytestmark = pytest.mark.gpu


def _tiny_ctx():
    os.environ["ZDX_START_CAPACITY"] = "tiny"
    try:
        return zdx.Context(0)
    finally:
        del os.environ["ZDX_START_CAPACITY"]


def _tiny(fn):
    os.environ["ZDX_START_CAPACITY"] = "tiny"
    try:
        return fn()
    finally:
        del os.environ["ZDX_START_CAPACITY"]


@pytest.mark.parametrize("mode", [0, 1])
def test_tiny_capacities_cluster_box_bitwise(golden_sims, ref_inputs, mode):
    case = [s for s in golden_sims if s["name"] == "cluster_box"][0]
    files = {fn: t for n, text, fs in GI.sim_cases(ref_inputs) if n == "cluster_box" for fn, t in fs.items()}
    from zdx import synth

    with tempfile.TemporaryDirectory() as td:
        os.symlink(os.path.join(ref_inputs, "geometry"), os.path.join(td, "geometry"))
        for fn, tris in files.items():
            synth.write_stl_ascii(os.path.join(td, fn), tris)
        cfg = os.path.join(td, "cluster_box.txt")
        with open(cfg, "w") as f:
            f.write(case["config"])
        ca, cb = zdx.Context(0), zdx.Context(0)
        a = zdx.Sim.from_config(ca, cfg, td)
        b = _tiny(lambda: zdx.Sim.from_config(cb, cfg, td))
        b.stage_times()
        for _ in range(8):
            x = a.step(50, mode)
            y = b.step(50, mode)
            assert x == y
            assert np.array_equal(a.state(), b.state())
        st = b.stage_times()
        print(f"[capacity] cluster_box mode {mode}: {st['retries']} steps run again", flush=True)
        assert st["retries"] >= 3
        ca.close()
        cb.close()


def test_tiny_capacities_bed_2k_bitwise():
    wl = W.granular_bed(2000, box=0.06, speed=0.8)
    ca, cb = zdx.Context(0), zdx.Context(0)
    a, _ = W.build_sim(ca, wl)
    b, _ = _tiny(lambda: W.build_sim(cb, wl))
    b.stage_times()
    tot = 0
    for it in range(12):
        mode = it % 2
        x = a.step(25, mode)
        y = b.step(25, mode)
        assert x == y
        assert np.array_equal(a.state(), b.state())
        tot += x["contacts"] + x["wall_contacts"]
    st = b.stage_times()
    print(f"[capacity] bed 2k: {st['retries']} steps run again, contacts {tot}", flush=True)
    assert tot > 100
    assert st["retries"] >= 3
    ca.close()
    cb.close()


def test_tiny_capacities_uneven_pairs_bitwise():
    meshes, names, pairs = W.uneven_pairs(512, seed=5)
    out = []
    for ctx in (zdx.Context(0), _tiny_ctx()):
        ids = [ctx.register_mesh(m) for m in meshes]
        r = ctx.contact_pairs([ids[p[0]] for p in pairs], [ids[p[1]] for p in pairs], [p[2] for p in pairs],
                              [p[3] for p in pairs], [p[4] for p in pairs], [p[5] for p in pairs], split=True,
                              mode=zdx.MODE_BATCHED, k=1.0, segments=True)
        out.append((r, ctx.contact_times()["attempts"]))
        ctx.close()
    (r0, a0), (r1, a1) = out
    print(f"[capacity] uneven 512: attempts {a0} / {a1}", flush=True)
    assert a1 > 1
    for key in ("tri_checks", "n_segments", "n_comps", "n_loops_raw", "loop_offset", "loops", "loop_vtx",
                "segments"):
        assert np.array_equal(r0[key], r1[key]), key


def test_contact_pairs_reports_capacity_and_totals():
    """The raw ABI: loop, vertex and segment arrays one entry too small give
    ZDX_CAPACITY with n_*_total filled; exactly sized arrays then succeed."""
    ctx = zdx.Context(0)
    meshes, names, pairs = W.uneven_pairs(64, seed=6)
    ids = [ctx.register_mesh(m) for m in meshes]
    n = len(pairs)
    arr = np.zeros(n, dtype=zdx.PAIR_IN_DTYPE)
    arr["mesh_a"] = [ids[p[0]] for p in pairs]
    arr["mesh_b"] = [ids[p[1]] for p in pairs]
    arr["pos_a"] = [p[2] for p in pairs]
    arr["quat_a"] = [p[3] for p in pairs]
    arr["pos_b"] = [p[4] for p in pairs]
    arr["quat_b"] = [p[5] for p in pairs]
    opts = zdx.ContactOpts(1, 1, 1.0, 0, 0, 0.0)
    L = zdx.lib()

    def call(lc, vc, sc):
        lo = np.zeros(n + 1, dtype=np.int32)
        so = np.zeros(n + 1, dtype=np.int32)
        rec = np.zeros((max(lc, 1), 28))
        vo = np.zeros(max(lc, 1) + 1, dtype=np.int32)
        vtx = np.zeros((max(vc, 1), 3))
        seg = np.zeros((max(sc, 1), 6))
        out = zdx.PairOut(None, None, None, None, lo.ctypes.data, so.ctypes.data, rec.ctypes.data, lc,
                          vo.ctypes.data, vtx.ctypes.data, vc, seg.ctypes.data, sc, 0, 0, 0, None)
        st = L.zdx_contact_pairs(ctx.h, arr.ctypes.data, n, C.byref(opts), C.byref(out))
        return st, out, rec, vtx, seg

    st, out, _, _, _ = call(1, 1, 1)
    assert st == zdx.CAPACITY
    nl, nv, ns = out.n_loops_total, out.n_vtx_total, out.n_seg_total
    assert nl > 10 and nv > nl and ns > nl
    for caps in ((nl - 1, nv, ns), (nl, nv - 1, ns), (nl, nv, ns - 1)):
        st, out, _, _, _ = call(*caps)
        assert st == zdx.CAPACITY, caps
        assert (out.n_loops_total, out.n_vtx_total, out.n_seg_total) == (nl, nv, ns)
    st, out, rec, vtx, seg = call(nl, nv, ns)
    assert st == zdx.OK
    ref = ctx.contact_pairs(list(arr["mesh_a"]), list(arr["mesh_b"]), arr["pos_a"], arr["quat_a"], arr["pos_b"],
                            arr["quat_b"], split=True, mode=1, k=1.0, segments=True)
    assert np.array_equal(rec, ref["loops"]) and np.array_equal(vtx, ref["loop_vtx"])
    assert np.array_equal(seg, ref["segments"])
    ctx.close()
