"""Broad phase on the device against the reference's pair test
(src/main.cpp:2049-2059: every i < j, dpos = x_j - x_i, dist2 summed x, y, z,
rejected when dist2 > (r_i + r_j)^2): the cell-binned pair set must be
exactly that set.

* A hub of one large body (radius 10 mm) ringed by 300 small hulls whose
  bounding spheres reach it: the hub has 300 partners, far more than the
  first-pass partner list holds, so its partners come from the cell rescan.
  Pair set exact; trajectories against the oracle for 20 steps.
* The 100k granular bed after 300 settle steps (about 12k candidate pairs):
  the pair set of the next step exact against a k-d tree superset filtered
  with the reference's test; plain and batched steps bitwise equal over 50
  steps.
"""
import math
import os
import tempfile

import numpy as np
import pytest

import oracle as O
from zdx import synth
from zdx import workloads as W

zdx = pytest.importorskip("zdx")
pytestmark = pytest.mark.gpu


def exact_pairs(P, rad, cand):
    """The reference's sphere test on candidate pairs (a < b), same doubles."""
    if len(cand) == 0:
        return cand.reshape(0, 2)
    dp = P[cand[:, 1]] - P[cand[:, 0]]
    d2 = dp[:, 0] * dp[:, 0] + dp[:, 1] * dp[:, 1] + dp[:, 2] * dp[:, 2]
    rs = rad[cand[:, 0]] + rad[cand[:, 1]]
    keep = cand[~(d2 > rs * rs)]
    return keep[np.lexsort((keep[:, 1], keep[:, 0]))]


def all_pairs(n):
    i, j = np.triu_indices(n, 1)
    return np.stack([i, j], axis=1)


def hub_workload(n_small=300, seed=12):
    g = np.random.Generator(np.random.PCG64(seed))
    big = synth.icosphere(3)
    small = synth.sphere_hull(64, seed)
    rs = W._centred_radius(small)
    tris = [big, small]
    d = g.normal(size=(n_small, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = np.concatenate([np.zeros((1, 3)), d * 0.0102], axis=0)
    n = n_small + 1
    q = g.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    init = np.concatenate([pos, q, np.zeros((n, 6))], axis=1)
    tpl = np.array([0] + [1] * n_small)
    scale = np.array([0.01] + [round(0.5e-3 / rs, 6)] * n_small)
    mat = np.tile(np.array([2500.0, 1e7, 0.25, 0.5, 0.5]), (n, 1))
    return W.Workload("hub", tris, ["ico3", "hull64"], tpl, scale, init, mat, [], (0.0, 0.0, 0.0), 1e-5, 20, True)


def test_hub_with_300_partners_matches_reference_test_and_oracle():
    wl = hub_workload()
    with tempfile.TemporaryDirectory() as td:
        cfg = W.write_reference_config(wl, range(wl.n), td, steps=20, output_interval=1000)
        ctx = zdx.Context(0)
        sim = zdx.Sim.from_config(ctx, cfg, td)
        osim = O.Sim(cfg, td)
        s0 = sim.state()
        so0 = osim.state()
        assert np.array_equal(s0, so0)
        # bounding radii as the step uses them: build_mesh of the scaled
        # templates (the oracle's is bit-exact with the reference's)
        r_t = {t: O.build_mesh(wl.templates[t] * wl.scale[np.argmax(wl.tpl == t)])[1][4] for t in (0, 1)}
        rad = np.array([r_t[wl.tpl[i]] for i in range(wl.n)])
        ctr = sim.step(1, zdx.MODE_BATCHED)
        got = sim.pairs()
        want = exact_pairs(s0[:, :3], rad, all_pairs(wl.n))
        hub = int((want[:, 0] == 0).sum())
        print(f"[hub] {len(want)} pairs, hub partners {hub}, active contacts {ctr['contacts']}", flush=True)
        assert hub >= 250
        assert np.array_equal(got, want), (len(got), len(want))
        oc = osim.step(1)
        assert ctr["pairs"] == len(want)
        for _ in range(19):
            sim.step(1, zdx.MODE_BATCHED)
            osim.step(1)
        so, sg = osim.state(), sim.state()
        err = np.max(np.abs(so - sg) / np.maximum(np.abs(so), 1e-3))
        assert err <= 1e-9, err
        ctx.close()


@pytest.fixture(scope="module")
def settled_bed():
    ctx = zdx.Context(0)
    wl = W.granular_bed(100_000)
    sim, _ = W.build_sim(ctx, wl)
    sim.step(300, zdx.MODE_BATCHED)
    yield ctx, wl, sim
    ctx.close()


def test_bed_pair_set_is_the_reference_set(settled_bed):
    from scipy.spatial import cKDTree

    ctx, wl, sim = settled_bed
    s = sim.state()
    rad = np.array([ctx.mesh_props[m][4] for m in sim.mesh_id])
    sim.step(1, zdx.MODE_BATCHED)
    got = sim.pairs()
    cand = np.array(sorted(cKDTree(s[:, :3]).query_pairs(2.0 * rad.max() * (1 + 1e-9) + 1e-12)), dtype=np.int64)
    want = exact_pairs(s[:, :3], rad, cand)
    deg = np.bincount(want.ravel(), minlength=wl.n)
    print(f"[bed] {len(want)} candidate pairs, most partners of one particle {deg.max()}", flush=True)
    assert len(want) > 5000
    assert np.array_equal(got, want)


def test_bed_plain_equals_batched_50_steps(settled_bed):
    ctx, wl, sim = settled_bed
    with tempfile.TemporaryDirectory() as td:
        ck = os.path.join(td, "ck.bin")
        sim.checkpoint(ck)
        cb = [sim.step(1, zdx.MODE_BATCHED) for _ in range(50)]
        sb = sim.state()
        sim.restore(ck)
        cp = [sim.step(1, zdx.MODE_PLAIN) for _ in range(50)]
        sp = sim.state()
    assert cb == cp
    assert sum(c["contacts"] for c in cb) > 10000
    assert np.array_equal(sb, sp)
