"""The large single-GPU configuration: 1M concave stars (16 templates of
320-2560 triangles, radii 1-4 mm) in a 0.5 m box with a 12-triangle box wall,
gravity, dt 5e-6 (zdx.workloads.concave_packing), stepped on one GPU.

Checks, on the first five steps: the first step's candidate pairs are
exactly the reference's sphere-test set (src/main.cpp:2049-2059) of the
seeded state (about 2.1M pairs); every step has contacts and finite state;
two batched steps equal two plain steps bitwise from the same checkpoint.
The device memory the simulation holds is printed and must stay far below
the 288 GB of one MI355X.
"""
import os
import tempfile
import time

import numpy as np
import pytest

from zdx import workloads as W

zdx = pytest.importorskip("zdx")
pytestmark = pytest.mark.gpu


def test_packing_1m_steps_on_one_gpu():
    from scipy.spatial import cKDTree
    import torch

    t0 = time.time()
    wl = W.concave_packing()
    free0, total = torch.cuda.mem_get_info()
    ctx = zdx.Context(0)
    sim, n_meshes = W.build_sim(ctx, wl)
    print(f"[1m] set-up {time.time() - t0:.1f} s: {wl.n} particles, {n_meshes} meshes", flush=True)
    rad = np.array([ctx.mesh_props[m][4] for m in sim.mesh_id])
    s0 = sim.state()
    assert np.array_equal(s0[:, :3], wl.init[:, :3])
    t1 = time.time()
    c = sim.step(1, zdx.MODE_BATCHED)
    print(f"[1m] step 1: {time.time() - t1:.2f} s wall, {c}", flush=True)
    got = sim.pairs()
    cand = cKDTree(s0[:, :3]).query_pairs(2.0 * rad.max() * (1 + 1e-9) + 1e-12, output_type="ndarray")
    cand = cand[cand[:, 0] < cand[:, 1]] if len(cand) else cand
    dp = s0[cand[:, 1], :3] - s0[cand[:, 0], :3]
    d2 = dp[:, 0] * dp[:, 0] + dp[:, 1] * dp[:, 1] + dp[:, 2] * dp[:, 2]
    rs = rad[cand[:, 0]] + rad[cand[:, 1]]
    want = cand[~(d2 > rs * rs)]
    want = want[np.lexsort((want[:, 1], want[:, 0]))]
    print(f"[1m] pairs: device {len(got)}, reference test {len(want)}", flush=True)
    assert len(want) > 1_000_000
    assert np.array_equal(got, want)
    assert c["pairs"] == len(want) and c["contacts"] > 10000
    for k in range(2):
        t1 = time.time()
        c = sim.step(1, zdx.MODE_BATCHED)
        print(f"[1m] step {k + 2}: {time.time() - t1:.3f} s wall, contacts {c['contacts']} wall contacts "
              f"{c['wall_contacts']} segments {c['segments']}", flush=True)
        assert c["contacts"] > 10000
    st = sim.stage_times()
    free1, _ = torch.cuda.mem_get_info()
    used = (free0 - free1) / 2 ** 30
    print(f"[1m] device memory held: {used:.1f} GiB of {total / 2 ** 30:.0f}; stages {st}", flush=True)
    assert used < 200
    with tempfile.TemporaryDirectory() as td:
        ck = os.path.join(td, "ck.bin")
        sim.checkpoint(ck)
        cb = [sim.step(1, zdx.MODE_BATCHED) for _ in range(2)]
        sb = sim.state()
        sim.restore(ck)
        cp = [sim.step(1, zdx.MODE_PLAIN) for _ in range(2)]
        sp = sim.state()
    assert cb == cp
    assert np.isfinite(sp).all()
    assert np.array_equal(sb, sp)
    ctx.close()
