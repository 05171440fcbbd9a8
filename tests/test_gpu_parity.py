"""GPU parity: the HIP path (through the C ABI of libzdx.so) against the CPU
oracle on the same inputs.

Bit-exact: segments (count, order and every coordinate), components, traced
loops, loop vertices, Sn, Gn, n, xc0, area and the minimal-energy force
(Fn = k/2 |Sn|) -- all of these use only + - * / sqrt floor, which are
correctly rounded on both sides.  Trajectories of the full step go through
cbrt/pow/log of the contact force (not correctly rounded in either libm), so
they are compared with a relative tolerance stated per test.
"""
import os
import tempfile

import numpy as np
import pytest

import golden_inputs as GI
import oracle as O
from zdx import synth

zdx = pytest.importorskip("zdx")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = zdx.Context(0)
    yield c
    c.close()


def _oracle_pair(ta, tb, pose, pa, split, k=1.0):
    return O.pair_contact(ta, tb, *pose, pa[5], pa[6], split=split, k=k)


def _check_against_oracle(res, p, ref, name):
    lo, hi = res["loop_offset"][p], res["loop_offset"][p + 1]
    so, se = res["seg_offset"][p], res["seg_offset"][p + 1]
    assert res["tri_checks"][p] == ref["tri_checks"], name
    assert res["n_segments"][p] == len(ref["segments"]), name
    assert np.array_equal(res["segments"][so:se], ref["segments"]), name
    if len(ref["segments"]):
        assert res["n_comps"][p] == ref["comps"], name
        assert res["n_loops_raw"][p] == ref["loops_raw"], name
    assert hi - lo == len(ref["loops"]), name
    got = res["loops"][lo:hi]
    # Sn Gn n xc0 area Fn F torqueA torqueB comp nverts raw
    assert np.array_equal(got[:, :26], ref["loops"][:, :26]), name
    v0, v1 = res["vtx_offset"][lo], res["vtx_offset"][hi]
    assert np.array_equal(res["loop_vtx"][v0:v1], ref["loop_vtx"]), name


@pytest.mark.parametrize("mode", [0, 1])
def test_pairs_match_oracle_golden_cases(ctx, ref_inputs, mode):
    cases = GI.pair_cases(ref_inputs)
    for split in (True, False):
        ids, refs = [], []
        for name, mA, mB, pose, sp in cases:
            if sp != split:
                continue
            a, b = ctx.register_mesh(mA), ctx.register_mesh(mB)
            ta, pa, _ = O.build_mesh(mA)
            tb, _, _ = O.build_mesh(mB)
            ids.append((name, a, b, pose))
            refs.append(_oracle_pair(ta, tb, pose, pa, split, k=2.5))
        res = ctx.contact_pairs([x[1] for x in ids], [x[2] for x in ids], [x[3][0] for x in ids],
                                [x[3][1] for x in ids], [x[3][2] for x in ids], [x[3][3] for x in ids], split=split,
                                mode=mode, k=2.5, segments=True)
        for p, (x, ref) in enumerate(zip(ids, refs)):
            _check_against_oracle(res, p, ref, x[0])


def _oracle_tol(t):
    """t nudged up to the nearest double that the oracle's rule (mean edge *
    0.1) reproduces exactly from the mean edge t / 0.1."""
    for _ in range(64):
        if (t / 0.1) * 0.1 == t:
            return t
        t = float(np.nextafter(t, np.inf))
    raise AssertionError(t)


@pytest.mark.parametrize("factor", [0.5, 2.0, 4.0])
def test_pairs_tol_override_matches_oracle(ctx, ref_inputs, factor):
    """opts.tol (polyline.py --tol) replaces mesh A's mean_edge * 0.1 for every
    pair; the oracle gets a mean edge whose tenth is exactly that tol.  The
    tolerances are kept near the default one: the reference's T-junction
    search visits (edge length / tol)^3 grid cells per edge, which the oracle
    repeats (the device does not scan cells)."""
    cases = GI.pair_cases(ref_inputs)
    for split in (True, False):
        ids, refs, tols = [], [], []
        for name, mA, mB, pose, sp in cases:
            a, b = ctx.register_mesh(mA), ctx.register_mesh(mB)
            ta, pa, _ = O.build_mesh(mA)
            tb, _, _ = O.build_mesh(mB)
            tol = _oracle_tol(factor * (pa[5] * 0.1))
            ids.append((name, a, b, pose))
            tols.append(tol)
            refs.append(O.pair_contact(ta, tb, *pose, tol / 0.1, pa[6], split=split, k=1.5))
        # one call per pair: the tolerance is per call
        for x, ref, tol in zip(ids, refs, tols):
            res = ctx.contact_pairs([x[1]], [x[2]], [x[3][0]], [x[3][1]], [x[3][2]], [x[3][3]], split=split,
                                    mode=1, k=1.5, segments=True, tol=tol)
            _check_against_oracle(res, 0, ref, (x[0], factor, split))


def test_pairs_random_batch_plain_equals_batched_and_oracle(ctx):
    """A batch of uneven pairs (icospheres of 80-1280 triangles and concave
    stars) in one call; plain and batched must be bitwise identical, and each
    pair must equal the oracle."""
    rng = synth.Rng(2024)
    meshes = {"ico1": synth.icosphere(1), "ico2": synth.icosphere(2), "ico3": synth.icosphere(3),
              "star2a": synth.star_mesh(7, 2), "star3": synth.star_mesh(9, 3), "hull64": synth.sphere_hull(64, 5)}
    ids = {k: ctx.register_mesh(v) for k, v in meshes.items()}
    built = {k: O.build_mesh(v) for k, v in meshes.items()}
    names = list(meshes)
    A, B, poses = [], [], []
    for it in range(96):
        a = names[int(rng.uniform(1)[0] * len(names))]
        b = names[int(rng.uniform(1)[0] * len(names))]
        dep = 0.01 + 0.09 * float(rng.uniform(1)[0]) if it % 3 else 0.1 + 0.3 * float(rng.uniform(1)[0])
        poses.append(GI._pose(rng, built[a][1][4], built[b][1][4], dep, True))
        A.append(a)
        B.append(b)
    args = ([ids[a] for a in A], [ids[b] for b in B], [p[0] for p in poses], [p[1] for p in poses],
            [p[2] for p in poses], [p[3] for p in poses])
    r0 = ctx.contact_pairs(*args, split=True, mode=0, k=1.0, segments=True)
    r1 = ctx.contact_pairs(*args, split=True, mode=1, k=1.0, segments=True)
    for key in ("tri_checks", "n_segments", "n_comps", "n_loops_raw", "loop_offset", "loops", "loop_vtx", "segments"):
        assert np.array_equal(r0[key], r1[key]), key
    n_loops = 0
    for p in range(len(A)):
        ta, pa, _ = built[A[p]]
        tb, _, _ = built[B[p]]
        ref = _oracle_pair(ta, tb, poses[p], pa, True)
        _check_against_oracle(r1, p, ref, f"{A[p]}-{B[p]}-{p}")
        n_loops += len(ref["loops"])
    assert n_loops > 40


def test_pairs_empty_and_disjoint(ctx):
    m = ctx.register_mesh(synth.icosphere(1))
    I = [1.0, 0.0, 0.0, 0.0]
    res = ctx.contact_pairs([m, m], [m, m], [[0, 0, 0], [0, 0, 0]], [I, I], [[5.0, 0, 0], [0.0, 0.0, 2.5]], [I, I],
                            segments=True)
    assert res["n_segments"].tolist() == [0, 0]
    assert len(res["loops"]) == 0
    res = ctx.contact_pairs([], [], [], [], [], [], segments=True)
    assert len(res["loops"]) == 0


def _sim_case(golden_sims, ref_inputs, name, td):
    case = [s for s in golden_sims if s["name"] == name][0]
    files = {fn: t for n, text, fs in GI.sim_cases(ref_inputs) if n == name for fn, t in fs.items()}
    os.symlink(os.path.join(ref_inputs, "geometry"), os.path.join(td, "geometry"))
    for fn, tris in files.items():
        synth.write_stl_ascii(os.path.join(td, fn), tris)
    cfg = os.path.join(td, name + ".txt")
    with open(cfg, "w") as f:
        f.write(case["config"])
    return case, cfg


# relative tolerance of trajectories against the oracle after the stated
# number of steps: forces go through cbrt(pow(A, 1.5)) and log(e), whose last
# bit differs between the device libm and glibc
SIM_CASES = [("cfg1_contact", 1000, 1e-9), ("star_nosplit", 300, 1e-9), ("wall_drop", 4000, 1e-7),
             ("cluster_box", 400, 1e-6), ("example_sim", 6000, 1e-8)]


@pytest.mark.parametrize("name,steps,rtol", SIM_CASES)
def test_sim_matches_oracle(ctx, golden_sims, ref_inputs, name, steps, rtol):
    with tempfile.TemporaryDirectory() as td:
        case, cfg = _sim_case(golden_sims, ref_inputs, name, td)
        osim = O.Sim(cfg, td)
        gsim = zdx.Sim.from_config(ctx, cfg, td)
        chunk = 50
        done = 0
        topo_equal_steps = 0
        while done < steps:
            k = min(chunk, steps - done)
            for _ in range(k):
                oc = osim.step(1)
                gc = gsim.step(1, zdx.MODE_PLAIN)
                if [gc["contacts"], gc["tri_checks"], gc["segments"], gc["comps"], gc["loops"]] == oc[:5].tolist():
                    topo_equal_steps += 1
            done += k
            so, sg = osim.state(), gsim.state()
            scale = np.maximum(np.abs(so), 1e-3)
            err = np.max(np.abs(so - sg) / scale)
            assert err <= rtol, (name, done, err)
        # counters identical on (almost) every step: a flip would need a
        # decision within ~1e-12 of its threshold
        assert topo_equal_steps >= steps - 2, (name, topo_equal_steps, steps)
        F1, T1, c1 = osim.forces()
        F2, T2, c2 = gsim.forces()
        assert np.array_equal(c1, c2)


def test_sim_plain_equals_batched(ctx, golden_sims, ref_inputs):
    with tempfile.TemporaryDirectory() as td:
        case, cfg = _sim_case(golden_sims, ref_inputs, "cluster_box", td)
        a = zdx.Sim.from_config(ctx, cfg, td)
        b = zdx.Sim.from_config(ctx, cfg, td)
        for _ in range(5):
            ca = a.step(40, zdx.MODE_PLAIN)
            cb = b.step(40, zdx.MODE_BATCHED)
            assert ca == cb
            assert np.array_equal(a.state(), b.state())
        Fa, Ta, _ = a.forces()
        Fb, Tb, _ = b.forces()
        assert np.array_equal(Fa, Fb) and np.array_equal(Ta, Tb)


def test_sim_checkpoint_restore_is_exact(ctx, golden_sims, ref_inputs):
    with tempfile.TemporaryDirectory() as td:
        case, cfg = _sim_case(golden_sims, ref_inputs, "cluster_box", td)
        a = zdx.Sim.from_config(ctx, cfg, td)
        a.step(100, zdx.MODE_PLAIN)
        ck = os.path.join(td, "ck.bin")
        a.checkpoint(ck)
        a.step(60, zdx.MODE_PLAIN)
        sa = a.state()
        b = zdx.Sim.from_config(ctx, cfg, td)
        b.restore(ck)
        b.step(60, zdx.MODE_BATCHED)
        assert np.array_equal(sa, b.state())


@pytest.mark.parametrize("depth,off,pa", [(0.2, (0.0, 0.0), (0.0, 0.0)), (0.3, (0.2, 0.2), (0.3, -0.7)),
                                           (0.45, (0.3, -0.3), (-1.1, 0.25))])
def test_box_on_box_closed_form(ctx, depth, off, pa):
    """|Sn| = overlap area, Fn = k/2 |Sn| (= k/2 dOmega/ddelta with Omega =
    delta A), n from A to B, xc0 = -2/3 of the rectangle centre's in-plane
    part -- the closed form the oracle is pinned with, on the device path."""
    k = 3.0
    a, b = ctx.register_mesh(synth.box_wall(1.0, 1.2, 0.8)), ctx.register_mesh(synth.box_wall(2, 2, 2))
    pA = np.array([0.0, pa[0], pa[1]])
    pB = pA + np.array([0.5 - depth + 1.0, off[0], off[1]])
    for mode in (zdx.MODE_PLAIN, zdx.MODE_BATCHED):
        r = ctx.contact_pairs([a], [b], [pA], [[1, 0, 0, 0]], [pB], [[1, 0, 0, 0]], mode=mode, k=k)
        assert len(r["loops"]) == 1
        L = r["loops"][0]
        A = 1.2 * 0.8
        assert L[12] == pytest.approx(A, rel=1e-13)
        assert np.allclose(L[6:9], [1.0, 0.0, 0.0], atol=1e-13)
        assert L[13] == pytest.approx(0.5 * k * A, rel=1e-13)
        assert np.allclose(L[9:12], -2.0 / 3.0 * np.array([0.0, pa[0], pa[1]]), atol=1e-13)
