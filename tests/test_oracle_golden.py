"""Pins the CPU oracle to the golden vectors recorded from the compiled
reference (tests/golden/make_golden.py).  Everything here is bit-exact: the
oracle restates src/main.cpp with the same operation order."""
import hashlib
import os
import tempfile

import numpy as np
import pytest

import golden_inputs as GI
import oracle as O
from zdx import synth


def unhex(lst):
    return np.array([float.fromhex(x) for x in lst])


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def test_mesh_prep_matches_reference(golden_geometry, ref_inputs):
    inputs = dict(GI.mesh_inputs(ref_inputs))
    checked = 0
    for g in golden_geometry["meshes"]:
        if g["name"].startswith("load:"):
            tris = O.load_stl(os.path.join(ref_inputs, "geometry", g["name"][5:]))
            assert tris.shape[0] == g["T"]
            assert sha(tris) == g["tris_sha"]
            assert np.array_equal(tris[0].reshape(-1), unhex(g["first"]))
            continue
        t = inputs[g["name"]]
        assert sha(t) == g["input_sha"], "input mesh drifted: " + g["name"]
        ot, props, nv = O.build_mesh(t, bool(g["center"]))
        assert nv == g["nverts"], g["name"]
        assert np.array_equal(props, unhex(g["props"])), g["name"]
        assert sha(ot) == g["tris_sha"], g["name"]
        checked += 1
    assert checked >= 20


def test_tri_tri_matches_reference(golden_geometry):
    g = golden_geometry["tri_tri"]
    A, B = GI.tri_pairs(g["seed"], g["n"])
    assert sha(np.concatenate([A, B])) == g["input_sha"]
    L = O.lib()
    flags, segs = [], []
    for i in range(len(A)):
        o = np.zeros(6)
        f = L.orc_tri_tri(np.ascontiguousarray(A[i].reshape(-1)), np.ascontiguousarray(B[i].reshape(-1)), o)
        flags.append(int(f))
        segs.append(o if f else np.zeros(6))
    segs = np.asarray(segs)
    assert flags == g["flags"]
    assert sum(flags) > 500            # plenty of intersecting pairs
    for i, s in enumerate(g["first"]):
        assert np.array_equal(segs[i], unhex(s))
    assert sha(segs) == g["segs_sha"]


def _pair_case_map(ref_inputs):
    return {c[0]: c for c in GI.pair_cases(ref_inputs)}


def test_pair_pipeline_matches_reference(golden_geometry, ref_inputs):
    cases = _pair_case_map(ref_inputs)
    n_loops = 0
    for g in golden_geometry["pairs"]:
        name, mA, mB, pose, split = cases[g["name"]]
        assert sha(mA) == g["inA_sha"] and sha(mB) == g["inB_sha"]
        for p, h in zip(pose, g["pose"]):
            assert np.array_equal(np.asarray(p, dtype=np.float64), unhex(h))
        ta, pa, _ = O.build_mesh(mA)
        tb, _, _ = O.build_mesh(mB)
        r = O.pair_contact(ta, tb, *pose, pa[5], pa[6], split=bool(split))
        assert r["tri_checks"] == g["tri_checks"], name
        assert len(r["segments"]) == g["segments"], name
        assert sha(r["segments"]) == g["segs_sha"], name
        assert r["comps"] == len(g["comps"]), name
        assert r["loops_raw"] == len(g["loops"]), name
        acc = [lo for lo in g["loops"] if lo["accepted"]]
        assert len(r["loops"]) == len(acc), name
        off = 0
        for rec, lo in zip(r["loops"], acc):
            nvtx = int(rec[24])
            pts = r["loop_vtx"][off:off + nvtx]
            off += nvtx
            assert int(rec[23]) == lo["comp"] and int(rec[25]) == lo["raw"], name
            assert np.array_equal(pts.reshape(-1), unhex(lo["pts"])), name
            assert np.array_equal(rec[0:3], unhex(lo["Sn"])), name
            assert np.array_equal(rec[3:6], unhex(lo["Gn"])), name
            assert np.array_equal(rec[6:9], unhex(lo["n"])), name
            assert np.array_equal(rec[9:12], unhex(lo["xc0"])), name
            assert rec[12] == float.fromhex(lo["area"]), name
            n_loops += 1
    assert n_loops >= 15


def run_oracle_sim(case, files, ref_inputs, tmp):
    os.symlink(os.path.join(ref_inputs, "geometry"), os.path.join(tmp, "geometry"))
    for fn, tris in files.items():
        synth.write_stl_ascii(os.path.join(tmp, fn), tris)
    cfg = os.path.join(tmp, case["name"] + ".txt")
    with open(cfg, "w") as f:
        f.write(case["config"])
    return O.Sim(cfg, tmp)


@pytest.mark.parametrize("name", ["cfg1_contact", "example_sim", "wall_drop", "cluster_box", "star_nosplit"])
def test_sim_trajectory_matches_reference(golden_sims, ref_inputs, name):
    case = [s for s in golden_sims if s["name"] == name][0]
    files = {fn: t for n, text, fs in GI.sim_cases(ref_inputs) if n == name for fn, t in fs.items()}
    for fn, h in case["files"].items():
        assert sha(files[fn]) == h
    with tempfile.TemporaryDirectory() as td:
        sim = run_oracle_sim(case, files, ref_inputs, td)
        snaps = {s["step"]: unhex(s["state"]).reshape(sim.n, 13) for s in case["snapshots"]}
        ctrs = {c["step"]: c for c in case["counters"]}
        last = max(snaps)
        for step in range(last + 1):
            ctr = sim.step(1)
            if step in ctrs:
                c = ctrs[step]
                got = dict(contacts=ctr[0], tri_checks=ctr[1], segments=ctr[2], comps=ctr[3], loops=ctr[4])
                for k, v in got.items():
                    assert v == c[k], (name, step, k)
                st = sim.state()
                mine = np.concatenate([st[:, 0:3], st[:, 3:7], st[:, 7:10], st[:, 10:13]], axis=1)
                assert np.array_equal(mine, snaps[step]), (name, step, np.abs(mine - snaps[step]).max())
