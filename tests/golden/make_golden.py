"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Run in the container that holds the reference source (it needs
oracle/_ref/libzdx_ref.so, built by `make -C oracle ref`):

    python tests/golden/make_golden.py

Every expected output comes from the reference's own functions (src/main.cpp,
exposed by oracle/ref_harness.cpp) or from its own zdem_cpu main loop run on a
config file; nothing here computes an expected value itself.  Inputs are
either data files the reference ships (geometry/*.stl, config/*.txt, copied
verbatim into ref_inputs/) or seeded synthetic meshes from zdx.synth, whose
bytes are fingerprinted (sha256) in the fixture so input drift is detected.
Doubles are stored as float.hex() strings (exact).
"""
from __future__ import annotations

import contextlib
import ctypes as C
import hashlib
import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "zedm-x_amd"))
sys.path.insert(0, HERE)

import oracle as O  # noqa: E402
from zdx import synth  # noqa: E402
import golden_inputs as GI  # noqa: E402

REF_GEOM = os.path.join(O.REF_DIR, "geometry")
REF_CFG = os.path.join(O.REF_DIR, "config")
INPUTS = os.path.join(HERE, "ref_inputs")


def hx(a) -> list:
    return [float(x).hex() for x in np.asarray(a, dtype=np.float64).reshape(-1)]


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


@contextlib.contextmanager
def capture_fd1(path):
    sys.stdout.flush()
    fd = os.dup(1)
    with open(path, "w") as f:
        os.dup2(f.fileno(), 1)
        try:
            yield
        finally:
            C.CDLL(None).fflush(None)
            os.dup2(fd, 1)
            os.close(fd)


def copy_inputs():
    os.makedirs(os.path.join(INPUTS, "geometry"), exist_ok=True)
    os.makedirs(os.path.join(INPUTS, "config"), exist_ok=True)
    for f in ["low-poly-banbana.stl", "convex-particle.stl", "wall_plane.stl", "wall_box.stl", "table.stl"]:
        shutil.copyfile(os.path.join(REF_GEOM, f), os.path.join(INPUTS, "geometry", f))
    for f in ["example_sim.txt", "wall_test.txt"]:
        shutil.copyfile(os.path.join(REF_CFG, f), os.path.join(INPUTS, "config", f))


def gold_meshes(R):
    out = []
    for name, tris in GI.mesh_inputs(INPUTS):
        for center in (1, 0):
            t = np.ascontiguousarray(tris, dtype=np.float64)
            ot = np.zeros_like(t)
            props = np.zeros(16)
            nv = R.ref_build_mesh(t.reshape(-1), t.shape[0], center, ot.reshape(-1), props)
            out.append(dict(name=name, center=center, T=int(t.shape[0]), input_sha=sha(t), nverts=int(nv),
                            props=hx(props), tris_sha=sha(ot)))
    # STL loader: doubles read by the reference loader
    for f in ["low-poly-banbana.stl", "wall_box.stl"]:
        p = os.path.join(INPUTS, "geometry", f)
        T = R.ref_load_stl(p.encode(), None, 0)
        buf = np.zeros((T, 3, 3))
        R.ref_load_stl(p.encode(), buf.ctypes.data, T)
        out.append(dict(name="load:" + f, T=int(T), tris_sha=sha(buf), first=hx(buf[0])))
    return out


def gold_tri_tri(R):
    A, B = GI.tri_pairs(seed=11, n=4000)
    flags, segs = [], []
    for i in range(len(A)):
        o = np.zeros(6)
        f = R.ref_tri_tri(np.ascontiguousarray(A[i].reshape(-1)), np.ascontiguousarray(B[i].reshape(-1)), o)
        flags.append(int(f))
        segs.append(o if f else np.zeros(6))
    segs = np.asarray(segs)
    return dict(seed=11, n=len(A), input_sha=sha(np.concatenate([A, B])), flags=flags,
                segs_sha=sha(segs), first=[hx(s) for s in segs[:64]])


def ref_pair(R, mA, mB, pose, split=True):
    """Reference per-pair result composed of the reference's own functions."""
    tA, propsA = mA
    tB, _ = mB
    posA, qA, posB, qB = [np.ascontiguousarray(x, dtype=np.float64) for x in pose]
    cap = 100000
    seg = np.zeros((cap, 6))
    chk = C.c_int64(0)
    ns = R.ref_pair_segments(tA.reshape(-1), tA.shape[0], tB.reshape(-1), tB.shape[0], posA, qA, posB, qB,
                             seg.reshape(-1), cap, C.byref(chk))
    seg = seg[:ns].copy()
    mean_edge, bbox = propsA[5], propsA[6]
    tol = mean_edge * 0.1 if mean_edge > 0 else bbox * 1e-2
    if tol <= 0:
        tol = 1e-6
    res = dict(tri_checks=int(chk.value), segments=int(ns), segs_sha=sha(seg), segs_first=[hx(s) for s in seg[:8]],
               comps=[], loops=[])
    if ns == 0:
        return res
    if split:
        sizes = np.zeros(ns, dtype=np.int32)
        order = np.zeros(ns, dtype=np.int32)
        nc = R.ref_components(seg.reshape(-1), ns, tol, sizes, order)
        comps, k = [], 0
        for c in range(nc):
            comps.append(order[k:k + sizes[c]].tolist())
            k += sizes[c]
    else:
        comps = [list(range(ns))]
    res["comps"] = comps
    area_eps = mean_edge * mean_edge * 1e-4 if mean_edge > 0 else 1e-12
    raw = 0
    for ci, comp in enumerate(comps):
        cs = np.ascontiguousarray(seg[comp])
        V = np.zeros((4 * len(comp) + 8, 3))
        nv = C.c_int(0)
        llen = np.zeros(4 * len(comp) + 8, dtype=np.int32)
        lv = np.zeros(8 * len(comp) + 8, dtype=np.int32)
        nl = R.ref_loops(cs.reshape(-1), len(comp), tol, V.reshape(-1), C.byref(nv), llen, lv, lv.size)
        k = 0
        for li in range(nl):
            vids = lv[k:k + llen[li]].tolist()
            k += llen[li]
            entry = dict(comp=ci, raw=raw + li, vids=vids, accepted=False)
            if len(vids) >= 3:
                pts = np.ascontiguousarray(V[vids])
                Sn, Gn, xc, nA, area = np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(1)
                ok = R.ref_sn_xc0(pts.reshape(-1), len(vids), Sn, Gn, xc, nA, area)
                if ok and area[0] >= area_eps:
                    if np.dot(nA, posB - posA) < 0.0:
                        nA = nA * -1.0
                    entry.update(accepted=True, pts=hx(pts), Sn=hx(Sn), Gn=hx(Gn), xc0=hx(xc), n=hx(nA),
                                 area=float(area[0]).hex())
            res["loops"].append(entry)
        raw += nl
    return res


def gold_pairs(R):
    out = []
    for name, mA, mB, pose, split in GI.pair_cases(INPUTS):
        ta, pa, _ = O.build_mesh(mA, True)
        tb, pb, _ = O.build_mesh(mB, True)
        # the reference-side mesh prep is pinned separately; here feed both the same centred tris
        rA = np.zeros_like(ta)
        propsA = np.zeros(16)
        R.ref_build_mesh(np.ascontiguousarray(mA).reshape(-1), mA.shape[0], 1, rA.reshape(-1), propsA)
        rB = np.zeros_like(tb)
        propsB = np.zeros(16)
        R.ref_build_mesh(np.ascontiguousarray(mB).reshape(-1), mB.shape[0], 1, rB.reshape(-1), propsB)
        r = ref_pair(R, (rA, propsA), (rB, propsB), pose, split)
        r.update(name=name, split=int(split), inA_sha=sha(mA), inB_sha=sha(mB), pose=[hx(p) for p in pose])
        out.append(r)
    return out


def run_ref_config(R, cfg_text: str, workdir: str, name: str):
    cfg_path = os.path.join(workdir, name + ".txt")
    with open(cfg_path, "w") as f:
        f.write(cfg_text)
    log = os.path.join(workdir, name + ".log")
    cwd = os.getcwd()
    os.chdir(workdir)
    try:
        argv = (C.c_char_p * 3)(b"zdem_cpu", b"--config", cfg_path.encode())
        with capture_fd1(log):
            rc = R.ref_main(3, argv)
    finally:
        os.chdir(cwd)
    assert rc == 0, rc
    return log


def gold_sims(R):
    out = []
    with tempfile.TemporaryDirectory() as td:
        os.symlink(os.path.join(INPUTS, "geometry"), os.path.join(td, "geometry"))
        for name, text, files in GI.sim_cases(INPUTS):
            for fn, tris in files.items():
                synth.write_stl_ascii(os.path.join(td, fn), tris)
            log = run_ref_config(R, text, td, name)
            lines = [ln.split() for ln in open(log) if ln.startswith("step=")]
            ctr = [{kv.split("=")[0]: int(kv.split("=")[1]) for kv in ln if kv.split("=")[0] in
                    ("step", "contacts", "tri_checks", "segments", "comps", "loops")} for ln in lines]
            prefix = [ln for ln in text.splitlines() if ln.startswith("vtk_prefix")][0].split("=")[1].strip()
            snaps = []
            for c in ctr:
                st = O.parse_state_txt(os.path.join(td, "out", f"{prefix}_{c['step']:06d}.txt"))
                snaps.append(dict(step=c["step"], state=hx(st)))
            out.append(dict(name=name, config=text, files={k: sha(v) for k, v in files.items()},
                            counters=ctr, snapshots=snaps))
    return out


# (case, step) whose VTK and state txt files are kept byte for byte
OUTPUT_FILES = [("wall_drop", 2000), ("cluster_box", 50)]


def gold_outputs():
    """Copies the output files (write_vtk_particles, write_particle_state_txt)
    of a few snapshots written by the reference's unmodified zdem_cpu program
    (oracle/_ref/zdem_cpu; the harness build prints more digits) into output/."""
    import subprocess

    od = os.path.join(HERE, "output")
    os.makedirs(od, exist_ok=True)
    want = dict(OUTPUT_FILES)
    with tempfile.TemporaryDirectory() as td:
        os.symlink(os.path.join(INPUTS, "geometry"), os.path.join(td, "geometry"))
        for name, text, files in GI.sim_cases(INPUTS):
            if name not in want:
                continue
            for fn, tris in files.items():
                synth.write_stl_ascii(os.path.join(td, fn), tris)
            cfg_path = os.path.join(td, name + ".txt")
            with open(cfg_path, "w") as f:
                f.write(text)
            subprocess.run([O.REF_BIN, "--config", cfg_path], cwd=td, check=True, stdout=subprocess.DEVNULL)
            prefix = [ln for ln in text.splitlines() if ln.startswith("vtk_prefix")][0].split("=")[1].strip()
            for ext in ("vtk", "txt"):
                src = os.path.join(td, "out", f"{prefix}_{want[name]:06d}.{ext}")
                shutil.copyfile(src, os.path.join(od, f"{name}_{want[name]:06d}.{ext}"))


def main():
    if "--outputs-only" in sys.argv:
        gold_outputs()
        return
    if not O.ref_available():
        O.build(ref=True)
    R = O.ref()
    copy_inputs()
    data = dict(meshes=gold_meshes(R), tri_tri=gold_tri_tri(R), pairs=gold_pairs(R))
    with open(os.path.join(HERE, "golden_geometry.json"), "w") as f:
        json.dump(data, f, separators=(",", ":"))
    sims = gold_sims(R)
    with open(os.path.join(HERE, "golden_sims.json"), "w") as f:
        json.dump(sims, f, separators=(",", ":"))
    gold_outputs()
    print("pairs:", [(p["name"], p["segments"], len(p["comps"]), len(p["loops"])) for p in data["pairs"]])
    print("sims:", [(s["name"], len(s["snapshots"])) for s in sims])


if __name__ == "__main__":
    main()
