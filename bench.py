#!/usr/bin/env python3
"""Throughput of the Route-B contact step on one MI355X.

Workload: the granular bed of 100k polyhedra (8 templates: convex hulls of
32-512 sphere points and radial-noise stars of 320-1280 triangles, radii
log-uniform 0.5-2 mm, density 2500, young 1e7) in a 0.2 m box bounded by a
12-triangle box wall, gravity -9.81, dt 1e-5 -- see zdx.workloads.granular_bed.

A step is one full time step (transform, cell-binned broad phase, narrow phase,
topology, forces with tangential history, walls, ordered force fold,
integrator).  300 untimed steps settle the bed first, whatever --warmup
says (the lattice-seeded bed reaches its colliding state of about 12k
candidate pairs per step within that), the autotuner picks the narrow-phase
shape on the settled state, W more untimed warmup steps follow, then K timed
steps in batched mode (persistent work queue, hipGraph-captured step), then
the same K steps again in plain mode (one kernel per stage, host-sized
launches) from a checkpoint of the post-warmup state; the two trajectories
must be bitwise equal.  value = contact pairs/s of the batched step, where a
contact pair is a broad-phase candidate pair evaluated by the narrow phase.
The line also carries steps/s, the per-step counters, the steps run again
after a buffer growth and the graph captures inside the timed window, the
narrow phase's roofline and the single-thread reference CPU baseline.

Prints one JSON line.  --gpus N with N > 1 exits: the step runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "zedm-x_amd"))

# MI355X published peaks used for the roofline
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# algorithmic work of one narrow-phase launch (per step), from the counters:
# FP64 ops: 24 per triangle scanned by the cull (3 corner loads, 12 min/max,
#   6 compares, 6 outward roundings), 6 per executed AABB test, 200 per hit
#   (exact AABB re-test 30, two planes 70, tri-tri 100), 100 per segment;
# HBM bytes: 36 per triangle scanned (12 B of corner indices, 24 B of world
#   vertex data: each vertex is shared by ~2 triangles) + 48 per segment written.
OPS_PER_SCANNED, OPS_PER_TEST, OPS_PER_HIT, OPS_PER_SEG = 24, 6, 200, 100
BYTES_PER_SCANNED, BYTES_PER_SEG = 36, 48


def narrow_model(ctr: dict):
    ops = (OPS_PER_SCANNED * ctr["tri_scanned"] + OPS_PER_TEST * ctr["tri_tests"] + OPS_PER_HIT * ctr["tri_hits"] +
           OPS_PER_SEG * ctr["segments"])
    byts = BYTES_PER_SCANNED * ctr["tri_scanned"] + BYTES_PER_SEG * ctr["segments"]
    return ops, byts


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def _candidate_pairs(P, rad):
    """Candidate pairs (a < b) of positions P with radii rad under the
    reference's exact sphere test (src/main.cpp:2054-2059, dpos = P[b] - P[a],
    rejected when dist2 > (ra + rb)^2), from a k-d tree superset."""
    from scipy.spatial import cKDTree

    pr = np.array(sorted(cKDTree(P).query_pairs(2.0 * float(rad.max()) * (1 + 1e-9) + 1e-12)), dtype=np.int64)
    if len(pr) == 0:
        return pr.reshape(0, 2)
    dp = P[pr[:, 1]] - P[pr[:, 0]]
    d2 = dp[:, 0] * dp[:, 0] + dp[:, 1] * dp[:, 1] + dp[:, 2] * dp[:, 2]
    rs = rad[pr[:, 0]] + rad[pr[:, 1]]
    return pr[~(d2 > rs * rs)]


def pick_cpu_sample(wl, state, min_pairs, max_particles):
    """Particles of the densest part of the bed: the k nearest (Chebyshev) to
    the coarse-grid cell holding the most candidate-pair midpoints, k doubled
    from 2000 until the sample holds min_pairs candidate pairs or reaches
    max_particles.  Returns (sorted indices, candidate pairs of the sample)."""
    pos = state[:, :3]
    rad = np.array([wl.scale[i] * _tpl_radius(wl, wl.tpl[i]) for i in range(wl.n)])
    allp = _candidate_pairs(pos, rad)
    if len(allp) == 0:
        raise RuntimeError("the bed has no candidate pairs: nothing for the reference to evaluate")
    mid = 0.5 * (pos[allp[:, 0]] + pos[allp[:, 1]])
    lo, hi = pos.min(axis=0), pos.max(axis=0)
    g = 10
    cell = np.clip(((mid - lo) / np.maximum(hi - lo, 1e-12) * g).astype(np.int64), 0, g - 1)
    key = (cell[:, 0] * g + cell[:, 1]) * g + cell[:, 2]
    best = np.bincount(key, minlength=g ** 3).argmax()
    centre = mid[key == best].mean(axis=0)
    d = np.max(np.abs(pos - centre), axis=1)
    order = np.argsort(d, kind="stable")
    k = 2000
    while True:
        idx = np.sort(order[:min(k, wl.n)])
        pairs = _candidate_pairs(pos[idx], rad[idx])
        if len(pairs) >= min_pairs or k >= max_particles or k >= wl.n:
            return idx, pairs
        k *= 2


def cpu_baseline(wl, state, min_pairs, max_particles, timeout_s):
    """Times the reference program (zdem_cpu, compiled from the reference's
    own source into oracle/_ref/) for one step on the densest part of the
    settled bed, single-threaded.  value = the sample's candidate pairs
    divided by the step's contact time as the program reports it (tri_ms +
    rebuild_ms: narrow phase, topology and forces; its O(N^2) pair loop, the
    walls and its output are not counted).  The program's VTK / state
    output is pointed at a directory that does not exist, so it costs
    nothing (the reference reports the failed open and goes on).  Returns
    None when the program is not built; raises when the sample has no
    candidate pair or the program prints no step line."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from zdx import workloads as W

    if not os.path.exists(O.REF_BIN):
        return None
    idx, pairs = pick_cpu_sample(wl, state, min_pairs, max_particles)
    n_pairs = len(pairs)
    if n_pairs == 0:
        raise RuntimeError(f"cpu baseline sample of {len(idx)} particles has no candidate pair")
    tmp = tempfile.mkdtemp(prefix="zdx_cpu_")
    try:
        cfg = W.write_reference_config(wl, idx, tmp, steps=1, output_interval=1, state=state,
                                       prefix="sample", vtk_prefix=os.path.join("no_such_dir", "sample"))
        env = dict(os.environ)
        env["OMP_NUM_THREADS"] = "1"
        cmd = [O.REF_BIN, "--config", cfg]
        if shutil.which("taskset"):
            cmd = ["taskset", "-c", "0"] + cmd
        t0 = time.time()
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s, env=env, cwd=tmp)
        wall = time.time() - t0
        m = re.search(r"step=0 contacts=(\d+) step_ms=(\d+) tri_ms=(\d+) rebuild_ms=(\d+) output_ms=(\d+) "
                      r"tri_checks=(\d+)", out.stdout)
        if not m:
            raise RuntimeError("reference run produced no step line: " + out.stderr[-400:])
        contact_ms = int(m.group(3)) + int(m.group(4))
        if contact_ms <= 0:
            raise RuntimeError(f"reference contact time of {n_pairs} pairs below its 1 ms resolution")
        return dict(value=n_pairs / contact_ms * 1000.0, unit="contact pairs/s", cores=1, kind="reference",
                    sample=f"{len(idx)} particles of the densest part of the settled bed, {n_pairs} candidate "
                           f"pairs, 1 step of zdem_cpu (reference src/main.cpp, g++ -O2, one thread pinned to "
                           f"core 0): narrow phase + topology + forces {contact_ms} ms (its tri_ms + rebuild_ms), "
                           f"tri_checks {m.group(6)}, active contacts {m.group(1)}, step {m.group(2)} ms incl. "
                           f"its O(N^2) pair loop, process wall {wall:.1f} s incl. set-up",
                    cpu=cpu_model(), ms_per_pair=contact_ms / n_pairs)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def uneven_block(ctx, reps: int, cpu_sample: int, log):
    """The uneven-pair batch (zdx.workloads.uneven_pairs: 4096 independent
    pairs, icospheres of 80-5120 triangles and concave stars of 1280) in one
    zdx_contact_pairs call per repetition, batched and plain.  Reports pairs/s
    on the device time of the call (transform + narrow phase + topology, HIP
    events) and end to end (host packing and PCIe transfers included), and a
    single-thread CPU baseline: the oracle (the FP64 restatement of the
    reference's per-pair body, src/main.cpp:2061-2219) on a seeded sample of
    pairs, extrapolated to the batch by the reference's tri_checks counter."""
    import zdx
    from zdx import workloads as W

    meshes, names, pairs = W.uneven_pairs(4096)
    ids = [ctx.register_mesh(m) for m in meshes]
    P = dict(mesh_a=[ids[p[0]] for p in pairs], mesh_b=[ids[p[1]] for p in pairs],
             pos_a=np.array([p[2] for p in pairs]), quat_a=np.array([p[3] for p in pairs]),
             pos_b=np.array([p[4] for p in pairs]), quat_b=np.array([p[5] for p in pairs]))
    out = {"workload": "4096 independent pairs: 2048 icosphere pairs (levels 1-4 per body, 80-5120 triangles), "
                       "2048 concave star pairs (1280 triangles, radius x (1 + 0.3 noise)), random quaternions, "
                       "overlap 1-10 % of the smaller radius along the line of centres; split contacts, FP64 "
                       "per-loop Sn, Gn, n, xc0, area, Fn, force and torques"}
    ref = None
    for mode, name in ((zdx.MODE_BATCHED, "batched"), (zdx.MODE_PLAIN, "plain")):
        ctx.contact_pairs(**P, split=True, mode=mode, k=1.0)      # untimed: buffers grow here
        dev, wall = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            r = ctx.contact_pairs(**P, split=True, mode=mode, k=1.0)
            wall.append(time.perf_counter() - t0)
            t = ctx.contact_times()
            assert t["attempts"] == 1
            dev.append((t["transform"] + t["narrow"] + t["topology"], t["narrow"], t["topology"]))
        d = np.median(np.array(dev), axis=0)
        if ref is None:
            ref = r
        else:
            out["plain_equals_batched"] = bool(all(np.array_equal(ref[k], r[k]) for k in
                                                   ("tri_checks", "n_segments", "loop_offset", "loops")))
        out[name] = dict(pairs_per_s=4096 / (d[0] * 1e-3), device_ms=d[0], narrow_ms=d[1], topology_ms=d[2],
                         pairs_per_s_end_to_end=4096 / float(np.median(wall)),
                         end_to_end_ms=float(np.median(wall)) * 1e3)
        log(f"uneven {name}: device {d[0]:.2f} ms (narrow {d[1]:.2f}, topology {d[2]:.2f}), "
            f"end to end {np.median(wall) * 1e3:.1f} ms")
    out["batched_over_plain"] = out["plain"]["device_ms"] / out["batched"]["device_ms"]
    tc = ref["tri_checks"].astype(np.float64)
    out["tri_checks"] = float(tc.sum())
    out["segments"] = int(ref["n_segments"].sum())
    out["loops"] = int(len(ref["loops"]))
    if cpu_sample > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        built = [O.build_mesh(m) for m in meshes]
        sel = np.random.default_rng(11).choice(4096, cpu_sample, replace=False)
        t0 = time.perf_counter()
        for p in sel:
            a, b, pa, qa, pb, qb = pairs[p]
            O.pair_contact(built[a][0], built[b][0], pa, qa, pb, qb, built[a][1][5], built[a][1][6], split=True,
                           k=1.0)
        ts = time.perf_counter() - t0
        est = ts * tc.sum() / tc[sel].sum()
        out["cpu_baseline"] = dict(value=4096 / est, unit="contact pairs/s", cores=1, kind="port",
                                   sample=f"{cpu_sample} seeded pairs of the batch through the oracle (FP64 "
                                          f"restatement of src/main.cpp:2061-2219, g++ -O2, one thread) in "
                                          f"{ts:.2f} s, extrapolated to the 4096 pairs by tri_checks "
                                          f"({tc[sel].sum():.3g} of {tc.sum():.3g}): {est:.1f} s per batch")
        log(f"uneven cpu: {cpu_sample} pairs {ts:.2f} s, batch estimate {est:.1f} s")
    return out


def packing_block(ctx, steps: int, log):
    """The 1M-particle concave packing (zdx.workloads.concave_packing: 16
    star templates of 320-2560 triangles, radii 1-4 mm, 0.5 m box, dt 5e-6)
    on the same GPU: 2 untimed batched steps (buffer growth, graph capture),
    then `steps` timed batched steps."""
    import zdx
    from zdx import workloads as W

    t0 = time.time()
    wl = W.concave_packing()
    sim, n_meshes = W.build_sim(ctx, wl)
    setup = time.time() - t0
    sim.step(2, zdx.MODE_BATCHED)
    sim.stage_times()
    t, c = run_steps(sim, steps, zdx.MODE_BATCHED)
    st = sim.stage_times()
    sim.close()
    log(f"packing 1M: {steps} steps {t:.3f} s, pairs/step {c['pairs'] / steps:.0f}")
    return {"workload": f"1M concave stars (16 templates of 320-2560 triangles, radii log-uniform 1-4 mm on 256 "
                        f"levels), 0.5 m box, 12-triangle box wall, gravity -9.81, dt 5e-6; hexagonal layers "
                        f"with bounding spheres overlapping up to 12 %; timed after 2 untimed steps",
            "particles": wl.n, "meshes": n_meshes, "setup_s": setup, "steps": steps,
            "contact_pairs_per_s": c["pairs"] / t, "steps_per_s": steps / t, "ms_per_step": t / steps * 1e3,
            "contact_pairs_per_step": c["pairs"] / steps, "active_contacts_per_step": c["contacts"] / steps,
            "segments_per_step": c["segments"] / steps, "timed_retries": st["retries"],
            "stage_ms": {k: st[k] for k in zdx.Sim.STAGES}}


_TPL_RAD = {}


def _tpl_radius(wl, t):
    from zdx import workloads as W

    if t not in _TPL_RAD:
        _TPL_RAD[t] = W._centred_radius(wl.templates[t])
    return _TPL_RAD[t]


COPY_BYTES = 256 << 20   # plain copy used to scale the byte counters


def copy_calibration():
    """A device-to-device copy of COPY_BYTES by an elementwise copy kernel
    (torch), for scaling FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch

    n = COPY_BYTES // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()


def _pmc_by_kernel(d, counter):
    """Per-dispatch values of `counter` (summed over its instances) from a
    rocprofv3 --pmc csv directory: {kernel name: [values in dispatch order]}."""
    import csv
    import glob

    out = {}
    for f in glob.glob(os.path.join(d, "*counter_collection.csv")):
        per = {}
        for r in csv.DictReader(open(f)):
            if r.get("Counter_Name") != counter:
                continue
            key = (int(r["Dispatch_Id"]), r["Kernel_Name"])
            per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
        for (did, name), v in sorted(per.items()):
            out.setdefault(name, []).append(v)
    return out


def _trace_by_kernel(d, pred):
    """Durations (us) of the dispatches of each kernel matching pred, in
    dispatch order, from a rocprofv3 --kernel-trace csv directory."""
    import csv
    import glob

    rows = {}
    for f in glob.glob(os.path.join(d, "*kernel_trace.csv")):
        for r in csv.DictReader(open(f)):
            if pred(r["Kernel_Name"]):
                rows.setdefault(r["Kernel_Name"], []).append(
                    (int(r["Dispatch_Id"]), (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0))
    return {k: [t for _, t in sorted(v)] for k, v in rows.items()}


def _short(name):
    return name.replace("zdx::(anonymous namespace)::", "").split("(")[0].replace("void ", "")


# the narrow phase (one launch each per step): cull + sweep with one wave per
# pair, then a workgroup per pair for the pairs it leaves, exact tests,
# compaction, the topology's work lists
NARROW_KERNELS = ("k_narrow_wave", "k_narrow<", "k_narrow_exact", "k_narrow_compact", "k_topo_lists")
NARROW_DESC = ("narrow phase: k_narrow_wave + k_narrow (cull + sweep) + k_narrow_exact + k_narrow_compact + "
               "k_topo_lists")


def _is_narrow(name):
    return any(k in name for k in NARROW_KERNELS)


def summarize_profiles(stats_dir, fetch_dir, write_dir, command="", timed_stats=20, timed_pmc=5, bench_json=None):
    """profiles/ summaries: per-kernel durations of the bench
    (rocprofv3 --kernel-trace --stats) and the HBM traffic of the narrow-phase
    kernel from FETCH_SIZE / WRITE_SIZE (KiB per dispatch), each scaled by what
    the same counter reads for a plain copy of COPY_BYTES.  bench_json: the
    JSON line the profiled bench printed, whose algorithmic bytes per launch
    are recorded beside the counted ones."""
    import csv
    import glob

    prof = os.path.join(ROOT, "profiles")
    os.makedirs(prof, exist_ok=True)
    stats = []
    for f in glob.glob(os.path.join(stats_dir, "*kernel_stats.csv")):
        stats += list(csv.DictReader(open(f)))
    stats.sort(key=lambda r: -float(r["TotalDurationNs"]))
    lines = ["# Kernel durations -- granular bed, 100k polyhedra, batched step (graph capture off)", "",
             f"Command: `{command}`" if command else "", "",
             "| kernel | calls | average us | total % |", "|---|---|---|---|"]
    for r in stats[:30]:
        lines.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1000:.1f} | "
                     f"{float(r['Percentage']):.2f} |")
    tn = {k: v[-timed_stats:] for k, v in _trace_by_kernel(stats_dir, _is_narrow).items()}
    if tn:
        lines += ["", f"Narrow phase over the last {timed_stats} steps (the timed steps), average per launch:"]
        for k, v in sorted(tn.items()):
            lines.append(f"- `{_short(k)}`: {sum(v) / len(v):.1f} us (min {min(v):.1f}, max {max(v):.1f})")
        lines.append(f"- sum of these kernels: {sum(sum(v) / len(v) for v in tn.values()):.1f} us per step")
    open(os.path.join(prof, "kernels_granular_bed_100k.md"), "w").write("\n".join(lines) + "\n")
    fetch = _pmc_by_kernel(fetch_dir, "FETCH_SIZE")
    write = _pmc_by_kernel(write_dir, "WRITE_SIZE")

    def pick(d, pred):
        vals = [v for k, vs in d.items() if pred(k) for v in vs]
        return vals

    is_copy = lambda k: ("copy" in k.lower() or "elementwise" in k.lower()) and "zdx" not in k
    cf, cw = pick(fetch, is_copy), pick(write, is_copy)

    def per_step(d):   # mean per launch of each narrow kernel over the timed steps, summed
        vals = {k: v[-timed_pmc:] for k, v in d.items() if _is_narrow(k)}
        return sum(sum(v) / len(v) for v in vals.values()) if vals else 0.0, vals

    nf_kib, nf = per_step(fetch)
    nw_kib, nwr = per_step(write)
    res = {}
    if cf and cw and nf and nwr:
        kib = COPY_BYTES / 1024.0
        f_scale = kib / max(cf)     # bytes the copy moved / what the counter says
        w_scale = kib / max(cw)
        fetch_b = nf_kib * f_scale * 1024.0
        write_b = nw_kib * w_scale * 1024.0
        res = dict(kernel=NARROW_DESC,
                   dispatches=timed_pmc,
                   fetch_bytes_per_launch=fetch_b,
                   write_bytes_per_launch=write_b, bytes_per_launch=fetch_b + write_b,
                   fetch_scale=f_scale, write_scale=w_scale, copy_bytes=COPY_BYTES,
                   copy_fetch_kib=max(cf), copy_write_kib=max(cw))
        if bench_json:
            line = [x for x in open(bench_json).read().splitlines() if x.startswith("{")][-1]
            rf = json.loads(line)["roofline"]
            res.update(algorithmic_bytes_per_launch=rf["bytes_per_launch"], profiled_avg_launch_ms=rf["avg_launch_ms"])
        json.dump(res, open(os.path.join(prof, "narrow_traffic.json"), "w"), indent=1)
        md = ["# HBM traffic of the narrow phase -- granular bed, 100k polyhedra", "",
              f"Commands: `{command}` with `--pmc FETCH_SIZE` and `--pmc WRITE_SIZE` (separate runs)", "",
              f"- kernels: {', '.join(_short(k) for k in sorted(nf))}, one launch each per step; the last "
              f"{timed_pmc} steps (the timed steps)",
              *[f"  - `{_short(k)}`: FETCH {sum(v) / len(v):.0f} KiB, WRITE {sum(nwr.get(k, [0])) / max(1, len(nwr.get(k, [0]))):.0f} KiB"
                for k, v in sorted(nf.items())],
              f"- FETCH_SIZE per step: {nf_kib:.0f} KiB x scale {f_scale:.3f} = {fetch_b / 1e6:.1f} MB",
              f"- WRITE_SIZE per step: {nw_kib:.0f} KiB x scale {w_scale:.3f} = {write_b / 1e6:.1f} MB",
              f"- scale = bytes of a plain {COPY_BYTES >> 20} MiB device copy / what the counter reads for it "
              f"(fetch {max(cf):.0f} KiB, write {max(cw):.0f} KiB)",
              f"- traffic per launch: {(fetch_b + write_b) / 1e6:.1f} MB",
              *([f"- algorithmic bytes per launch of the profiled run (its counters): "
                 f"{res['algorithmic_bytes_per_launch'] / 1e6:.1f} MB"] if "algorithmic_bytes_per_launch" in res
                else [])]
        open(os.path.join(prof, "narrow_traffic_granular_bed_100k.md"), "w").write("\n".join(md) + "\n")
    print(json.dumps(res))


VALU_COUNTERS = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_WAVE_CYCLES",
                 "SQ_BUSY_CYCLES", "SQ_WAIT_INST_ANY", "SQ_INSTS_VALU_FLOPS_FP64", "GRBM_GUI_ACTIVE")


def summarize_valu(pmc_dir, stats_dir, command="", timed=5, n_cu=256):
    """profiles/ summary of the VALU and occupancy counters (one rocprofv3
    --pmc run with VALU_COUNTERS) per kernel, averaged over the last `timed`
    dispatches of each: the rocprofv3 derived metrics computed from the raw
    counters -- VALUBusy = 100 sum(SQ_ACTIVE_INST_VALU) / CU_NUM / G,
    VALUUtilization = 100 sum(SQ_THREAD_CYCLES_VALU) / (sum(SQ_ACTIVE_INST_VALU)
    64), OccupancyPercent = 400 sum(SQ_WAVE_CYCLES) / G / CU_NUM / 32 -- with
    G the kernel's GPU cycles: GRBM_GUI_ACTIVE as rocprofv3 reports it per
    dispatch on MI355X is the sum over the 8 XCDs (Num_Xcc of the agent
    information; it equals 8 x the kernel's duration x the 2.4 GHz clock), so G
    = GRBM_GUI_ACTIVE / Num_Xcc -- and the FP64 VALU rate from
    SQ_INSTS_VALU_FLOPS_FP64 over the kernel's average duration in the kernel
    trace of stats_dir."""
    import csv
    import glob

    n_xcc = 8
    for f in glob.glob(os.path.join(pmc_dir, "*agent_info.csv")):
        for r in csv.DictReader(open(f)):
            if int(r.get("Num_Xcc", 0) or 0) > 0:
                n_xcc = int(r["Num_Xcc"])
    per = {}
    for f in glob.glob(os.path.join(pmc_dir, "*counter_collection.csv")):
        for r in csv.DictReader(open(f)):
            key = (int(r["Dispatch_Id"]), r["Kernel_Name"])
            d = per.setdefault(key, {})
            c, v = r["Counter_Name"], float(r["Counter_Value"])
            d[c] = max(d.get(c, 0.0), v) if c.startswith("GRBM") else d.get(c, 0.0) + v
    by = {}
    for (did, name), d in sorted(per.items()):
        by.setdefault(name, []).append(d)
    dur = {k: v[-timed:] for k, v in _trace_by_kernel(stats_dir, lambda n: True).items()}
    rows = []
    for name, ds in by.items():
        ds = ds[-timed:]
        m = {c: sum(d.get(c, 0.0) for d in ds) / len(ds) for c in VALU_COUNTERS}
        g = max(m["GRBM_GUI_ACTIVE"] / n_xcc, 1.0)
        t_us = sum(dur.get(name, [0.0])) / max(1, len(dur.get(name, [0.0])))
        rows.append(dict(name=name, t_us=t_us, waves=m["SQ_WAVES"], gpu_mhz=g / max(t_us, 1e-9),
                         valu_busy=100.0 * m["SQ_ACTIVE_INST_VALU"] / n_cu / g,
                         valu_util=100.0 * m["SQ_THREAD_CYCLES_VALU"] / max(m["SQ_ACTIVE_INST_VALU"] * 64.0, 1.0),
                         occupancy=400.0 * m["SQ_WAVE_CYCLES"] / g / n_cu / 32.0,
                         wait_frac=m["SQ_WAIT_INST_ANY"] / max(m["SQ_WAVE_CYCLES"], 1.0),
                         fp64=m["SQ_INSTS_VALU_FLOPS_FP64"],
                         fp64_tflops=m["SQ_INSTS_VALU_FLOPS_FP64"] / (t_us * 1e-6) / 1e12 if t_us > 0 else 0.0))
    rows.sort(key=lambda r: -r["t_us"])
    lines = ["# VALU and occupancy -- granular bed, 100k polyhedra, batched step (graph capture off)", "",
             f"Command: `{command}`" if command else "", "",
             f"Counters: {' '.join(VALU_COUNTERS)} (one --pmc pass), per kernel averaged over its last {timed} "
             f"dispatches (the timed steps); durations from the kernel trace of the same bench.  VALUBusy, "
             f"VALUUtilization and OccupancyPercent are rocprofv3's derived metrics computed from these raw counters "
             f"(CU_NUM = {n_cu}) with the kernel's GPU cycles G = GRBM_GUI_ACTIVE / {n_xcc} (rocprofv3 sums "
             f"GRBM_GUI_ACTIVE over the {n_xcc} XCDs; the column MHz = G / duration checks it against the 2.4 GHz "
             f"clock); wait = SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES; FP64 = SQ_INSTS_VALU_FLOPS_FP64 per "
             f"dispatch and its rate over the kernel's duration (peak 78.6 TFLOP/s).", "",
             "| kernel | us | MHz | waves | VALUBusy % | VALUUtil % | Occupancy % | wait | FP64 GFLOP | FP64 TFLOP/s |",
             "|---|---|---|---|---|---|---|---|---|---|"]
    for r in rows[:16]:
        lines.append(f"| `{_short(r['name'])}` | {r['t_us']:.1f} | {r['gpu_mhz']:.0f} | {r['waves']:.0f} | "
                     f"{r['valu_busy']:.1f} | {r['valu_util']:.1f} | {r['occupancy']:.1f} | {r['wait_frac']:.2f} | "
                     f"{r['fp64'] / 1e9:.3f} | {r['fp64_tflops']:.2f} |")
    open(os.path.join(ROOT, "profiles", "valu_occupancy_granular_bed_100k.md"), "w").write("\n".join(lines) + "\n")
    print(json.dumps(rows[:16]))


def run_steps(sim, k, mode):
    import zdx

    tot = {c: 0 for c in zdx.COUNTERS}
    t0 = time.perf_counter()
    for _ in range(k):
        c = sim.step(1, mode)
        for key in tot:
            tot[key] += c[key]
    return time.perf_counter() - t0, tot


# untimed steps that settle the bed before anything is timed or tuned, whatever
# --warmup asks: the lattice-seeded bed starts with no overlapping bounding
# spheres and reaches its colliding state (about 12k candidate pairs and 1.5k
# active contacts per step) within about 250 steps
SETTLE_STEPS = 300


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed steps after the settle and the autotune (at least 3 run)")
    ap.add_argument("--settle", type=int, default=SETTLE_STEPS, help=argparse.SUPPRESS)
    ap.add_argument("--particles", type=int, default=100_000)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tiles", default="", help="narrow-phase shape 'threads*100+tests,groups_per_cu' (no autotune)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-uneven", action="store_true", help="skip the 4096-pair uneven batch block")
    ap.add_argument("--uneven-reps", type=int, default=5)
    ap.add_argument("--uneven-cpu-pairs", type=int, default=256)
    ap.add_argument("--no-packing", action="store_true", help="skip the 1M-particle packing block")
    ap.add_argument("--packing-steps", type=int, default=5)
    ap.add_argument("--cpu-pairs", type=int, default=4000, help="candidate pairs wanted in the CPU sample")
    ap.add_argument("--cpu-max-particles", type=int, default=32000)
    ap.add_argument("--cpu-timeout", type=float, default=240.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain device copy of known size (for scaling hardware byte counters)")
    ap.add_argument("--summarize-profiles", nargs=3, metavar=("STATS_DIR", "FETCH_DIR", "WRITE_DIR"),
                    help="write profiles/ summaries from rocprofv3 --stats / --pmc FETCH_SIZE / WRITE_SIZE "
                         "output directories of this bench, then exit")
    ap.add_argument("--profile-command", default="", help="command line recorded in the summaries")
    ap.add_argument("--profile-bench-json", default=None, help="JSON line printed by the profiled bench run")
    ap.add_argument("--summarize-valu", nargs=2, metavar=("PMC_DIR", "STATS_DIR"),
                    help="write the profiles/ VALU and occupancy summary from a rocprofv3 --pmc run with "
                         "VALU_COUNTERS and the --kernel-trace --stats run of the same bench, then exit")
    args = ap.parse_args()
    if args.summarize_valu:
        summarize_valu(*args.summarize_valu, command=args.profile_command)
        return
    if args.summarize_profiles:
        summarize_profiles(*args.summarize_profiles, command=args.profile_command, bench_json=args.profile_bench_json)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the step runs on a single MI355X: there is no collective or multi-GPU "
                                   "path; run with --gpus 1"}))
        sys.exit(2)

    import zdx
    from zdx import workloads as W

    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)
    if args.copy_calibration:
        copy_calibration()
    t0 = time.time()
    wl = W.granular_bed(n=args.particles)
    ctx = zdx.Context(0)
    sim, n_meshes = W.build_sim(ctx, wl)
    log(f"setup {time.time() - t0:.1f}s: {wl.n} particles, {n_meshes} meshes")

    # settle (untimed, fixed length), then the autotune on the settled
    # state, then the warmup (the first step after the autotune captures the
    # graphs again)
    tiles = None
    if args.tiles:
        tiles = [int(x) for x in args.tiles.split(",")]
        sim.set_tiles(tiles[0], tiles[1])
    sim.step(args.settle, zdx.MODE_BATCHED)
    st_settle = sim.stage_times()
    log(f"settle {args.settle} steps done {time.time() - t0:.1f}s, retries {st_settle['retries']}, "
        f"graph captures {st_settle['graph_captures']}")
    if not args.tiles and not args.no_autotune:
        tiles = list(sim.autotune())
        log(f"autotune: threads*100+tests={tiles[0]} groups/CU={tiles[1]}")
    warm = max(args.warmup, 3)
    sim.step(warm, zdx.MODE_BATCHED)
    sim.stage_times()
    log(f"warmup {warm} steps done {time.time() - t0:.1f}s")
    tmpdir = tempfile.mkdtemp(prefix="zdx_bench_")
    ck = os.path.join(tmpdir, "warm.ckpt")
    sim.checkpoint(ck)
    state_warm = sim.state()

    # timed: batched
    sim.stage_times()  # reset the per-stage accumulators
    tb, cb = run_steps(sim, args.steps, zdx.MODE_BATCHED)
    st_b = sim.stage_times()
    state_b = sim.state()
    log(f"batched {args.steps} steps {tb:.3f}s pairs/step {cb['pairs'] / args.steps:.0f} stages {st_b}")

    plain = None
    if not args.no_plain:
        sim.restore(ck)
        sim.step(1, zdx.MODE_PLAIN)     # untimed: the plain sequence's first launches
        sim.restore(ck)
        sim.stage_times()
        tp, cp = run_steps(sim, args.steps, zdx.MODE_PLAIN)
        st_p = sim.stage_times()
        state_p = sim.state()
        plain = dict(seconds=tp, counters=cp, stages=st_p, bitwise_equal=bool(np.array_equal(state_p, state_b)))
        log(f"plain {tp:.3f}s equal={plain['bitwise_equal']} retries {st_p['retries']}")

    # roofline of the dominant kernel of the batched step, from the live events
    stage_ms = {k: st_b[k] for k in zdx.Sim.STAGES}
    dom = max(stage_ms, key=stage_ms.get)
    per_step = {k: (v if k == "seg_max" else v / args.steps) for k, v in st_b["counters"].items()}
    ops, byts = narrow_model(per_step)
    t_np = st_b["narrow_kernels"] * 1e-3   # the narrow-phase kernels alone (device clock stamps)
    ach_tflops = ops / t_np / 1e12 if t_np > 0 else 0.0
    ach_gbs = byts / t_np / 1e9 if t_np > 0 else 0.0
    f_flops, f_hbm = ach_tflops / FP64_PEAK_TFLOPS, ach_gbs / HBM_PEAK_GBS
    if f_hbm >= f_flops:
        roof = dict(bound="hbm", achieved=ach_gbs, peak=HBM_PEAK_GBS, unit="GB/s", frac=f_hbm, traffic=None)
    else:
        roof = dict(bound="mfma", achieved=ach_tflops, peak=FP64_PEAK_TFLOPS, unit="TFLOP/s", frac=f_flops,
                    traffic=None)
    roof.update(kernel=NARROW_DESC,
                compute="FP64 VALU (no matrix-core work in this path)",
                avg_launch_ms=st_b["narrow_kernels"], fp64_ops_per_launch=ops, bytes_per_launch=byts,
                frac_fp64=f_flops, frac_hbm=f_hbm, dominant_stage=dom, stage_ms=stage_ms)
    traffic_file = os.path.join(ROOT, "profiles", "narrow_traffic.json")
    if os.path.exists(traffic_file):
        try:
            tr = json.load(open(traffic_file))
            roof["traffic"] = tr.get("bytes_per_launch")
            roof["traffic_source"] = (f"profiles/narrow_traffic.json: rocprofv3 FETCH_SIZE + WRITE_SIZE of the "
                                      f"narrow-phase kernels on this workload after the same settle; there "
                                      f"{tr.get('algorithmic_bytes_per_launch', 0) / 1e6:.0f} MB algorithmic")
        except (OSError, ValueError):
            pass

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        try:
            cpu = cpu_baseline(wl, state_warm, args.cpu_pairs, args.cpu_max_particles, args.cpu_timeout)
        except (RuntimeError, subprocess.TimeoutExpired) as e:
            log(f"cpu baseline FAILED: {e}")
            cpu = dict(value=None, unit="contact pairs/s", cores=1, kind="reference", error=str(e)[:300],
                       sample="failed")
    if cpu is None:
        cpu = dict(value=None, unit="contact pairs/s", cores=1, kind="reference",
                   sample="not run (the reference program is not built)")

    uneven = None
    if not args.no_uneven:
        uneven = uneven_block(ctx, args.uneven_reps, 0 if args.no_cpu_baseline else args.uneven_cpu_pairs, log)
    packing = None
    if not args.no_packing:
        packing = packing_block(ctx, args.packing_steps, log)

    pairs_b = cb["pairs"]
    res = {
        "metric": "contact pairs/sec + steps/sec @100k polyhedra, batched vs plain GPU",
        "value": pairs_b / tb,
        "unit": "contact pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tb / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"granular bed, 100k polyhedra (8 templates, 60-1280 triangles), 0.2 m box, "
                               f"12-triangle box wall, gravity -9.81, dt 1e-5, agitated non-overlapping seeding; "
                               f"timed after {args.settle} untimed settle steps, the autotune on the settled "
                               f"state and {warm} warmup steps",
                   "particles": wl.n, "meshes": n_meshes, "mode": "batched (persistent queue + hipGraph)",
                   "tiles": tiles, "settle_steps": args.settle},
        "steps_per_s": args.steps / tb,
        "contact_pairs_per_step": pairs_b / args.steps,
        "active_contacts_per_step": cb["contacts"] / args.steps,
        "counters_per_step": per_step,
        "timed_retries": st_b["retries"],
        "timed_graph_captures": st_b["graph_captures"],
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if plain:
        res["plain"] = {"contact_pairs_per_s": plain["counters"]["pairs"] / plain["seconds"],
                        "steps_per_s": args.steps / plain["seconds"],
                        "ms_per_step": plain["seconds"] / args.steps * 1e3,
                        "stage_ms": {k: plain["stages"][k] for k in zdx.Sim.STAGES},
                        "timed_retries": plain["stages"]["retries"],
                        "bitwise_equal_to_batched": plain["bitwise_equal"]}
        res["batched_over_plain"] = (pairs_b / tb) / (plain["counters"]["pairs"] / plain["seconds"])
    if uneven:
        res["uneven_pairs"] = uneven
    if packing:
        res["packing_1m"] = packing
    shutil.rmtree(tmpdir, ignore_errors=True)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
