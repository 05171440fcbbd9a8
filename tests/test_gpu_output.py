"""GPU: the per-step output files and the zdx command-line driver.

zdx_sim_write_vtk / zdx_sim_write_state_txt after N device steps are compared
with the files the reference's own zdem_cpu wrote at the same step
(tests/golden/output/): every non-numeric token is identical and the numbers
agree to the trajectory tolerance of the device step (its contact force goes
through cbrt/pow/log, so the state differs from the reference's in the last
bits, which can move the last printed digit).  The zdx CLI is run on the same
configuration and its step lines are checked against the oracle's counters.
"""
import gzip
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import golden_inputs as GI
import oracle as O
from zdx import synth

zdx = pytest.importorskip("zdx")
pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = [("wall_drop", 2000), ("cluster_box", 50)]


@pytest.fixture(scope="module")
def ctx():
    c = zdx.Context(0)
    yield c
    c.close()


def _case_dir(name, td):
    _, text, files = {c[0]: c for c in GI.sim_cases(os.path.join(GOLD, "ref_inputs"))}[name]
    os.symlink(os.path.join(GOLD, "ref_inputs", "geometry"), os.path.join(td, "geometry"))
    for fn, tris in files.items():
        synth.write_stl_ascii(os.path.join(td, fn), tris)
    cfg = os.path.join(td, name + ".txt")
    with open(cfg, "w") as f:
        f.write(text)
    return cfg, text


def _expected(name, step, ext):
    p = os.path.join(GOLD, "output", f"{name}_{step:06d}.{ext}")
    if os.path.exists(p):
        return open(p).read()
    return gzip.open(p + ".gz", "rt").read()


def _num(tok):
    try:
        return float(tok)
    except ValueError:
        return None


def _compare_text(got, want, rtol, atol_frac):
    """Header lines (first token a word, no '=') are identical; the numbers of
    data lines and of 'key = values' lines are close, grouped by section
    (the header or key above them); atol = atol_frac x the section's max |value|."""
    gl, wl = got.splitlines(), want.splitlines()
    assert len(gl) == len(wl)
    sections, cur = {}, "head"
    for a, b in zip(gl, wl):
        ta, tb = a.split(), b.split()
        assert len(ta) == len(tb), (a, b)
        if ta and _num(ta[0]) is None:
            if len(ta) < 2 or ta[1] != "=":
                assert a == b, (a, b)
                if ta[0] != "LOOKUP_TABLE":
                    cur = a
                continue
            cur = ta[0]
        for x, y in zip(ta, tb):
            nx, ny = _num(x), _num(y)
            if nx is None or ny is None:
                assert x == y, (a, b)
            else:
                sections.setdefault(cur, []).append((nx, ny))
    for sec, v in sections.items():
        v = np.asarray(v)
        atol = atol_frac * max(np.max(np.abs(v[:, 1])), 1e-30)
        bad = np.abs(v[:, 0] - v[:, 1]) > atol + rtol * np.abs(v[:, 1])
        assert not bad.any(), (sec, v[bad][:5])


@pytest.mark.parametrize("name,step", CASES)
def test_output_files_match_reference(ctx, name, step):
    with tempfile.TemporaryDirectory() as td:
        cfg, _ = _case_dir(name, td)
        sim = zdx.Sim.from_config(ctx, cfg, td)
        sim.step(step + 1, zdx.MODE_PLAIN)
        _, _, cnt = sim.forces()
        assert cnt.sum() > 0
        vtk, txt = os.path.join(td, "o.vtk"), os.path.join(td, "o.txt")
        sim.write_vtk(vtk)
        sim.write_state_txt(txt)
        _compare_text(open(vtk).read(), _expected(name, step, "vtk"), rtol=2e-5, atol_frac=1e-5)
        # 9 fixed decimals: positions/velocities agree to the trajectory tolerance
        _compare_text(open(txt).read(), _expected(name, step, "txt"), rtol=1e-6, atol_frac=1e-6)


def test_cli_step_lines_match_oracle_counters():
    """zdx --config prints the reference's step line; its counters equal the
    oracle's and the per-step files are written under output_dir."""
    with tempfile.TemporaryDirectory() as td:
        cfg, text = _case_dir("cluster_box", td)
        text = re.sub(r"steps = \d+", "steps = 60", text)
        text = re.sub(r"output_interval = \d+", "output_interval = 20", text)
        with open(cfg, "w") as f:
            f.write(text)
        r = subprocess.run([zdx.CLI_PATH, "--config", cfg], cwd=td, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        lines = [ln for ln in r.stdout.splitlines() if ln.startswith("step=")]
        assert r.stdout.rstrip().endswith("done")
        assert [int(ln.split()[0].split("=")[1]) for ln in lines] == [0, 20, 40]
        osim = O.Sim(cfg, td)
        want = {}
        for s in range(60):
            c = osim.step(1)
            if s % 20 == 0:
                want[s] = c[:5].tolist()
        for ln in lines:
            kv = dict(t.split("=") for t in ln.split())
            got = [int(kv[k]) for k in ("contacts", "tri_checks", "segments", "comps", "loops")]
            assert got == want[int(kv["step"])], ln
        for s in (0, 20, 40):
            for ext in ("vtk", "txt"):
                assert os.path.exists(os.path.join(td, "out", f"cl_{s:06d}.{ext}"))
