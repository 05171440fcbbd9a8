"""The reference's per-step output files (write_vtk_particles
src/main.cpp:1710-1840, write_particle_state_txt 1842-1870) written by the
product's writers (zedm-x_amd/csrc/zdx_output.cpp, host build) from the
oracle's state, compared byte for byte with files the reference's own
zdem_cpu wrote for the same configuration and step (tests/golden/output/)."""
import ctypes as C
import gzip
import os
import subprocess

import numpy as np
import pytest

import golden_inputs as GI
import oracle as O
from zdx import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "zedm-x_amd", "csrc")
GOLD = os.path.join(ROOT, "tests", "golden")
CASES = [("wall_drop", 2000), ("cluster_box", 50)]


@pytest.fixture(scope="module")
def W(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("outtest") / "liboutputtest.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", out,
                    os.path.join(ROOT, "tests", "csrc", "outtest.cpp"), os.path.join(CSRC, "zdx_output.cpp"),
                    os.path.join(CSRC, "zdx_host.cpp")], check=True)
    L = C.CDLL(out)
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    L.ot_write.argtypes = [C.c_char_p, C.c_char_p, C.c_int, dp, dp, dp,
                           np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"), C.c_char_p, C.c_char_p]
    return L


def _expected(name, step, ext):
    p = os.path.join(GOLD, "output", f"{name}_{step:06d}.{ext}")
    if os.path.exists(p):
        return open(p, "rb").read()
    return gzip.open(p + ".gz", "rb").read()


@pytest.mark.parametrize("name,step", CASES)
def test_output_files_match_reference_bytes(W, tmp_path, name, step):
    case = {c[0]: c for c in GI.sim_cases(os.path.join(GOLD, "ref_inputs"))}[name]
    _, text, files = case
    os.symlink(os.path.join(GOLD, "ref_inputs", "geometry"), tmp_path / "geometry")
    for fn, tris in files.items():
        synth.write_stl_ascii(str(tmp_path / fn), tris)
    cfg = tmp_path / "case.txt"
    cfg.write_text(text)
    sim = O.Sim(str(cfg), str(tmp_path))
    sim.step(step + 1)                     # files of step s are written after step s (steps count from 0)
    state = np.ascontiguousarray(sim.state())
    F, T, cnt = sim.forces()
    vtk, txt = tmp_path / "o.vtk", tmp_path / "o.txt"
    rc = W.ot_write(str(cfg).encode(), str(tmp_path).encode(), sim.n, state, np.ascontiguousarray(F),
                    np.ascontiguousarray(T), np.ascontiguousarray(cnt, dtype=np.int32), str(vtk).encode(),
                    str(txt).encode())
    assert rc == 0
    assert np.any(cnt > 0), "the snapshot should hold contacts"
    assert vtk.read_bytes() == _expected(name, step, "vtk")
    # the state file names the STL as configured (relative to the config's directory)
    assert txt.read_bytes() == _expected(name, step, "txt")
