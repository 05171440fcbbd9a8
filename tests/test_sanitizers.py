"""CPU: the host C++ under AddressSanitizer and UndefinedBehaviorSanitizer
(g++ -fsanitize=address,undefined -fno-sanitize-recover=all): the config
parser and scene set-up, STL loader, build_mesh, output writers and
polyline.py helpers of the library, the per-pair geometry and topology that
the kernels run (compiled for the host), and the oracle's time step, driven
by tests/csrc/sanitize_main.cpp over the reference's own configs and meshes
and the golden simulation configs.  Any out-of-bounds access, use after
free, leak or undefined behaviour fails the run."""
import os
import shutil
import subprocess

import pytest

import golden_inputs as GI
from zdx import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "zedm-x_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path, ref_inputs, golden_sims):
    exe = str(tmp_path / "sanitize_main")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-w", "-o", exe,
           os.path.join(ROOT, "tests", "csrc", "sanitize_main.cpp"), os.path.join(CSRC, "zdx_host.cpp"),
           os.path.join(CSRC, "zdx_output.cpp"), os.path.join(ROOT, "oracle", "zdx_oracle.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    # the reference's configs and meshes, and the golden simulations' configs
    # (paths in the configs are relative to a directory holding geometry/)
    cfgs = []
    d = tmp_path / "sims"
    d.mkdir()
    os.symlink(os.path.join(ref_inputs, "geometry"), str(d / "geometry"))
    for name in ("example_sim.txt", "wall_test.txt"):
        shutil.copy(os.path.join(ref_inputs, "config", name), str(d / name))
        cfgs.append(str(d / name))
    for name, text, files in GI.sim_cases(ref_inputs):
        for fn, tris in files.items():
            synth.write_stl_ascii(str(d / fn), tris)
        (d / f"{name}.txt").write_text(text)
        cfgs.append(str(d / f"{name}.txt"))
    stls = [os.path.join(ref_inputs, "geometry", f) for f in sorted(os.listdir(os.path.join(ref_inputs, "geometry")))]
    out = tmp_path / "out"
    out.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(out)] + cfgs + ["--"] + stls, capture_output=True, text=True, env=env,
                       cwd=str(tmp_path), timeout=600)
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert "sanitize_main ok" in r.stdout
