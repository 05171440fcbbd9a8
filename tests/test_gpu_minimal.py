"""GPU: the time step with the minimal-energy particle-particle force
(``force_model minimal``): closed form on a box pair (at rest and moving:
the force does not depend on velocity), trajectories against the oracle's
step, plain against batched.  The differentiable model of the same step is
compared with the device in test_gpu_adjoint."""
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


def _write_case(td, text, files, ref_inputs=None):
    if ref_inputs is not None:
        os.symlink(os.path.join(ref_inputs, "geometry"), os.path.join(td, "geometry"))
    for fn, tris in files.items():
        synth.write_stl_ascii(os.path.join(td, fn), tris)
    cfg = os.path.join(td, "case.txt")
    with open(cfg, "w") as f:
        f.write(text)
    return cfg


@pytest.mark.parametrize("depth,off,pa,moving", [(0.2, (0.0, 0.0), (0.0, 0.0), False),
                                                 (0.3, (0.2, 0.2), (0.3, -0.7), False),
                                                 (0.3, (0.2, 0.2), (0.3, -0.7), True)])
def test_box_pair_closed_form(ctx, depth, off, pa, moving):
    """F = k/2 A n on B, -F on A, torques about each centre at the overlap
    rectangle's centre; with the bodies approaching at 3 m/s and spinning the
    force is the same (the minimal force has no velocity term)."""
    k = 5.0
    text, files, pA, pB, c = GI.box_pair_case(depth, off, pa, k)
    if moving:
        text = text.replace("vel = 0 0 0", "vel = 1.5 0 0", 1).replace("vel = 0 0 0", "vel = -1.5 0 0", 1)
        text = text.replace("omega = 0 0 0", "omega = 0.3 -0.2 0.1", 1)
        assert "vel = 1.5 0 0" in text and "vel = -1.5 0 0" in text
    with tempfile.TemporaryDirectory() as td:
        sim = zdx.Sim.from_config(ctx, _write_case(td, text, files), td)
        ctr = sim.step(1, zdx.MODE_PLAIN)
        assert ctr["contacts"] == 1
        F, T, cnt = sim.forces()
    Fb = 0.5 * k * (1.2 * 0.8) * np.array([1.0, 0.0, 0.0])
    assert np.allclose(F[1], Fb, rtol=1e-12, atol=1e-12) and np.allclose(F[0], -Fb, rtol=1e-12, atol=1e-12)
    assert np.allclose(T[0], np.cross(c - pA, -Fb), atol=1e-12)
    assert np.allclose(T[1], np.cross(c - pB, Fb), atol=1e-12)
    assert list(cnt) == [1, 1]


# (case, steps, k, rtol): with no walls the minimal force and the integrator
# use only + - * / sqrt, correctly rounded on both sides, so the trajectories
# agree to round-off; cluster_box has walls, whose force goes through
# cbrt/pow/log (last bits differ between the device libm and glibc)
MIN_CASES = [("cfg1_contact", 1000, 1e6, 1e-12), ("star_nosplit", 300, 1e5, 1e-12), ("cluster_box", 400, 1e6, 1e-6)]


def _minimal_text(ref_inputs, name, k):
    text, files = [(t, fs) for n, t, fs in GI.sim_cases(ref_inputs) if n == name][0]
    return text + f"force_model = minimal\nminimal_k = {k:.17g}\n", files


@pytest.mark.parametrize("name,steps,k,rtol", MIN_CASES)
def test_sim_minimal_matches_oracle(ctx, ref_inputs, name, steps, k, rtol):
    text, files = _minimal_text(ref_inputs, name, k)
    with tempfile.TemporaryDirectory() as td:
        cfg = _write_case(td, text, files, ref_inputs)
        osim = O.Sim(cfg, td)
        gsim = zdx.Sim.from_config(ctx, cfg, td)
        done, equal_steps, contacts = 0, 0, 0
        while done < steps:
            m = min(50, steps - done)
            for _ in range(m):
                oc = osim.step(1)
                gc = gsim.step(1, zdx.MODE_PLAIN)
                contacts += gc["contacts"]
                if [gc["contacts"], gc["segments"], gc["loops"]] == [oc[0], oc[2], oc[4]]:
                    equal_steps += 1
            done += m
            so, sg = osim.state(), gsim.state()
            err = np.max(np.abs(so - sg) / np.maximum(np.abs(so), 1e-3))
            assert err <= rtol, (name, done, err)
        assert contacts > 0
        assert equal_steps >= steps - 2, (name, equal_steps, steps)


def test_sim_minimal_plain_equals_batched(ctx, ref_inputs):
    text, files = _minimal_text(ref_inputs, "cluster_box", 1e6)
    with tempfile.TemporaryDirectory() as td:
        cfg = _write_case(td, text, files, ref_inputs)
        a = zdx.Sim.from_config(ctx, cfg, td)
        b = zdx.Sim.from_config(ctx, cfg, td)
        for _ in range(4):
            assert a.step(50, zdx.MODE_PLAIN) == b.step(50, zdx.MODE_BATCHED)
            assert np.array_equal(a.state(), b.state())
        # switching the model back drops the captured graphs: both follow
        a.set_force_model(zdx.FORCE_ZDEM)
        b.set_force_model(zdx.FORCE_ZDEM)
        assert a.step(20, zdx.MODE_PLAIN) == b.step(20, zdx.MODE_BATCHED)
        assert np.array_equal(a.state(), b.state())
        with pytest.raises(zdx.ZdxError):
            a.set_force_model(zdx.FORCE_MINIMAL, -1.0)
