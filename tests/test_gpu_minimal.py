"""GPU: the time step with the minimal-energy particle-particle force
(``force_model minimal``): closed form on a box pair, trajectories against the
oracle's step, plain against batched, and the torch model of the same step
(zdx.autograd.step, the forward of the adjoint) against the device."""
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


@pytest.mark.parametrize("depth,off,pa", [(0.2, (0.0, 0.0), (0.0, 0.0)), (0.3, (0.2, 0.2), (0.3, -0.7))])
def test_box_pair_closed_form(ctx, depth, off, pa):
    k = 5.0
    text, files, pA, pB, c = GI.box_pair_case(depth, off, pa, k)
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


# (case, steps, k, rtol): the damping goes through log(e), whose last bit
# differs between the device libm and glibc, and the step amplifies it
MIN_CASES = [("cfg1_contact", 1000, 1e6, 1e-9), ("star_nosplit", 300, 1e5, 1e-9), ("cluster_box", 400, 1e6, 1e-6)]


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


def test_torch_model_matches_device_step(ctx):
    """zdx.autograd.step (torch, the forward the adjoint differentiates) and
    the device step with the minimal force on 8 overlapping hulls, no floor,
    no gravity: same trajectories to round-off after 200 steps."""
    torch = pytest.importorskip("torch")
    from zdx.autograd import Scene, body_inverse_inertia, simulate

    hull = synth.sphere_hull(128, 7) * 0.005
    mid = ctx.register_mesh(hull)
    props = ctx.mesh_props[mid]
    dens, e, k = 2500.0, 0.5, 4e6
    rng = np.random.default_rng(11)
    n = 8
    pos = np.array([[(i & 1) * 0.0094, ((i >> 1) & 1) * 0.0094, (i >> 2) * 0.0094] for i in range(n)])
    pos += rng.uniform(-2e-4, 2e-4, size=pos.shape)
    quat = rng.normal(size=(n, 4))
    quat /= np.linalg.norm(quat, axis=1, keepdims=True)
    vel = rng.normal(scale=0.05, size=(n, 3))
    om = rng.normal(scale=2.0, size=(n, 3))
    init = np.concatenate([pos, quat, vel, om], axis=1)
    mat = np.tile([dens, 1e7, 0.25, 0.5, e], (n, 1))
    sim = zdx.Sim.create(ctx, np.full(n, mid), init, mat, (0.0, 0.0, 0.0), 1e-5, True)
    sim.set_force_model(zdx.FORCE_MINIMAL, k)
    s0 = sim.state()
    steps = 200
    contacts = sum(sim.step(1, zdx.MODE_PLAIN)["contacts"] for _ in range(steps))
    assert contacts > 50
    sg = sim.state()
    sim.close()

    V, I = float(props[0]), np.asarray(props[7:16]).reshape(3, 3) * dens
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64))
    scene = Scene(zctx=ctx, mesh=np.full(n, mid), radius=np.full(n, float(props[4])), mass=t(np.full(n, dens * V)),
                  Ibinv=t(np.broadcast_to(body_inverse_inertia(I), (n, 3, 3)).copy()),
                  req=t(np.full(n, np.cbrt(3.0 * V / (4.0 * np.pi)))), floor_mesh=None, floor_pos=None,
                  floor_top=0.0, gravity=t([0.0, 0.0, 0.0]), dt=1e-5, split=True)
    with torch.no_grad():
        st = simulate(scene, t(k), t(e), t(s0[:, 0:3]), t(s0[:, 3:7]), t(s0[:, 7:10]), t(s0[:, 10:13]),
                      t(s0[:, 13:16]), steps, 50)
    st = np.concatenate([x.numpy() for x in st], axis=1)
    scale = np.maximum(np.abs(sg), 1e-3)
    err = np.abs(st - sg) / scale
    assert np.max(err) <= 1e-8, (np.max(err), np.unravel_index(np.argmax(err), err.shape))
    assert np.max(np.abs(sg[:, 7:10] - s0[:, 7:10])) > 1e-3      # the contacts did change the velocities
