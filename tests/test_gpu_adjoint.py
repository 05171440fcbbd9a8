"""GPU: sensitivities of the contact pipeline and the differentiable
minimal-energy step, against central finite differences (relative step
1e-6 on parameters, small absolute steps on poses) of the same device
forward."""
import numpy as np
import pytest

import golden_inputs as GI
import oracle as O
from zdx import synth

zdx = pytest.importorskip("zdx")
torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = zdx.Context(0)
    yield c
    c.close()


def _qexp(w):
    th = np.linalg.norm(w)
    if th == 0.0:
        return np.array([1.0, 0.0, 0.0, 0.0])
    return np.concatenate([[np.cos(th / 2)], np.sin(th / 2) * w / th])


def _qmul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


CASES = [("ellipsoid200", "ellipsoid200", [1.95, 0.0, 0.0], [1, 0, 0, 0]),
         ("hull128", "star2a", [1.5, 0.3, -0.1], [0.9, 0.1, 0.3, -0.2]),
         ("star2a", "star2b", [1.45, 0.1, -0.05], [0.8, -0.3, 0.2, 0.4])]


def test_loop_jacobians_match_finite_differences(ctx):
    m = GI.synth_meshes()
    for na, nb, pB, qB in CASES:
        a, b = ctx.register_mesh(m[na]), ctx.register_mesh(m[nb])
        pA, qA = np.zeros(3), np.array([1.0, 0.0, 0.0, 0.0])
        pB, qB = np.asarray(pB, float), np.asarray(qB, float) / np.linalg.norm(qB)
        r0 = ctx.contact_pairs([a], [b], [pA], [qA], [pB], [qB], sensitivities=True, vertices=False)
        n0 = len(r0["loops"])
        assert n0 >= 1
        h = 1e-7
        num = np.zeros((n0, 72))
        for k in range(12):
            out = []
            for sgn in (1, -1):
                xi = np.zeros(12)
                xi[k] = sgn * h
                pa, pb = pA + xi[0:3], pB + xi[6:9]
                qa, qb = _qmul(_qexp(xi[3:6]), qA), _qmul(_qexp(xi[9:12]), qB)
                r = ctx.contact_pairs([a], [b], [pa], [qa], [pb], [qb], vertices=False)
                assert len(r["loops"]) == n0
                out.append(r["loops"])
            d = (out[0] - out[1]) / (2 * h)
            num[:, k:36:12] = d[:, 0:3]
            num[:, 36 + k:72:12] = d[:, 3:6]
        ana = r0["loop_jac"]
        for l in range(n0):
            for part in (slice(0, 36), slice(36, 72)):
                scale = np.max(np.abs(ana[l, part]))
                assert np.max(np.abs(ana[l, part] - num[l, part])) <= 2e-5 * scale, (na, nb, l)


def test_pair_contact_backward_matches_finite_differences(ctx):
    from zdx.autograd import PairContact

    m = GI.synth_meshes()
    ids = [(ctx.register_mesh(m[na]), ctx.register_mesh(m[nb])) for na, nb, _, _ in CASES]
    pA = torch.zeros(3, 3, dtype=torch.float64)
    qA = torch.tensor([[1.0, 0, 0, 0], [0.9, 0.2, -0.1, 0.3], [1.0, 0, 0, 0]], dtype=torch.float64)
    pB = torch.tensor([c[2] for c in CASES], dtype=torch.float64)
    qB = torch.tensor([c[3] for c in CASES], dtype=torch.float64)
    ma = np.array([i[0] for i in ids])
    mb = np.array([i[1] for i in ids])
    rng = np.random.default_rng(4)

    def f(pA, qA, pB, qB, w):
        Sc, xc, lp = PairContact.apply(pA, qA, pB, qB, ctx, ma, mb, True)
        return (Sc * w[0][: len(Sc)]).sum() + (xc * w[1][: len(xc)]).sum()

    w = (torch.from_numpy(rng.normal(size=(64, 3))), torch.from_numpy(rng.normal(size=(64, 3))))
    X = [t.clone().requires_grad_(True) for t in (pA, qA, pB, qB)]
    f(*X, w).backward()
    for idx, name in enumerate(("posA", "quatA", "posB", "quatB")):
        g = X[idx].grad.numpy()
        num = np.zeros_like(g)
        base = [t.detach().clone() for t in (pA, qA, pB, qB)]
        h = 1e-7
        for pos in np.ndindex(*g.shape):
            vals = []
            for sgn in (1, -1):
                Y = [t.clone() for t in base]
                Y[idx][pos] += sgn * h
                with torch.no_grad():
                    vals.append(float(f(*Y, w)))
            num[pos] = (vals[0] - vals[1]) / (2 * h)
        scale = np.max(np.abs(num)) + 1e-12
        assert np.max(np.abs(g - num)) <= 1e-5 * scale, (name, g, num)


def _drop_scene(ctx, n_side=2, seed=4):
    """Hulls of 128 sphere points (r = 5 mm) just above a floor slab."""
    from zdx.autograd import Scene, body_inverse_inertia

    rng = np.random.default_rng(seed)
    hull = synth.sphere_hull(128, 7) * 0.005
    mid = ctx.register_mesh(hull)
    _, props, _ = O.build_mesh(hull)
    V = props[0]
    dens = 2500.0
    I = props[7:16].reshape(3, 3) * dens
    floor = synth.box_wall(0.2, 0.2, 0.02)
    fid = ctx.register_mesh(floor)
    n = n_side * n_side
    pos = np.zeros((n, 3))
    for k in range(n):
        pos[k] = [(k % n_side) * 0.012, (k // n_side) * 0.012, 0.0052 + 0.0003 * rng.random()]
    quat = rng.normal(size=(n, 4))
    quat /= np.linalg.norm(quat, axis=1, keepdims=True)
    rad = np.full(n, props[4])
    scene = Scene(zctx=ctx, mesh=np.full(n, mid), radius=rad, mass=torch.full((n,), dens * V, dtype=torch.float64),
                  Ibinv=torch.from_numpy(body_inverse_inertia(I)).expand(n, 3, 3).clone(),
                  req=torch.full((n,), (3 * V / (4 * np.pi)) ** (1 / 3), dtype=torch.float64), floor_mesh=fid,
                  floor_pos=np.array([0.0, 0.0, -0.01]), floor_top=0.0,
                  gravity=torch.tensor([0.0, 0.0, -9.81], dtype=torch.float64), dt=1e-5)
    return scene, torch.from_numpy(pos), torch.from_numpy(quat)


def test_minimal_model_gradients_match_finite_differences(ctx):
    """d loss / d (k, e, v0) through 300 steps with checkpoints every 50
    (adjoint) against central differences with relative step 1e-6."""
    from zdx.autograd import simulate

    scene, pos0, quat0 = _drop_scene(ctx)
    n = pos0.shape[0]
    target = pos0.clone() + torch.tensor([0.0, 0.0, 0.0002])
    v0 = torch.zeros(n, 3, dtype=torch.float64)
    v0[:, 2] = -0.4
    k0 = torch.tensor(8e6, dtype=torch.float64)
    e0 = torch.tensor(0.5, dtype=torch.float64)

    def loss(k, e, v):
        z = torch.zeros(n, 3, dtype=torch.float64)
        p = simulate(scene, k, e, pos0, quat0, v, z, z, 300, 50)[0]
        return ((p - target) ** 2).sum()

    k, e, v = k0.clone().requires_grad_(True), e0.clone().requires_grad_(True), v0.clone().requires_grad_(True)
    L0 = loss(k, e, v)
    L0.backward()
    with torch.no_grad():
        assert loss(k0, e0, v0) == L0.detach()     # deterministic forward
        rel = 1e-6
        checks = [("k", k.grad.item(), lambda s: loss(k0 * (1 + s * rel), e0, v0), float(k0) * rel),
                  ("e", e.grad.item(), lambda s: loss(k0, e0 * (1 + s * rel), v0), float(e0) * rel)]
        for i in range(n):
            def fv(s, i=i):
                vv = v0.clone()
                vv[i, 2] *= (1 + s * rel)
                return loss(k0, e0, vv)
            checks.append((f"v0[{i}].z", v.grad[i, 2].item(), fv, float(v0[i, 2]) * rel))
        for name, g, fn, hstep in checks:
            fd = (float(fn(1)) - float(fn(-1))) / (2 * hstep)
            assert abs(g - fd) <= 1e-4 * max(abs(fd), abs(g)) + 1e-12, (name, g, fd)


def test_calibration_reduces_loss(ctx):
    """Drop-on-plane calibration (16 hulls of 128 points, 300 steps): the
    positions of a run with (k*, e*, v0*) are the target; Adam on (log k,
    logit e, v0) with adjoint gradients from a perturbed start lowers the loss."""
    from zdx import workloads as W
    from zdx.autograd import calibrate, scene_from_workload, simulate

    wl = W.drop_on_plane(16)
    scene, pos0, quat0, vel0 = scene_from_workload(ctx, wl)
    n = pos0.shape[0]
    k_true, e_true = torch.tensor(4e6, dtype=torch.float64), torch.tensor(0.6, dtype=torch.float64)
    v_true = vel0.clone()
    v_true[:, 2] = -0.3
    with torch.no_grad():
        z = torch.zeros(n, 3, dtype=torch.float64)
        target = simulate(scene, k_true, e_true, pos0, quat0, v_true, z, z, 300, 50)[0]
    v_start = v_true.clone()
    v_start[:, 2] = -0.25
    hist, k_fit, e_fit, v_fit = calibrate(scene, pos0, quat0, target, 2e6, 0.4, v_start, 300, iters=6, lr=0.05)
    losses = [h[0] for h in hist]
    assert losses[0] > 0.0
    assert min(losses[1:]) < 0.5 * losses[0], losses


def test_drop_on_plane_64_gradients_and_calibration(ctx):
    """The drop-on-plane workload at full size (64 hulls of 128 points, r = 5 mm,
    500 steps, dt 1e-5): adjoint d loss / d (k, e, v0) against central
    differences with relative step 1e-6, then a calibration from a perturbed
    start that lowers the loss."""
    import time

    from zdx import workloads as W
    from zdx.autograd import calibrate, scene_from_workload, simulate

    wl = W.drop_on_plane(64)
    scene, pos0, quat0, vel0 = scene_from_workload(ctx, wl)
    n, steps = pos0.shape[0], 500
    assert n == 64
    z = torch.zeros(n, 3, dtype=torch.float64)
    k_true, e_true = torch.tensor(4e6, dtype=torch.float64), torch.tensor(0.6, dtype=torch.float64)
    v_true = vel0.clone()
    v_true[:, 2] = -0.3
    t0 = time.time()
    with torch.no_grad():
        target = simulate(scene, k_true, e_true, pos0, quat0, v_true, z, z, steps, 50)[0]
    print(f"forward {steps} steps: {time.time() - t0:.2f} s", flush=True)

    k0, e0 = torch.tensor(3e6, dtype=torch.float64), torch.tensor(0.5, dtype=torch.float64)
    v0 = v_true.clone()
    v0[:, 2] = -0.28

    def loss(k, e, v):
        return ((simulate(scene, k, e, pos0, quat0, v, z, z, steps, 50)[0] - target) ** 2).sum()

    k, e, v = k0.clone().requires_grad_(True), e0.clone().requires_grad_(True), v0.clone().requires_grad_(True)
    t0 = time.time()
    loss(k, e, v).backward()
    print(f"forward + adjoint: {time.time() - t0:.2f} s", flush=True)
    rel = 1e-6
    with torch.no_grad():
        checks = [("k", k.grad.item(), lambda s: loss(k0 * (1 + s * rel), e0, v0), float(k0) * rel),
                  ("e", e.grad.item(), lambda s: loss(k0, e0 * (1 + s * rel), v0), float(e0) * rel)]
        for i in (0, 27, 63):
            def fv(s, i=i):
                vv = v0.clone()
                vv[i, 2] *= (1 + s * rel)
                return loss(k0, e0, vv)
            checks.append((f"v0[{i}].z", v.grad[i, 2].item(), fv, float(v0[i, 2]) * rel))
        for name, g, fn, h in checks:
            fd = (float(fn(1)) - float(fn(-1))) / (2 * h)
            print(name, g, fd, flush=True)
            assert abs(g - fd) <= 1e-4 * max(abs(fd), abs(g)) + 1e-14, (name, g, fd)

    hist, k_fit, e_fit, v_fit = calibrate(scene, pos0, quat0, target, float(k0), float(e0), v0, steps, iters=6,
                                          lr=0.05)
    losses = [h[0] for h in hist]
    print("calibration losses", losses, "k", k_fit, "e", e_fit, flush=True)
    assert min(losses[1:]) < 0.5 * losses[0], losses
