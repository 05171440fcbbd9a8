"""GPU: derivatives of the contact geometry and of the differentiable model
(zdx.autograd: the device step with the minimal-energy force and the
reference's wall force, hand-written adjoint) against central finite
differences of the same device forward.

* Loop Jacobians (dSn/dxi, dGn/dxi) and wall-contact Jacobians (d area/dxi,
  d point/dxi) for rigid motions xi of the bodies.
* The drop-on-plane model: d loss / d (k, e, v0) with checkpoints, against
  central differences: 16 bodies over 300 steps, 1e-6 relative against the
  fourth-order difference at relative step 1e-4 (consistent with the same
  stencil at 5e-5), and within the difference's own rounding noise at
  relative step 1e-6; 64 bodies over 500 steps, whose loss is smooth only at
  much smaller steps, at relative step 1e-6 to 2e-5; then a calibration that
  lowers the loss.
"""
import numpy as np
import pytest

import golden_inputs as GI
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
                assert np.max(np.abs(ana[l, part] - num[l, part])) <= 1e-6 * scale, (na, nb, l)


def test_wall_jacobians_match_finite_differences(ctx):
    """Sections of hulls and stars by wall triangles (one clipped by the
    triangle's edges): d area / dxi and d point / dxi against central
    differences of zdx_wall_contacts."""
    m = GI.synth_meshes()
    rng = np.random.default_rng(5)
    tri = np.array([[[-3.0, -3.0, 0.0], [3.0, -3.0, 0.0], [0.0, 3.0, 0.0]],
                    [[-0.2, -3.0, 0.0], [3.0, -3.0, 0.0], [0.1, 3.0, 0.0]]])      # the second clips
    nrm = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]])
    checked = 0
    for name in ("hull128", "star2a", "ellipsoid200"):
        mid = ctx.register_mesh(m[name])
        for _ in range(3):
            q = rng.normal(size=4)
            q /= np.linalg.norm(q)
            p = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(-0.5, 0.5)])
            r0 = ctx.wall_contacts([mid], [p], [q], tri, nrm, jac=True)[0]
            h = 1e-7
            num = np.zeros((2, 4, 6))
            ok = np.ones(2, dtype=bool)
            for k in range(6):
                out = []
                for sgn in (1, -1):
                    xi = np.zeros(6)
                    xi[k] = sgn * h
                    out.append(ctx.wall_contacts([mid], [p + xi[0:3]], [_qmul(_qexp(xi[3:6]), q)], tri, nrm)[0])
                ok &= (out[0][:, 0] == r0[:, 0]) & (out[1][:, 0] == r0[:, 0])
                num[:, 0, k] = (out[0][:, 1] - out[1][:, 1]) / (2 * h)
                num[:, 1:4, k] = (out[0][:, 2:5] - out[1][:, 2:5]) / (2 * h)
            for t in range(2):
                if r0[t, 0] < 0.5 or not ok[t]:
                    continue
                ana = np.concatenate([r0[t, 8:14][None], r0[t, 14:32].reshape(3, 6)], 0)
                scale = np.max(np.abs(ana))
                assert np.max(np.abs(ana - num[t])) <= 1e-6 * scale, (name, t, ana, num[t])
                checked += 1
    assert checked >= 8


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
        assert np.max(np.abs(g - num)) <= 1e-6 * scale, (name, g, num)


def _fd_checks(loss, k0, e0, v0, kg, eg, vg, particles, fourth=True):
    """Central differences of the loss in k, e and some initial velocity
    components: the second-order difference at relative step 1e-6 and (with
    fourth) the fourth-order stencil at relative steps 1e-4 and 5e-5.
    Returns [(name, adjoint, fd(1e-6), fd4(1e-4), fd4(5e-5))].

    The loss of a contact trajectory carries rounding differences that the
    contact dynamics amplify over the steps, so a difference quotient's own
    error grows like 1/h: about 1e-5 relative at the step 1e-6 (measured on
    these runs), a few 1e-7 at 1e-4.  The fourth-order stencil keeps the
    truncation error at 1e-4 negligible, and its two steps agreeing shows the
    loss is smooth at that scale (no contact starting or ending inside the
    step) -- true for the shorter run, not over 500 steps of 64 bodies, where
    only the step 1e-6 stays inside one contact sequence."""
    res = []

    def diffs(fn, base):
        d2 = (float(fn(1e-6)) - float(fn(-1e-6))) / (2e-6 * base)
        if not fourth:
            return d2, float("nan"), float("nan")
        d4 = lambda r: (-float(fn(2 * r)) + 8 * float(fn(r)) - 8 * float(fn(-r)) + float(fn(-2 * r))) / (12 * r * base)
        return d2, d4(1e-4), d4(5e-5)

    with torch.no_grad():
        res.append(("k", kg, *diffs(lambda s: loss(k0 * (1 + s), e0, v0), float(k0))))
        res.append(("e", eg, *diffs(lambda s: loss(k0, e0 * (1 + s), v0), float(e0))))
        for i, c in particles:
            def fv(s, i=i, c=c):
                vv = v0.clone()
                vv[i, c] *= (1 + s)
                return loss(k0, e0, vv)
            res.append((f"v0[{i},{c}]", float(vg[i, c]), *diffs(fv, float(v0[i, c]))))
    for name, g, d6, da, db in res:
        print(f"[adjoint] {name}: adjoint {g:.12e}; relative to it: fd(1e-6) {(d6 - g) / g:+.2e}, "
              f"fd4(1e-4) {(da - g) / g:+.2e}, fd4(5e-5) {(db - g) / g:+.2e}", flush=True)
    return res


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b))


def _drop(ctx, n):
    from zdx import workloads as W
    from zdx.autograd import scene_from_workload

    wl = W.drop_on_plane(n)
    return scene_from_workload(ctx, wl)


def test_drop_gradients_match_finite_differences(ctx):
    """16 hulls dropped onto the plane wall, 300 steps, checkpoints every 50:
    adjoint d loss / d (k, e, v0) against central differences (see
    _fd_checks): 1e-6 relative where the difference itself is that accurate."""
    from zdx.autograd import simulate

    scene, pos0, quat0, vel0 = _drop(ctx, 16)
    n = pos0.shape[0]
    z = torch.zeros(n, 3, dtype=torch.float64)
    target = pos0 + torch.tensor([0.0, 0.0, 0.0002], dtype=torch.float64)
    k0, e0 = torch.tensor(4e6, dtype=torch.float64), torch.tensor(0.6, dtype=torch.float64)
    v0 = vel0.clone()
    v0[:, 2] = -0.3

    def loss(k, e, v):
        return ((simulate(scene, k, e, pos0, quat0, v, z, z, 300, 50)[0] - target) ** 2).sum()

    k, e, v = k0.clone().requires_grad_(True), e0.clone().requires_grad_(True), v0.clone().requires_grad_(True)
    L0 = loss(k, e, v)
    L0.backward()
    with torch.no_grad():
        assert float(loss(k0, e0, v0)) == float(L0)     # deterministic forward
    assert abs(k.grad.item()) > 0 and abs(e.grad.item()) > 0
    for name, g, d6, da, db in _fd_checks(loss, k0, e0, v0, k.grad.item(), e.grad.item(), v.grad,
                                          [(0, 0), (5, 1), (15, 2)]):
        assert _rel(da, db) <= 1e-6, (name, "loss not smooth at the steps 5e-5..2e-4", da, db)
        assert _rel(g, da) <= 1e-6, (name, g, da)
        assert _rel(g, d6) <= 2e-4, (name, g, d6)     # the step 1e-6: within the difference's own noise


def test_model_matches_device_step(ctx):
    """The model's forward and zdx_sim_step with the minimal force on the
    same drop (pairs and wall contacts): same trajectories to the last bits
    of cbrt/pow/log in the wall force."""
    from zdx import workloads as W
    from zdx.autograd import simulate

    wl = W.drop_on_plane(16)
    wl.init[:, 9] = -0.3          # down at 0.3 m/s: the wall is reached within the steps
    scene, pos0, quat0, vel0 = _drop(ctx, 16)
    sim, _ = W.build_sim(ctx, wl)
    k = 4e6
    sim.set_force_model(zdx.FORCE_MINIMAL, k)
    s0 = sim.state()
    steps = 300
    tot = {"contacts": 0, "wall_contacts": 0}
    for _ in range(steps):
        c = sim.step(1, zdx.MODE_BATCHED)
        tot["contacts"] += c["contacts"]
        tot["wall_contacts"] += c["wall_contacts"]
    sg = sim.state()
    sim.close()
    print(f"[adjoint] device drop: {tot}", flush=True)
    assert tot["contacts"] > 20 and tot["wall_contacts"] > 100
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64))
    with torch.no_grad():
        st = simulate(scene, t(k), t(0.5), t(s0[:, 0:3]), t(s0[:, 3:7]), t(s0[:, 7:10]), t(s0[:, 10:13]),
                      t(s0[:, 13:16]), steps, 50)
    st = np.concatenate([x.numpy() for x in st], axis=1)
    err = np.abs(st - sg) / np.maximum(np.abs(sg), 1e-3)
    assert np.max(err) <= 1e-9, (np.max(err), np.unravel_index(np.argmax(err), err.shape))


def test_drop_on_plane_64_gradients_and_calibration(ctx):
    """The drop-on-plane workload at full size (64 hulls of 128 points, r = 5 mm,
    500 steps, dt 1e-5): adjoint d loss / d (k, e, v0) against central
    differences (see _fd_checks), then a calibration from a perturbed start
    that lowers the loss."""
    import time

    from zdx.autograd import calibrate, simulate

    scene, pos0, quat0, vel0 = _drop(ctx, 64)
    n, steps = pos0.shape[0], 500
    assert n == 64
    z = torch.zeros(n, 3, dtype=torch.float64)
    k_true, e_true = torch.tensor(4e6, dtype=torch.float64), torch.tensor(0.6, dtype=torch.float64)
    v_true = vel0.clone()
    v_true[:, 2] = -0.3
    t0 = time.time()
    with torch.no_grad():
        target = simulate(scene, k_true, e_true, pos0, quat0, v_true, z, z, steps, 50)[0]
    print(f"[adjoint] forward {steps} steps: {time.time() - t0:.2f} s", flush=True)

    k0, e0 = torch.tensor(3e6, dtype=torch.float64), torch.tensor(0.5, dtype=torch.float64)
    v0 = v_true.clone()
    v0[:, 2] = -0.28

    def loss(k, e, v):
        return ((simulate(scene, k, e, pos0, quat0, v, z, z, steps, 50)[0] - target) ** 2).sum()

    k, e, v = k0.clone().requires_grad_(True), e0.clone().requires_grad_(True), v0.clone().requires_grad_(True)
    t0 = time.time()
    loss(k, e, v).backward()
    print(f"[adjoint] forward + adjoint: {time.time() - t0:.2f} s", flush=True)
    # the loss is smooth only at steps far below 1e-4 here: the step 1e-6,
    # whose own rounding noise is about 1e-12 absolute (1e-5 relative for the
    # larger gradients); the velocity components checked are the three with
    # the largest gradients
    big = np.argsort(-np.abs(v.grad.numpy()).ravel(), kind="stable")[:3]
    for name, g, d6, _, _ in _fd_checks(loss, k0, e0, v0, k.grad.item(), e.grad.item(), v.grad,
                                        [(int(b) // 3, int(b) % 3) for b in big], fourth=False):
        assert _rel(g, d6) <= 2e-5, (name, g, d6)

    t0 = time.time()
    hist, k_fit, e_fit, v_fit = calibrate(scene, pos0, quat0, target, float(k0), float(e0), v0, steps, iters=16)
    losses = [h[0] for h in hist]
    print(f"[adjoint] calibration ({time.time() - t0:.1f} s): losses {losses}, k {k_fit}, e {e_fit}, "
          f"|v0 - v_true| {float((v_fit - v_true).norm()):.3e} (start {float((v0 - v_true).norm()):.3e})", flush=True)
    assert losses[-1] < 0.5 * losses[0], losses
