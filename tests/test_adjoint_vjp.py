"""CPU: the hand-written vector-Jacobian products of the differentiable model
(zdx.autograd) against torch autograd of the same forward formulas, on seeded
random inputs: the integrator (integrate_particle, src/main.cpp:1494-1510),
the reference's wall force (2363-2431) in all its branches (damping clamp,
Coulomb cap, spring reset) and the minimal-energy force (polyline.py:655-701
at the loop centroid).  Tolerance 1e-12 relative to the largest entry."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
from zdx import autograd as A  # noqa: E402

PI2 = 3.141592653589793 * 3.141592653589793


def _close(a, b, tol=1e-12):
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = max(np.max(np.abs(b)), 1e-300)
    assert np.max(np.abs(a - b)) <= tol * scale, (np.max(np.abs(a - b)), scale)


def _t(a):
    return torch.tensor(np.asarray(a, dtype=np.float64), requires_grad=True)


def test_integrator_vjp_matches_autograd():
    rng = np.random.default_rng(1)
    n = 7
    x, v, L = rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    q = rng.normal(size=(n, 4))
    q *= 1.0 + 1e-3 * rng.normal(size=(n, 1))      # not quite unit: the normalisations count
    F, T = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    im = rng.uniform(0.5, 2.0, n)
    B = rng.normal(size=(n, 3, 3))
    Ib = B @ np.swapaxes(B, 1, 2) + np.eye(3)
    g, dt = np.array([0.0, 0.0, -9.81]), 1e-3
    x1, q1, v1, om1, L1, c = A._integrate(x, q, v, L, F, T, im, Ib, g, dt)
    G = [rng.normal(size=a.shape) for a in (x1, q1, v1, om1, L1)]
    got = A._integrate_vjp(c, q, L, T, im, Ib, dt, *G)

    tx, tq, tv, tL, tF, tT = (_t(a) for a in (x, q, v, L, F, T))
    acc = tF * torch.tensor(im)[:, None] + torch.tensor(g)
    tv1 = tv + acc * dt
    tx1 = tx + tv1 * dt
    qn = tq / tq.norm(dim=-1, keepdim=True)
    w, qx, qy, qz = qn.unbind(-1)
    R = torch.stack([torch.stack([1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * w), 2 * (qx * qz + qy * w)], -1),
                     torch.stack([2 * (qx * qy + qz * w), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * w)], -1),
                     torch.stack([2 * (qx * qz - qy * w), 2 * (qy * qz + qx * w), 1 - 2 * (qx * qx + qy * qy)], -1)],
                    -2)
    Iinv = R @ torch.tensor(Ib) @ R.transpose(-1, -2)
    tL1 = tL + tT * dt
    om = (Iinv @ tL1.unsqueeze(-1)).squeeze(-1)
    dq = A.qmul(torch.cat([torch.zeros(n, 1, dtype=torch.float64), om], -1), tq)
    qt = tq + 0.5 * dq * dt
    tq1 = qt / qt.norm(dim=-1, keepdim=True)
    loss = sum((o * torch.tensor(gg)).sum() for o, gg in zip((tx1, tq1, tv1, om, tL1), G))
    loss.backward()
    for name, mine, ref in zip(("x", "q", "v", "L", "F", "T"), got, (tx, tq, tv, tL, tF, tT)):
        _close(mine, ref.grad.numpy())


def _wall_torch(area, xc, cn, x, v, om, ds, young, eqr, poisson, mass, mu, e, dt):
    rP = xc - x
    vrel = v + torch.cross(om, rP, dim=-1)
    vn = (vrel * cn).sum(-1)
    Rp, Ep = torch.tensor(np.maximum(eqr, 1e-12)), torch.tensor(np.maximum(young, 1e-12))
    kn = (Ep / Rp) * area
    kt = (Ep / (2.0 * (1.0 + torch.tensor(poisson)) * Rp)) * area
    ec = torch.clamp(e, 1e-6, 0.9999)
    le = torch.log(ec)
    sq = torch.sqrt(le * le + PI2)
    m = torch.tensor(mass)
    cnn = torch.abs(2.0 * le * torch.sqrt(kn * m) / sq)
    ct = torch.abs(2.0 * le * torch.sqrt(kt * m) / sq)
    fe = kn * torch.sqrt(area)
    fn = torch.clamp(fe - cnn * vn, min=0.0)
    vt = vrel - cn * vn[:, None]
    ds1 = ds + vt * dt
    ds2 = ds1 - cn * (ds1 * cn).sum(-1, keepdim=True)
    ft = -kt[:, None] * ds2 - ct[:, None] * vt
    ftn = ft.norm(dim=-1)
    ftm = torch.tensor(mu) * fn
    cap = (ftn > ftm) & (ftn > 1e-14)
    ft_out = torch.where(cap[:, None], ft * (ftm / torch.where(cap, ftn, torch.ones_like(ftn)))[:, None], ft)
    use = cap & (kt > 1e-12)
    ds_out = torch.where(use[:, None], -ft_out / torch.where(use, kt, torch.ones_like(kt))[:, None], ds2)
    return cn * fn[:, None] + ft_out, rP, ds_out, cap


def test_wall_force_vjp_matches_autograd():
    rng = np.random.default_rng(2)
    n = 64
    area = rng.uniform(1e-8, 4e-7, n)
    xc = rng.normal(scale=3e-3, size=(n, 3))
    cn = rng.normal(size=(n, 3))
    cn /= np.linalg.norm(cn, axis=1, keepdims=True)
    x = rng.normal(scale=3e-3, size=(n, 3))
    v = rng.normal(scale=0.3, size=(n, 3))
    om = rng.normal(scale=20.0, size=(n, 3))
    ds = rng.normal(scale=2e-6, size=(n, 3))
    ds[::3] = 0.0
    young, eqr, poisson = np.full(n, 1e7), rng.uniform(3e-3, 6e-3, n), np.full(n, 0.25)
    mass = rng.uniform(5e-4, 2e-3, n)
    mu = rng.uniform(0.05, 0.8, n)
    e, dt = 0.55, 1e-5
    f, rP, dso, c = A._wall_force(area, xc, cn, x, v, om, ds, young, eqr, poisson, mass, mu, e, dt)
    # every branch is exercised: capped and free springs, clamped normal force
    assert c["cap"].any() and (~c["cap"]).any() and c["clampn"].any() and (~c["clampn"]).any()
    gf, grP, gds = rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    got = A._wall_vjp(c, gf, grP, gds, dt)

    T = [_t(a) for a in (area, xc, cn, x, v, om, ds)]
    te = torch.tensor(e, dtype=torch.float64, requires_grad=True)
    tf, trP, tdso, tcap = _wall_torch(*T, young, eqr, poisson, mass, mu, te, dt)
    assert np.array_equal(tcap.numpy(), c["cap"])
    _close(tf.detach().numpy(), f, 1e-13)
    _close(tdso.detach().numpy(), dso, 1e-13)
    ((tf * torch.tensor(gf)).sum() + (trP * torch.tensor(grP)).sum() + (tdso * torch.tensor(gds)).sum()).backward()
    for name, key, t in (("area", "area", T[0]), ("xc", "xc", T[1]), ("x", "x", T[3]), ("v", "v", T[4]),
                         ("omega", "om", T[5]), ("ds", "ds", T[6])):
        _close(got[key], t.grad.numpy())
    _close(got["e"], te.grad.item())


def test_minimal_force_vjp_matches_autograd():
    rng = np.random.default_rng(3)
    n = 9
    Sc, xc0 = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    xA, xB = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    k = 3.7e6
    f = -0.5 * k * Sc
    c = -1.5 * xc0
    rA, rB = c - xA, c - xB
    G = [rng.normal(size=(n, 3)) for _ in range(4)]
    got = A._pp_vjp(Sc, rA, rB, f, k, *G)
    tS, tx, ta, tb = (_t(a) for a in (Sc, xc0, xA, xB))
    tk = torch.tensor(k, dtype=torch.float64, requires_grad=True)
    area = tS.norm(dim=-1, keepdim=True)
    nrm = -tS / area
    tf = 0.5 * tk * area * nrm
    tc = -1.5 * tx
    FA, FB = -tf, tf
    TA, TB = torch.cross(tc - ta, -tf, dim=-1), torch.cross(tc - tb, tf, dim=-1)
    sum((o * torch.tensor(g)).sum() for o, g in zip((FA, FB, TA, TB), G)).backward()
    for key, t in (("Sc", tS), ("xc0", tx), ("xA", ta), ("xB", tb)):
        _close(got[key], t.grad.numpy())
    _close(got["k"], tk.grad.item())
