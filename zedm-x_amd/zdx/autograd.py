"""Differentiable time step for calibration, with a hand-written adjoint.

The model is the device time step with the minimal-energy particle-particle
force (``zdx_sim_set_force_model(ZDX_FORCE_MINIMAL, k)``), term for term:

* candidate pairs by the reference's sphere test (src/main.cpp:2049-2059);
  their contact loops from ``zdx_contact_pairs`` (the device pipeline);
* particle-particle force of polyline.py:655-701, ``F_B = k/2 |Sn| n`` along
  the A->B normal, applied at the loop centroid (``-3/2 xc0`` up to a
  multiple of n, which the torque of a normal force does not see); no
  damping, no tangential force;
* particle-wall contacts after the reference's sphere and box pre-tests
  (src/main.cpp:2318-2343), their geometry from ``zdx_wall_contacts`` (the
  device's compute_wall_contact, 1205-1258), and the reference's wall force
  (2363-2431): ``kn = (E/R) A``, elastic ``kn A^(1/2)``, restitution damping
  ``-cn vn``, tangential spring per (particle, wall) with Coulomb cap, the
  spring erased when the pair has no contact (2461-2468);
* forces summed per particle in the reference's order (partners ascending,
  loops in order, then walls), then ``integrate_particle`` (1494-1510; omega
  from the old rotation, the reference's inverse-inertia rule).

The restitution enters through the wall force, the stiffness k through the
particle-particle force.

Adjoint (no torch autograd anywhere): ``Trajectory`` is a
``torch.autograd.Function`` over (k, e, initial state).  Its forward keeps the
state every ``ckpt_every`` steps; its backward re-runs each segment forward
from its checkpoint to record the segment's tape, then walks the tape back
with the vector-Jacobian products written below: the integrator
(``_integrate_vjp``), the wall force (``_wall_vjp``), the minimal force
(``_pp_vjp``), and the contact geometry -- the device's Jacobians of Sn and Gn
per loop and of the wall section's area and point, with respect to rigid
motions of the bodies, chained to positions and quaternions
(``_quat_grad``).  ``PairContact`` exposes the loop geometry alone as a
``torch.autograd.Function``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

PI2 = 3.141592653589793 * 3.141592653589793


def _skew(v):
    z = np.zeros(v.shape[:-1])
    return np.stack([np.stack([z, -v[..., 2], v[..., 1]], -1), np.stack([v[..., 2], z, -v[..., 0]], -1),
                     np.stack([-v[..., 1], v[..., 0], z], -1)], -2)


def _cross(a, b):
    """a x b with the component formulas of the reference's cross."""
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1], a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)


def _dot(a, b):
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def _quat_grad(q, g_w):
    """dL/dq for the world-frame rotation increment w (q' = exp(w/2) q, q
    normalised inside the rotation): 2 ((0, g_w) x q_hat) / |q|."""
    nq = np.linalg.norm(q, axis=-1, keepdims=True)
    qh = q / nq
    w, v = qh[..., :1], qh[..., 1:]
    gw = -(g_w * v).sum(-1, keepdims=True)
    gv = w * g_w + np.cross(g_w, v)
    return 2.0 * np.concatenate([gw, gv], -1) / nq


# ------------------------------------------------------------------ geometry
def _loop_geometry(zctx, mesh_a, mesh_b, pA, qA, pB, qB, split):
    """Contact loops of the pairs and their derivatives: per loop the pair
    index, the canonical area vector Sc = -|Sn| n (n the A->B normal), xc0,
    and dSc/dxi, dxc0/dxi (3 x 12 each, xi = (dA, wA, dB, wB))."""
    r = zctx.contact_pairs(mesh_a, mesh_b, pA, qA, pB, qB, split=split, k=1.0, vertices=False, sensitivities=True)
    rec, jac, lo = r["loops"], r["loop_jac"], r["loop_offset"]
    nl = len(rec)
    pair = np.repeat(np.arange(len(mesh_a)), np.diff(lo)).astype(np.int64)
    Sn, Gn, n, xc = rec[:, 0:3], rec[:, 3:6], rec[:, 6:9], rec[:, 9:12]
    area = rec[:, 12]
    s = np.sign((n * (-Sn)).sum(-1)) if nl else np.zeros(0)   # n = s (-Sn/|Sn|)
    dSn = jac[:, 0:36].reshape(nl, 3, 12)
    dGn = jac[:, 36:72].reshape(nl, 3, 12)
    # xc0 = -(Sn x Gn) / |Sn|^2
    a2 = (area ** 2)[:, None, None]
    c = np.cross(Sn, Gn)
    dc = -_skew(Gn) @ dSn + _skew(Sn) @ dGn
    dxc = -dc / a2 + 2.0 * c[:, :, None] * np.einsum("li,lij->lj", Sn, dSn)[:, None, :] / (a2 * a2)
    return pair, s[:, None] * Sn, xc.copy(), s[:, None, None] * dSn, dxc, n


class PairContact(torch.autograd.Function):
    """Per contact loop of a batch of pairs: the canonical area vector
    Sc = -|Sn| n and the contact point xc0 (src/main.cpp:1027-1052) as
    functions of the poses; forward zdx_contact_pairs, backward the device's
    loop Jacobians chained to positions and quaternions.  The loop topology
    is held fixed, as a derivative must."""

    @staticmethod
    def forward(ctx, posA, quatA, posB, quatB, zctx, mesh_a, mesh_b, split):
        pair, Sc, xc, dSc, dxc, _ = _loop_geometry(zctx, mesh_a, mesh_b, posA.detach().cpu().numpy(),
                                                   quatA.detach().cpu().numpy(), posB.detach().cpu().numpy(),
                                                   quatB.detach().cpu().numpy(), split)
        ctx.save_for_backward(torch.from_numpy(dSc), torch.from_numpy(dxc), torch.from_numpy(pair),
                              quatA.detach().cpu(), quatB.detach().cpu())
        ctx.n_pairs = len(mesh_a)
        ctx.dev = posA.device
        lp = torch.from_numpy(pair).to(posA.device)
        ctx.mark_non_differentiable(lp)
        return torch.from_numpy(Sc).to(posA), torch.from_numpy(xc).to(posA), lp

    @staticmethod
    def backward(ctx, gSc, gxc, _gpair):
        dSc, dxc, pair, qA, qB = ctx.saved_tensors
        G = _geometry_vjp(gSc.detach().cpu().numpy(), gxc.detach().cpu().numpy(), dSc.numpy(), dxc.numpy(),
                          pair.numpy(), ctx.n_pairs)
        gpA = torch.from_numpy(G[:, 0:3]).to(ctx.dev)
        gqA = torch.from_numpy(_quat_grad(qA.numpy(), G[:, 3:6])).to(ctx.dev)
        gpB = torch.from_numpy(G[:, 6:9]).to(ctx.dev)
        gqB = torch.from_numpy(_quat_grad(qB.numpy(), G[:, 9:12])).to(ctx.dev)
        return gpA, gqA, gpB, gqB, None, None, None, None


def _geometry_vjp(gSc, gxc, dSc, dxc, pair, n_pairs):
    """Per pair, the gradient with respect to xi = (dA, wA, dB, wB)."""
    gxi = np.einsum("li,lij->lj", gSc, dSc) + np.einsum("li,lij->lj", gxc, dxc)
    G = np.zeros((n_pairs, 12))
    np.add.at(G, pair, gxi)
    return G


def qmul(a, b):
    """Quaternion product (w, x, y, z), src/main.cpp:76-83, for torch tensors."""
    w1, x1, y1, z1 = a.unbind(-1)
    w2, x2, y2, z2 = b.unbind(-1)
    return torch.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                        w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], -1)


def body_inverse_inertia(I: np.ndarray) -> np.ndarray:
    """Inverse body inertia as the time step uses it: the reference's
    mat3_inverse (src/main.cpp:203-207) returns the identity when
    |det I| < 1e-18, which holds for every particle below about a centimetre
    at mineral densities (I ~ 1e-8 kg m^2), so their omega equals L."""
    I = np.asarray(I, dtype=np.float64)
    det = np.linalg.det(I)
    out = np.linalg.inv(np.where((np.abs(det) < 1e-18)[..., None, None], np.eye(3), I))
    return np.where((np.abs(det) < 1e-18)[..., None, None], np.eye(3), out)


# ------------------------------------------------------------------ scene
@dataclass
class Wall:
    """A wall as the step holds it: world triangles (T, 3, 3), their unit
    normals (T, 3), the wall position and bounding radius (sphere pre-test),
    friction coefficient."""
    tris: np.ndarray
    normals: np.ndarray
    pos: np.ndarray
    radius: float
    mu: float


@dataclass
class Scene:
    """Particles (registered meshes and their constants) and walls."""
    zctx: object                # zdx.Context: contact_pairs and wall_contacts
    mesh: np.ndarray            # per particle mesh id
    radius: np.ndarray          # bounding radius per particle (sphere pre-tests)
    mass: np.ndarray            # (N,)
    inv_mass: np.ndarray        # (N,)
    Ibinv: np.ndarray           # (N, 3, 3) body-frame inverse inertia (reference rule)
    eqr: np.ndarray             # (N,) equivalent-sphere radius cbrt(3V/4pi)
    young: np.ndarray
    poisson: np.ndarray
    mu: np.ndarray
    walls: list = field(default_factory=list)
    gravity: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, -9.81]))
    dt: float = 1e-5
    split: bool = True


def _rotate(q, v):
    """quat_rotate, src/main.cpp:93-98 (q normalised, q (0, v) q*)."""
    n = math.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3])
    w, x, y, z = q[0] / n, q[1] / n, q[2] / n, q[3] / n
    px, py, pz = v[..., 0], v[..., 1], v[..., 2]
    # (qn * p)
    aw = -x * px - y * py - z * pz
    ax = w * px + y * pz - z * py
    ay = w * py - x * pz + z * px
    az = w * pz + x * py - y * px
    # (qn * p) * conj(qn)
    cx, cy, cz = -x, -y, -z
    rx = aw * cx + ax * w + ay * cz - az * cy
    ry = aw * cy - ax * cz + ay * w + az * cx
    rz = aw * cz + ax * cy - ay * cx + az * w
    return np.stack([rx, ry, rz], -1)


def make_wall(zctx, tris, pos, quat, mu, center=True) -> Wall:
    """A wall as the reference builds it (src/main.cpp:1990-2021): its mesh
    through build_mesh (centred at its centroid when center), placed by
    transform_tris, normals by tri_normal (512-516)."""
    tris = np.asarray(tris, dtype=np.float64)
    mid = zctx.register_mesh(tris, center)
    props = zctx.mesh_props[mid]
    if center and props[0] > 0.0:
        tris = tris - props[1:4]
    w = _rotate(np.asarray(quat, dtype=np.float64), tris) + np.asarray(pos, dtype=np.float64)
    e1, e2 = w[:, 1] - w[:, 0], w[:, 2] - w[:, 0]
    n = _cross(e1, e2)
    n = n / np.sqrt(_dot(n, n))[:, None]
    return Wall(tris=w, normals=n, pos=np.asarray(pos, dtype=np.float64), radius=float(props[4]), mu=float(mu))


# ------------------------------------------------------------------ forces
def _wall_force(area, xc, cn, x, v, om, ds, young, eqr, poisson, mass, mu, e, dt):
    """The reference's particle-wall force (src/main.cpp:2363-2431),
    vectorised over contacts.  Returns (f, rP, ds_out, cache)."""
    rP = xc - x
    vrel = v + _cross(om, rP)
    vn = _dot(vrel, cn)
    Rp = np.maximum(eqr, 1e-12)
    Ep = np.maximum(young, 1e-12)
    kn = (Ep / Rp) * area
    kt = (Ep / (2.0 * (1.0 + poisson) * Rp)) * area
    ec = min(0.9999, max(1e-6, e))
    le = math.log(ec)
    sq = math.sqrt(le * le + PI2)
    cnn = np.abs(2.0 * le * np.sqrt(kn * mass) / sq)
    ct = np.abs(2.0 * le * np.sqrt(kt * mass) / sq)
    fe = kn * np.cbrt(np.power(area, 1.5))
    fd = -cnn * vn
    fn_raw = fe + fd
    clampn = fn_raw < 0.0
    fn = np.where(clampn, 0.0, fn_raw)
    fnv = cn * fn[:, None]
    vt = vrel - cn * vn[:, None]
    ds1 = ds + vt * dt
    ds2 = ds1 - cn * _dot(ds1, cn)[:, None]
    ft = ds2 * (-kt)[:, None] + vt * (-ct)[:, None]
    ftn = np.sqrt(_dot(ft, ft))
    ftm = mu * fn
    cap = (ftn > ftm) & (ftn > 1e-14)
    scale = np.where(cap, ftm / np.where(cap, ftn, 1.0), 1.0)
    ft_out = np.where(cap[:, None], ft * scale[:, None], ft)
    use_kt = cap & (kt > 1e-12)
    ds_out = np.where(use_kt[:, None], ft_out * (-1.0 / np.where(use_kt, kt, 1.0))[:, None], ds2)
    f = fnv + ft_out
    cache = dict(rP=rP, vrel=vrel, vn=vn, kn=kn, kt=kt, Ep=Ep, Rp=Rp, le=le, sq=sq, ec=ec, cnn=cnn, ct=ct,
                 clampn=clampn, fn=fn, vt=vt, ds2=ds2, ft=ft, ftn=ftn, ftm=ftm, cap=cap, use_kt=use_kt, ft_out=ft_out,
                 area=area, cn=cn, om=om, poisson=poisson, mass=mass, mu=mu, e=e)
    return f, rP, ds_out, cache


def _wall_vjp(c, g_f, g_rP, g_ds_out, dt):
    """Adjoint of _wall_force: gradients of area, xc, x, v, omega, ds and e
    from those of f, of the lever arm rP and of the spring ds_out."""
    cn = c["cn"]
    g_ftout = g_f.copy()
    g_fn = _dot(cn, g_f)
    g_kt = np.zeros(len(g_f))
    g_ds2 = np.where(c["use_kt"][:, None], 0.0, g_ds_out)
    # ds_out = -ft_out / kt where the cap set the spring
    kt_safe = np.where(c["use_kt"], c["kt"], 1.0)
    g_ftout = g_ftout + np.where(c["use_kt"][:, None], -g_ds_out / kt_safe[:, None], 0.0)
    g_kt += np.where(c["use_kt"], _dot(c["ft_out"], g_ds_out) / (kt_safe * kt_safe), 0.0)
    # Coulomb cap: ft_out = ftm ft / |ft|
    cap = c["cap"]
    ftn_safe = np.where(cap, c["ftn"], 1.0)
    u = c["ft"] / ftn_safe[:, None]
    g_ftm = np.where(cap, _dot(u, g_ftout), 0.0)
    g_u = c["ftm"][:, None] * g_ftout
    g_ft_cap = (g_u - u * _dot(u, g_u)[:, None]) / ftn_safe[:, None]
    g_ft = np.where(cap[:, None], g_ft_cap, g_ftout)
    g_fn = g_fn + c["mu"] * g_ftm
    # ft = -kt ds2 - ct vt
    g_kt += -_dot(c["ds2"], g_ft)
    g_ds2 = g_ds2 - c["kt"][:, None] * g_ft
    g_ct = -_dot(c["vt"], g_ft)
    g_vt = -c["ct"][:, None] * g_ft
    # ds2 = ds1 - cn (ds1 . cn); ds1 = ds + vt dt
    g_ds1 = g_ds2 - cn * _dot(cn, g_ds2)[:, None]
    g_ds = g_ds1
    g_vt = g_vt + dt * g_ds1
    # vt = vrel - cn vn
    g_vrel = g_vt.copy()
    g_vn = -_dot(cn, g_vt)
    # fn = max(0, fe + fd), fd = -cnn vn, fe = kn cbrt(area^1.5) = kn area^(1/2)
    g_raw = np.where(c["clampn"], 0.0, g_fn)
    g_cnn = -c["vn"] * g_raw
    g_vn = g_vn - c["cnn"] * g_raw
    sa = np.sqrt(c["area"])
    g_kn = sa * g_raw
    g_area = c["kn"] * 0.5 / sa * g_raw
    # cnn = -2 le sqrt(kn m) / sq, ct likewise with kt (le < 0)
    g_kn += np.where(c["kn"] > 0, g_cnn * c["cnn"] / (2.0 * np.where(c["kn"] > 0, c["kn"], 1.0)), 0.0)
    g_kt += np.where(c["kt"] > 0, g_ct * c["ct"] / (2.0 * np.where(c["kt"] > 0, c["kt"], 1.0)), 0.0)
    le, sq = c["le"], c["sq"]
    dC_dle = -2.0 * PI2 / (sq * sq * sq)       # d(-2 le / sq)/dle
    g_le = (g_cnn * np.sqrt(c["kn"] * c["mass"]) + g_ct * np.sqrt(c["kt"] * c["mass"])) * dC_dle
    g_e = float(g_le.sum()) / c["ec"] if 1e-6 < c["e"] < 0.9999 else 0.0
    # kn = (Ep/Rp) area, kt = (Ep/(2(1+nu)Rp)) area
    g_area = g_area + (c["Ep"] / c["Rp"]) * g_kn + (c["Ep"] / (2.0 * (1.0 + c["poisson"]) * c["Rp"])) * g_kt
    # vn = vrel . cn; vrel = v + om x rP; rP = xc - x
    g_vrel = g_vrel + cn * g_vn[:, None]
    g_v = g_vrel
    g_om = _cross(c["rP"], g_vrel)
    g_rP = g_rP + _cross(g_vrel, c["om"])
    return dict(area=g_area, xc=g_rP, x=-g_rP, v=g_v, om=g_om, ds=g_ds, e=g_e)


def _pp_vjp(Sc, rA, rB, f, k, gFa, gFb, gTa, gTb):
    """Adjoint of the minimal force f = -k/2 Sc at c = -3/2 xc0 (force on B,
    -f on A, torques rA x (-f) and rB x f): gradients of Sc, xc0, xA, xB, k."""
    g_f = gFb - gFa + _cross(gTb, rB) - _cross(gTa, rA)
    g_rA = _cross(-f, gTa)
    g_rB = _cross(f, gTb)
    g_c = g_rA + g_rB
    return dict(Sc=-0.5 * k * g_f, xc0=-1.5 * g_c, xA=-g_rA, xB=-g_rB, k=float(-0.5 * _dot(Sc, g_f).sum()))


# ------------------------------------------------------------------ integrator
def _qmat(q):
    """quat_to_mat3 (src/main.cpp:150-164), q normalised first."""
    n = np.sqrt(q[..., 0] * q[..., 0] + q[..., 1] * q[..., 1] + q[..., 2] * q[..., 2] + q[..., 3] * q[..., 3])[..., None]
    qn = q / n
    w, x, y, z = qn[..., 0], qn[..., 1], qn[..., 2], qn[..., 3]
    R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
                  np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
                  np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], -2)
    return R, qn, n[..., 0]


def _mat_mul(a, b):
    """mat3_mul (src/main.cpp:175-186): each entry summed k = 0, 1, 2 from 0."""
    r = np.zeros(np.broadcast_shapes(a.shape, b.shape))
    for i in range(3):
        for j in range(3):
            acc = np.zeros(r.shape[:-2])
            for kk in range(3):
                acc = acc + a[..., i, kk] * b[..., kk, j]
            r[..., i, j] = acc
    return r


def _integrate(x, q, v, L, F, T, inv_mass, Ibinv, g, dt):
    """integrate_particle, src/main.cpp:1494-1510, with its operation order."""
    acc = F * inv_mass[:, None] + g
    v1 = v + acc * dt
    x1 = x + v1 * dt
    R, qn, nq = _qmat(q)
    Iinv = _mat_mul(_mat_mul(R, Ibinv), np.swapaxes(R, -1, -2))
    L1 = L + T * dt
    om1 = np.stack([Iinv[:, i, 0] * L1[:, 0] + Iinv[:, i, 1] * L1[:, 1] + Iinv[:, i, 2] * L1[:, 2] for i in range(3)],
                   -1)
    w, qx, qy, qz = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    ox, oy, oz = om1[:, 0], om1[:, 1], om1[:, 2]
    # quat_mul((0, om), q)
    qd = np.stack([0.0 * w - ox * qx - oy * qy - oz * qz, 0.0 * qx + ox * w + oy * qz - oz * qy,
                   0.0 * qy - ox * qz + oy * w + oz * qx, 0.0 * qz + ox * qy - oy * qx + oz * w], -1)
    qt = q + 0.5 * qd * dt
    nt = np.sqrt(qt[:, 0] * qt[:, 0] + qt[:, 1] * qt[:, 1] + qt[:, 2] * qt[:, 2] + qt[:, 3] * qt[:, 3])
    q1 = qt / nt[:, None]
    return x1, q1, v1, om1, L1, dict(R=R, qn=qn, nq=nq, Iinv=Iinv, qt=qt, nt=nt)


# partial derivatives of quat_to_mat3's entries with respect to (w, x, y, z) of
# the normalised quaternion: dR[i][j] / dq[a] = sum_b C[i, j, a, b] q[b]
_C = np.zeros((3, 3, 4, 4))
for (i, j, terms) in [(0, 0, [(2, 2, -4), (3, 3, -4)]), (0, 1, [(0, 3, -2), (1, 2, 2), (2, 1, 2), (3, 0, -2)]),
                      (0, 2, [(0, 2, 2), (1, 3, 2), (2, 0, 2), (3, 1, 2)]),
                      (1, 0, [(0, 3, 2), (1, 2, 2), (2, 1, 2), (3, 0, 2)]), (1, 1, [(1, 1, -4), (3, 3, -4)]),
                      (1, 2, [(0, 1, -2), (1, 0, -2), (2, 3, 2), (3, 2, 2)]),
                      (2, 0, [(0, 2, -2), (1, 3, 2), (2, 0, -2), (3, 1, 2)]),
                      (2, 1, [(0, 1, 2), (1, 0, 2), (2, 3, 2), (3, 2, 2)]), (2, 2, [(1, 1, -4), (2, 2, -4)])]:
    for a, b, cval in terms:
        _C[i, j, a, b] = cval


def _integrate_vjp(c, q, L, T, inv_mass, Ibinv, dt, gx1, gq1, gv1, gom1, gL1):
    """Adjoint of _integrate: gradients of x, q, v, L, F, T."""
    # q1 = qt / |qt|
    q1 = c["qt"] / c["nt"][:, None]
    g_qt = (gq1 - q1 * (q1 * gq1).sum(-1, keepdims=True)) / c["nt"][:, None]
    # qt = q + dt/2 quat_mul((0, om1), q)
    g_qd = 0.5 * dt * g_qt
    L1 = L + T * dt
    Iinv = c["Iinv"]
    om1 = np.einsum("nij,nj->ni", Iinv, L1)
    v = q[:, 1:]
    g_om = gom1 - v * g_qd[:, :1] + q[:, :1] * g_qd[:, 1:] + np.cross(v, g_qd[:, 1:])
    g_q = g_qt.copy()
    g_q[:, 0] += _dot(om1, g_qd[:, 1:])
    g_q[:, 1:] += -om1 * g_qd[:, :1] + np.cross(g_qd[:, 1:], om1)
    # om1 = Iinv L1, Iinv = R A R^T with R = R(q / |q|)
    g_L1 = gL1 + np.einsum("nji,nj->ni", Iinv, g_om)
    G = np.einsum("ni,nj->nij", g_om, L1)
    R = c["R"]
    g_R = G @ R @ np.swapaxes(Ibinv, -1, -2) + np.swapaxes(G, -1, -2) @ R @ Ibinv
    qn = c["qn"]
    g_qn = np.einsum("nij,ijab,nb->na", g_R, _C, qn)
    g_q += (g_qn - qn * (qn * g_qn).sum(-1, keepdims=True)) / c["nq"][:, None]
    # L1 = L + T dt
    g_L = g_L1
    g_T = dt * g_L1
    # x1 = x + v1 dt; v1 = v + (F im + g) dt
    g_x = gx1
    g_v1 = gv1 + dt * gx1
    g_v = g_v1
    g_F = (inv_mass * dt)[:, None] * g_v1
    return g_x, g_q, g_v, g_L, g_F, g_T


# ------------------------------------------------------------------ step
def _pp_candidates(x, rad):
    """Pairs i < j whose bounding spheres meet (the reference's test, same
    doubles: dpos = x_j - x_i, rejected when dist2 > (r_i + r_j)^2)."""
    n = len(x)
    if n < 2:
        return np.zeros(0, dtype=np.int64), np.zeros(0, dtype=np.int64)
    if n <= 2048:
        ii, jj = np.triu_indices(n, 1)
    else:
        from scipy.spatial import cKDTree

        pr = cKDTree(x).query_pairs(2.0 * float(rad.max()) * (1 + 1e-9) + 1e-12, output_type="ndarray")
        pr = np.sort(pr, axis=1)
        ii, jj = pr[:, 0], pr[:, 1]
    dp = x[jj] - x[ii]
    d2 = dp[:, 0] * dp[:, 0] + dp[:, 1] * dp[:, 1] + dp[:, 2] * dp[:, 2]
    rs = rad[ii] + rad[jj]
    keep = ~(d2 > rs * rs)
    ii, jj = ii[keep], jj[keep]
    o = np.lexsort((jj, ii))
    return ii[o], jj[o]


def _wall_candidates(scene, x, w):
    """(particle, wall triangle) pairs passing the reference's sphere and
    box pre-tests (src/main.cpp:2318-2343), particle-major."""
    W = scene.walls[w]
    r = scene.radius
    dp = x - W.pos
    d2 = dp[:, 0] * dp[:, 0] + dp[:, 1] * dp[:, 1] + dp[:, 2] * dp[:, 2]
    rs = r + W.radius
    near = np.nonzero(~(d2 > rs * rs))[0]
    if len(near) == 0:
        return np.zeros((0, 2), dtype=np.int64)
    lo, hi = W.tris.min(axis=1), W.tris.max(axis=1)              # (T, 3)
    pmn = x[near] - r[near][:, None]
    pmx = x[near] + r[near][:, None]
    sep = ((pmx[:, None, :] < lo[None]) | (pmn[:, None, :] > hi[None])).any(-1)
    i_idx, t_idx = np.nonzero(~sep)
    return np.stack([near[i_idx], t_idx], axis=1)


def _forward_step(scene, k, e, st, tape=None):
    """One step; st = dict(x, q, v, om, L, ds (N, W, 3), act (N, W)).
    Appends the step's record to tape when given.  Returns the new state."""
    x, q, v, om, L = st["x"], st["q"], st["v"], st["om"], st["L"]
    N = len(x)
    ii, jj = _pp_candidates(x, scene.radius)
    rec = dict(ii=ii, jj=jj)
    contrib_i, contrib_f, contrib_t = [], [], []
    if len(ii):
        pair, Sc, xc0, dSc, dxc, n = _loop_geometry(scene.zctx, scene.mesh[ii], scene.mesh[jj], x[ii], q[ii], x[jj],
                                                    q[jj], scene.split)
        a, b = ii[pair], jj[pair]
        area = np.sqrt(_dot(Sc, Sc))           # |Sn|, as the loop record's area
        f = n * (0.5 * k * area)[:, None]
        c = xc0 * -1.5
        rA, rB = c - x[a], c - x[b]
        rec.update(pair=pair, Sc=Sc, xc0=xc0, dSc=dSc, dxc=dxc, a=a, b=b, f=f, rA=rA, rB=rB)
        # reference order: per loop F[a] -= f, F[b] += f (pairs ascending)
        idx = np.stack([a, b], 1).ravel()
        contrib_i.append(idx)
        contrib_f.append(np.stack([-f, f], 1).reshape(-1, 3))
        contrib_t.append(np.stack([_cross(rA, -f), _cross(rB, f)], 1).reshape(-1, 3))
    ds_new = np.zeros_like(st["ds"])
    act_new = np.zeros_like(st["act"])
    walls = []
    for w, W in enumerate(scene.walls):
        cand = _wall_candidates(scene, x, w)
        if len(cand) == 0:
            walls.append(None)
            continue
        bodies = np.unique(cand[:, 0])
        recs = scene.zctx.wall_contacts(scene.mesh[bodies], x[bodies], q[bodies], W.tris, W.normals, jac=True)
        row = np.searchsorted(bodies, cand[:, 0])
        r = recs[row, cand[:, 1]]
        hit = r[:, 0] > 0.5
        cand, r = cand[hit], r[hit]
        i = cand[:, 0]
        ds = np.where(st["act"][i, w][:, None], st["ds"][i, w], 0.0)
        # the spring of (i, w) is read and written triangle after triangle:
        # contacts of one particle come in triangle order, so process them
        # rank by rank (the k-th contact of every particle at once)
        rank = np.zeros(len(i), dtype=np.int64)
        if len(i):
            first = np.r_[True, i[1:] != i[:-1]]
            start = np.maximum.accumulate(np.where(first, np.arange(len(i)), 0))
            rank = np.arange(len(i)) - start
        outs = []
        cur = {}
        fw = np.zeros((len(i), 3))
        rw = np.zeros((len(i), 3))
        ds_in = np.zeros((len(i), 3))
        caches = []
        for kk in range(int(rank.max()) + 1 if len(i) else 0):
            sel = np.nonzero(rank == kk)[0]
            dsk = np.array([cur.get(int(i[s]), ds[s]) for s in sel]).reshape(-1, 3)
            f_k, rP_k, dso_k, cache = _wall_force(r[sel, 1], r[sel, 2:5], r[sel, 5:8], x[i[sel]], v[i[sel]], om[i[sel]],
                                                  dsk, scene.young[i[sel]], scene.eqr[i[sel]], scene.poisson[i[sel]],
                                                  scene.mass[i[sel]], np.minimum(scene.mu[i[sel]], W.mu), e, scene.dt)
            fw[sel], rw[sel], ds_in[sel] = f_k, rP_k, dsk
            for s_, d_ in zip(sel, dso_k):
                cur[int(i[s_])] = d_
            caches.append((sel, cache))
        for ip, d_ in cur.items():
            ds_new[ip, w] = d_
            act_new[ip, w] = True
        walls.append(dict(i=i, r=r, rank=rank, caches=caches, f=fw, rP=rw, ds_in=ds_in))
        contrib_i.append(i)
        contrib_f.append(fw)
        contrib_t.append(_cross(rw, fw))
    F = np.zeros((N, 3))
    T = np.zeros((N, 3))
    if contrib_i:
        idx = np.concatenate(contrib_i)
        # per particle: particle-particle contributions first, then walls
        np.add.at(F, idx, np.concatenate(contrib_f))
        np.add.at(T, idx, np.concatenate(contrib_t))
    x1, q1, v1, om1, L1, ic = _integrate(x, q, v, L, F, T, scene.inv_mass, scene.Ibinv, scene.gravity, scene.dt)
    if tape is not None:
        rec.update(st=st, walls=walls, T=T, ic=ic)
        tape.append(rec)
    return dict(x=x1, q=q1, v=v1, om=om1, L=L1, ds=ds_new, act=act_new)


def _backward_step(scene, k, e, rec, g):
    """Adjoint of one step: g holds the gradients of the step's outputs
    (x, q, v, om, L, ds); returns those of its inputs, and the parts of the
    gradients of k and e."""
    st = rec["st"]
    x, q, L = st["x"], st["q"], st["L"]
    N = len(x)
    g_x, g_q, g_v, g_L, g_F, g_T = _integrate_vjp(rec["ic"], q, L, rec["T"], scene.inv_mass, scene.Ibinv, scene.dt,
                                                  g["x"], g["q"], g["v"], g["om"], g["L"])
    g_om = np.zeros((N, 3))
    g_ds = np.zeros_like(st["ds"])
    g_k = g_e = 0.0
    g_xi = np.zeros((N, 6))         # rigid-motion gradients of the bodies (geometry)
    for w, wr in enumerate(rec["walls"]):
        if wr is None or len(wr["i"]) == 0:
            continue
        i, r = wr["i"], wr["r"]
        cur = {int(ip): g["ds"][ip, w].copy() for ip in np.unique(i)}   # gradient of each chain's last spring
        for sel, cache in reversed(wr["caches"]):
            ii = i[sel]
            g_f = g_F[ii]
            g_rP = _cross(wr["f"][sel], g_T[ii])
            g_f = g_f + _cross(g_T[ii], wr["rP"][sel])
            g_dso = np.array([cur[int(p)] for p in ii]).reshape(-1, 3)
            out = _wall_vjp(cache, g_f, g_rP, g_dso, scene.dt)
            for s_, p in enumerate(ii):
                cur[int(p)] = out["ds"][s_]
            np.add.at(g_x, ii, out["x"])
            np.add.at(g_v, ii, out["v"])
            np.add.at(g_om, ii, out["om"])
            g_e += out["e"]
            dA = r[sel, 8:14]
            dX = r[sel, 14:32].reshape(-1, 3, 6)
            np.add.at(g_xi, ii, out["area"][:, None] * dA + np.einsum("li,lij->lj", out["xc"], dX))
        for p, gd in cur.items():
            if st["act"][p, w]:
                g_ds[p, w] += gd
    if len(rec["ii"]) and "pair" in rec:
        a, b = rec["a"], rec["b"]
        out = _pp_vjp(rec["Sc"], rec["rA"], rec["rB"], rec["f"], k, g_F[a], g_F[b], g_T[a], g_T[b])
        g_k += out["k"]
        np.add.at(g_x, a, out["xA"])
        np.add.at(g_x, b, out["xB"])
        G = _geometry_vjp(out["Sc"], out["xc0"], rec["dSc"], rec["dxc"], rec["pair"], len(rec["ii"]))
        np.add.at(g_xi, rec["ii"], G[:, 0:6])
        np.add.at(g_xi, rec["jj"], G[:, 6:12])
    g_x = g_x + g_xi[:, 0:3]
    g_q = g_q + _quat_grad(q, g_xi[:, 3:6])
    return dict(x=g_x, q=g_q, v=g_v, om=g_om, L=g_L, ds=g_ds), g_k, g_e


def _state0(scene, x, q, v, om, L):
    N, W = len(x), len(scene.walls)
    return dict(x=x, q=q, v=v, om=om, L=L, ds=np.zeros((N, max(W, 1), 3)), act=np.zeros((N, max(W, 1)), dtype=bool))


class Trajectory(torch.autograd.Function):
    """steps time steps of the model from (x0, q0, v0, om0, L0) with
    parameters k and e; returns the final (x, q, v, om, L).  Backward: the
    hand-written adjoint, each segment of ckpt_every steps re-run from its
    checkpoint to record its tape."""

    @staticmethod
    def forward(ctx, k, e, x0, q0, v0, om0, L0, scene, steps, ckpt_every):
        kf, ef = float(k), float(e)
        st = _state0(scene, *(t.detach().cpu().numpy().astype(np.float64) for t in (x0, q0, v0, om0, L0)))
        ckpts = []
        for s in range(steps):
            if s % ckpt_every == 0:
                ckpts.append(st)
            st = _forward_step(scene, kf, ef, st)
        ctx.scene, ctx.steps, ctx.every, ctx.ckpts, ctx.k, ctx.e = scene, steps, ckpt_every, ckpts, kf, ef
        ctx.dev, ctx.dtype = x0.device, x0.dtype
        return tuple(torch.from_numpy(st[n]).to(x0) for n in ("x", "q", "v", "om", "L"))

    @staticmethod
    def backward(ctx, gx, gq, gv, gom, gL):
        scene = ctx.scene
        cv = lambda t, like: np.zeros_like(like) if t is None else t.detach().cpu().numpy().astype(np.float64)
        st0 = ctx.ckpts[0]
        g = dict(x=cv(gx, st0["x"]), q=cv(gq, st0["q"]), v=cv(gv, st0["v"]), om=cv(gom, st0["om"]),
                 L=cv(gL, st0["L"]), ds=np.zeros_like(st0["ds"]))
        g_k = g_e = 0.0
        for seg in range(len(ctx.ckpts) - 1, -1, -1):
            start = seg * ctx.every
            m = min(ctx.every, ctx.steps - start)
            tape = []
            st = ctx.ckpts[seg]
            for _ in range(m):
                st = _forward_step(scene, ctx.k, ctx.e, st, tape)
            for rec in reversed(tape):
                g, gk, ge = _backward_step(scene, ctx.k, ctx.e, rec, g)
                g_k += gk
                g_e += ge
        t = lambda a: torch.from_numpy(a).to(device=ctx.dev, dtype=ctx.dtype)
        return (torch.tensor(g_k, dtype=ctx.dtype), torch.tensor(g_e, dtype=ctx.dtype), t(g["x"]), t(g["q"]),
                t(g["v"]), t(g["om"]), t(g["L"]), None, None, None)


def simulate(scene: Scene, k, e, pos, quat, vel, omega, L, steps: int, ckpt_every: int = 50):
    """`steps` steps of the model; differentiable in k, e and the initial
    state through the hand-written adjoint (Trajectory), with checkpoints
    every ckpt_every steps (memory O(steps / ckpt_every) states plus one
    segment's tape).  Returns the final (pos, quat, vel, omega, L)."""
    as_t = lambda a: a if torch.is_tensor(a) else torch.tensor(float(a), dtype=torch.float64)
    return Trajectory.apply(as_t(k), as_t(e), pos, quat, vel, omega, L, scene, int(steps), int(ckpt_every))


def step(scene: Scene, k, e, pos, quat, vel, omega, L):
    """One step (float64 tensors in and out), through the same adjoint."""
    return simulate(scene, k, e, pos, quat, vel, omega, L, 1, 1)


def scene_from_workload(zctx, wl):
    """Scene of a workload (zdx.workloads), with its walls as the reference
    builds them.  Returns (scene, pos0, quat0, vel0) with float64 tensors."""
    N = wl.n
    mesh = np.zeros(N, dtype=np.int64)
    rad, mass, eqr = np.zeros(N), np.zeros(N), np.zeros(N)
    Ibinv = np.zeros((N, 3, 3))
    cache = {}
    for i in range(N):
        key = (int(wl.tpl[i]), float(wl.scale[i]))
        if key not in cache:
            mid = zctx.register_mesh(wl.templates[key[0]] * key[1])
            cache[key] = (mid, zctx.mesh_props[mid])   # volume, centroid(3), radius, mean edge, bbox, inertia(9)
        mid, props = cache[key]
        dens = float(wl.material[i, 0])
        V, I = float(props[0]), props[7:16].reshape(3, 3) * dens
        mesh[i] = mid
        rad[i] = float(props[4])
        mass[i] = dens * V
        Ibinv[i] = body_inverse_inertia(I)
        eqr[i] = np.cbrt(3.0 * V / (4.0 * math.pi))
    walls = [make_wall(zctx, tris, pos, quat, mu) for (tris, pos, quat, mu, _rest, _name) in wl.walls]
    inv_mass = np.where(mass > 0.0, 1.0 / np.where(mass > 0.0, mass, 1.0), 0.0)
    scene = Scene(zctx=zctx, mesh=mesh, radius=rad, mass=mass, inv_mass=inv_mass, Ibinv=Ibinv, eqr=eqr,
                  young=wl.material[:, 1].copy(), poisson=wl.material[:, 2].copy(), mu=wl.material[:, 3].copy(),
                  walls=walls, gravity=np.asarray(wl.gravity, dtype=np.float64), dt=float(wl.dt), split=bool(wl.split))
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64))
    return scene, t(wl.init[:, 0:3]), t(wl.init[:, 3:7]), t(wl.init[:, 7:10])


def calibrate(scene, pos0, quat0, target, k0, e0, v0, steps, iters=20, v_scale=0.1, ckpt_every=50, log=None):
    """Fits the stiffness k, the restitution e and the initial velocities v0 so
    that the positions after `steps` steps match `target` (loss = sum of
    squared distances), by L-BFGS with a strong-Wolfe line search
    (torch.optim.LBFGS; at most `iters` loss and gradient evaluations) in the
    scaled coordinates theta = (log k, logit e, v0 / v_scale), on the loss
    divided by its starting value, gradients from the adjoint.  A contact
    trajectory's loss is smooth only piecewise (contacts start and end), so
    the search settles where the pieces meet; the loss it reaches is the
    measure of the fit.  Returns the history [(loss, k, e)] of the
    evaluations (the start first) and the fitted k, e and v0."""
    n = pos0.shape[0]
    z = torch.zeros(n, 3, dtype=torch.float64)

    def unpack(th):
        return th[0].exp(), torch.sigmoid(th[1]), th[2:].reshape(n, 3) * v_scale

    def loss_of(th):
        k, e, v = unpack(th)
        p = simulate(scene, k, e, pos0, quat0, v, z, z, steps, ckpt_every)[0]
        return ((p - target) ** 2).sum()

    th = torch.cat([torch.tensor([math.log(float(k0)), math.log(float(e0) / (1.0 - float(e0)))], dtype=torch.float64),
                    (v0.detach().reshape(-1) / v_scale).to(torch.float64)]).requires_grad_(True)
    with torch.no_grad():
        L0 = float(loss_of(th))
    hist = []
    if L0 == 0.0:
        k, e, v = unpack(th.detach())
        return [(0.0, float(k), float(e))], float(k), float(e), v
    opt = torch.optim.LBFGS([th], lr=1.0, max_iter=iters, max_eval=iters, history_size=10,
                            line_search_fn="strong_wolfe", tolerance_grad=1e-12, tolerance_change=1e-14)

    def closure():
        opt.zero_grad()
        L = loss_of(th) / L0
        L.backward()
        k, e, _ = unpack(th.detach())
        hist.append((float(L.detach()) * L0, float(k), float(e)))
        if log:
            log(hist[-1])
        return L

    opt.step(closure)
    k, e, v = unpack(th.detach())
    return hist, float(k), float(e), v
