"""Differentiable contact step of the minimal-energy model (polyline.py:655-701)
on the device contact pipeline, for calibration.

* ``PairContact`` (a ``torch.autograd.Function``): per contact loop of a batch
  of pairs, the canonical area vector ``Sc = -|Sn| n`` (n: unit normal from A
  to B) and the contact point ``xc0`` (src/main.cpp:1042-1052).  Forward is
  ``zdx_contact_pairs``; backward is the hand-written adjoint: the device
  returns, per loop, dSn/dxi and dGn/dxi for rigid motions xi of the two
  bodies (each loop vertex is the mean of its snapped segment endpoints, each
  endpoint an edge of one body crossing a triangle plane of the other), and
  the chain to Sc, xc0 and to positions/quaternions is done here.  The
  loop topology is held fixed, as a derivative must.
* ``step``: forces from loops -- elastic ``F_B = k/2 |Sn| n`` (the
  energy identity f = -dPsi/dlambda n with Psi = k/2 Omega, dOmega/dlambda =
  |Sn|, polyline.py:655-701) plus normal damping ``-c (v_rel . n) n`` with
  ``c = 2 zeta(e) sqrt(m_eff k_c)``, ``zeta(e) = -ln e / sqrt(pi^2 + ln^2 e)``
  and the contact stiffness of a sphere of the effective radius
  R = R1 R2 / (R1 + R2), ``k_c = pi k R``, applied at the loop centroid
  (``-3/2 xc0`` up to a component along n; polyline.py's minimal model has no
  torques, and the world-frame ``xc0`` itself lies far from the loop for
  bodies away from the origin) -- and the rigid-body update of
  ``integrate_particle`` (src/main.cpp:1494-1510, omega from the old rotation,
  carried as state) in torch.  Without a floor this is the device time step
  with ``zdx_sim_set_force_model(ZDX_FORCE_MINIMAL, k)``, term for term.
* ``simulate``: K-step segments under ``torch.utils.checkpoint`` so memory is
  O(N steps / K) states, the contact calls of a segment being re-run in
  backward (they are deterministic).

The static floor is a registered mesh with infinite mass (a fixed body B).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch
from torch.utils.checkpoint import checkpoint


def _skew(v):
    z = np.zeros(v.shape[:-1])
    return np.stack([np.stack([z, -v[..., 2], v[..., 1]], -1), np.stack([v[..., 2], z, -v[..., 0]], -1),
                     np.stack([-v[..., 1], v[..., 0], z], -1)], -2)


def _quat_grad(q, g_w):
    """dL/dq for the world-frame rotation increment w (q' = exp(w/2) q, q
    normalised inside the rotation): 2 ((0, g_w) x q_hat) / |q|."""
    nq = np.linalg.norm(q, axis=-1, keepdims=True)
    qh = q / nq
    w, v = qh[..., :1], qh[..., 1:]
    gw = -(g_w * v).sum(-1, keepdims=True)
    gv = w * g_w + np.cross(g_w, v)
    return 2.0 * np.concatenate([gw, gv], -1) / nq


class PairContact(torch.autograd.Function):
    @staticmethod
    def forward(ctx, posA, quatA, posB, quatB, zctx, mesh_a, mesh_b, split):
        r = zctx.contact_pairs(mesh_a, mesh_b, posA.detach().cpu().numpy(), quatA.detach().cpu().numpy(),
                               posB.detach().cpu().numpy(), quatB.detach().cpu().numpy(), split=split, k=1.0,
                               vertices=False, sensitivities=True)
        rec, jac, lo = r["loops"], r["loop_jac"], r["loop_offset"]
        nl = len(rec)
        pair = np.repeat(np.arange(len(mesh_a)), np.diff(lo)).astype(np.int64)
        Sn, Gn, n, xc, area = rec[:, 0:3], rec[:, 3:6], rec[:, 6:9], rec[:, 9:12], rec[:, 12]
        s = np.sign((n * (-Sn)).sum(-1)) if nl else np.zeros(0)   # n = s (-Sn/|Sn|)
        Sc = s[:, None] * Sn
        dSn = jac[:, 0:36].reshape(nl, 3, 12)
        dGn = jac[:, 36:72].reshape(nl, 3, 12)
        # xc = -(Sn x Gn) / |Sn|^2
        a2 = (area ** 2)[:, None, None]
        c = np.cross(Sn, Gn)
        dc = -_skew(Gn) @ dSn + _skew(Sn) @ dGn
        dxc = -dc / a2 + 2.0 * c[:, :, None] * np.einsum("li,lij->lj", Sn, dSn)[:, None, :] / (a2 * a2)
        ctx.save_for_backward(torch.from_numpy(s[:, None, None] * dSn), torch.from_numpy(dxc),
                              torch.from_numpy(pair), quatA.detach().cpu(), quatB.detach().cpu())
        ctx.n_pairs = len(mesh_a)
        ctx.dev = posA.device
        lp = torch.from_numpy(pair).to(posA.device)
        ctx.mark_non_differentiable(lp)
        return torch.from_numpy(Sc).to(posA), torch.from_numpy(xc.copy()).to(posA), lp

    @staticmethod
    def backward(ctx, gSc, gxc, _gpair):
        dSc, dxc, pair, qA, qB = ctx.saved_tensors
        npair = ctx.n_pairs
        gxi = np.einsum("li,lij->lj", gSc.detach().cpu().numpy(), dSc.numpy())
        gxi += np.einsum("li,lij->lj", gxc.detach().cpu().numpy(), dxc.numpy())
        G = np.zeros((npair, 12))
        np.add.at(G, pair.numpy(), gxi)
        gpA = torch.from_numpy(G[:, 0:3]).to(ctx.dev)
        gqA = torch.from_numpy(_quat_grad(qA.numpy(), G[:, 3:6])).to(ctx.dev)
        gpB = torch.from_numpy(G[:, 6:9]).to(ctx.dev)
        gqB = torch.from_numpy(_quat_grad(qB.numpy(), G[:, 9:12])).to(ctx.dev)
        return gpA, gqA, gpB, gqB, None, None, None, None


def qmul(a, b):
    w1, x1, y1, z1 = a.unbind(-1)
    w2, x2, y2, z2 = b.unbind(-1)
    return torch.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                        w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], -1)


def qmat(q):
    q = q / q.norm(dim=-1, keepdim=True)
    w, x, y, z = q.unbind(-1)
    return torch.stack([torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
                        torch.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
                        torch.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], -2)


def body_inverse_inertia(I: np.ndarray) -> np.ndarray:
    """Inverse body inertia as the time step uses it: the reference's
    mat3_inverse (src/main.cpp:203-207) returns the identity when
    |det I| < 1e-18, which holds for every particle below about a centimetre
    at mineral densities (I ~ 1e-8 kg m^2), so their omega equals L."""
    I = np.asarray(I, dtype=np.float64)
    det = np.linalg.det(I)
    out = np.linalg.inv(np.where((np.abs(det) < 1e-18)[..., None, None], np.eye(3), I))
    return np.where((np.abs(det) < 1e-18)[..., None, None], np.eye(3), out)


@dataclass
class Scene:
    """Particles (registered meshes), optionally above a static floor mesh."""
    zctx: object
    mesh: np.ndarray            # per particle mesh id
    radius: np.ndarray          # bounding radius per particle (candidate test)
    mass: torch.Tensor          # (N,)
    Ibinv: torch.Tensor         # (N, 3, 3) body-frame inverse inertia
    req: torch.Tensor           # (N,) equivalent-sphere radius (cbrt(3V/4pi))
    floor_mesh: int | None
    floor_pos: np.ndarray | None
    floor_top: float            # z of the floor's top face
    gravity: torch.Tensor
    dt: float
    split: bool = True


def _candidates(scene: Scene, pos: np.ndarray):
    """Particle pairs whose bounding spheres overlap (the reference's broad
    phase test, src/main.cpp:2049-2059) and particles whose sphere reaches the
    floor."""
    d = pos[None, :, :] - pos[:, None, :]
    rs = scene.radius[:, None] + scene.radius[None, :]
    ii, jj = np.nonzero(np.triu((d * d).sum(-1) <= rs * rs, 1))
    if scene.floor_mesh is None:
        return ii, jj, np.zeros(0, dtype=np.int64)
    fl = np.nonzero(pos[:, 2] - scene.radius <= scene.floor_top)[0]
    return ii, jj, fl


def step(scene: Scene, k, e, pos, quat, vel, omega, L):
    """One step of the minimal model (the device step with the minimal-energy
    force, src/main.cpp:2034-2468 with polyline.py:655-701's force); all
    tensors float64, differentiable in k, e and the state (pos, quat, vel,
    omega, L).  Returns the new state."""
    ii, jj, fl = _candidates(scene, pos.detach().cpu().numpy())
    ia = np.concatenate([ii, fl]).astype(np.int64)
    ib = np.concatenate([jj, np.full(len(fl), -1)]).astype(np.int64)
    F = torch.zeros_like(pos)
    T = torch.zeros_like(pos)
    if len(ia):
        ident = torch.tensor([1.0, 0.0, 0.0, 0.0], dtype=pos.dtype, device=pos.device)
        fpos = torch.tensor(scene.floor_pos if scene.floor_mesh is not None else np.zeros(3), dtype=pos.dtype,
                            device=pos.device)
        tia = torch.from_numpy(ia).to(pos.device)
        tib = torch.from_numpy(np.where(ib >= 0, ib, 0)).to(pos.device)
        isfloor = torch.from_numpy(ib < 0).to(pos.device)
        pA, qA = pos[tia], quat[tia]
        pB = torch.where(isfloor[:, None], fpos.expand(len(ia), 3), pos[tib])
        qB = torch.where(isfloor[:, None], ident.expand(len(ia), 4), quat[tib])
        mb = np.where(ib >= 0, scene.mesh[np.where(ib >= 0, ib, 0)], -1 if scene.floor_mesh is None else
                      scene.floor_mesh)
        Sc, xc, lp = PairContact.apply(pA, qA, pB, qB, scene.zctx, scene.mesh[ia], mb, scene.split)
        if Sc.shape[0]:
            lp = lp.long()
            a = tia[lp]
            b = tib[lp]
            fb = isfloor[lp]
            area = Sc.norm(dim=-1, keepdim=True)
            nrm = -Sc / area
            # xc0 = (nA x Gn)/|Sn| = -2/3 of the loop centroid's part normal to n
            # (Stokes on Gn); -3/2 xc0 is the centroid up to a multiple of n,
            # which neither the torque of a normal force nor v_rel . n sees.
            xc = -1.5 * xc
            rA = xc - pos[a]
            va = vel[a] + torch.cross(omega[a], rA, dim=-1)
            rB = xc - torch.where(fb[:, None], fpos.expand_as(xc), pos[b])
            vb = torch.where(fb[:, None], torch.zeros_like(xc), vel[b] + torch.cross(omega[b], rB, dim=-1))
            vn = ((vb - va) * nrm).sum(-1, keepdim=True)
            ma = scene.mass[a][:, None]
            mbm = torch.where(fb[:, None], ma, scene.mass[b][:, None])
            meff = torch.where(fb[:, None], ma, ma * mbm / (ma + mbm))
            ra, rb = scene.req[a], scene.req[b]
            req = torch.where(fb, ra, ra * rb / (ra + rb))[:, None]
            kc = math.pi * k * req
            le = torch.log(e)
            zeta = -le / torch.sqrt(math.pi ** 2 + le * le)
            c = 2.0 * zeta * torch.sqrt(meff * kc)
            fB = (0.5 * k * area - c * vn) * nrm          # force on B
            F = F.index_add(0, a, -fB)
            T = T.index_add(0, a, torch.cross(rA, -fB, dim=-1))
            keep = ~fb
            F = F.index_add(0, b[keep], fB[keep])
            T = T.index_add(0, b[keep], torch.cross(rB[keep], fB[keep], dim=-1))
    # integrate_particle, src/main.cpp:1494-1510 (omega from the old rotation)
    vel = vel + (F / scene.mass[:, None] + scene.gravity) * scene.dt
    pos = pos + vel * scene.dt
    R = qmat(quat)
    Iinv = R @ scene.Ibinv @ R.transpose(-1, -2)
    L = L + T * scene.dt
    omega = (Iinv @ L.unsqueeze(-1)).squeeze(-1)
    dq = qmul(torch.cat([torch.zeros_like(omega[:, :1]), omega], -1), quat)
    quat = quat + 0.5 * dq * scene.dt
    quat = quat / quat.norm(dim=-1, keepdim=True)
    return pos, quat, vel, omega, L


def simulate(scene: Scene, k, e, pos, quat, vel, omega, L, steps: int, ckpt_every: int = 50):
    """`steps` steps; every `ckpt_every` steps form one checkpointed segment.
    Returns the final (pos, quat, vel, omega, L)."""
    s = 0
    state = (pos, quat, vel, omega, L)
    while s < steps:
        m = min(ckpt_every, steps - s)

        def seg(k_, e_, *st, m=m):
            for _ in range(m):
                st = step(scene, k_, e_, *st)
            return st

        if torch.is_grad_enabled():
            state = checkpoint(seg, k, e, *state, use_reentrant=False)
        else:
            state = seg(k, e, *state)
        s += m
    return state


def scene_from_workload(zctx, wl, floor_thickness: float = 0.02):
    """Scene of a drop-on-plane workload (zdx.workloads.drop_on_plane): the
    particles' meshes (one per template and scale), their mass properties, and
    a floor slab whose top face is the workload's plane wall (z = 0).
    Returns (scene, pos0, quat0, vel0) as float64 tensors."""
    from . import synth

    dens = float(wl.material[0, 0])
    mesh = np.zeros(wl.n, dtype=np.int64)
    rad = np.zeros(wl.n)
    mass = np.zeros(wl.n)
    Ibinv = np.zeros((wl.n, 3, 3))
    req = np.zeros(wl.n)
    cache = {}
    for i in range(wl.n):
        key = (int(wl.tpl[i]), float(wl.scale[i]))
        if key not in cache:
            tris = wl.templates[key[0]] * key[1]
            mid = zctx.register_mesh(tris)
            cache[key] = (mid, zctx.mesh_props[mid])   # volume, centroid(3), radius, mean edge, bbox, inertia(9)
        mid, props = cache[key]
        V, I = float(props[0]), props[7:16].reshape(3, 3) * dens
        mesh[i] = mid
        rad[i] = float(props[4])
        mass[i] = dens * V
        Ibinv[i] = body_inverse_inertia(I)
        req[i] = (3.0 * V / (4.0 * math.pi)) ** (1.0 / 3.0)
    tris, wpos, _, _, _, _ = wl.walls[0]
    ext = np.ptp(tris.reshape(-1, 3), axis=0)
    floor = synth.box_wall(float(ext[0]), float(ext[1]), floor_thickness)
    fid = zctx.register_mesh(floor)
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64))
    scene = Scene(zctx=zctx, mesh=mesh, radius=rad, mass=t(mass), Ibinv=t(Ibinv), req=t(req), floor_mesh=fid,
                  floor_pos=np.array([wpos[0], wpos[1], wpos[2] - 0.5 * floor_thickness]), floor_top=float(wpos[2]),
                  gravity=t(wl.gravity), dt=float(wl.dt), split=bool(wl.split))
    return scene, t(wl.init[:, 0:3]), t(wl.init[:, 3:7]), t(wl.init[:, 7:10])


def calibrate(scene, pos0, quat0, target, k0, e0, v0, steps, iters=10, lr=0.05, lr_v=None, ckpt_every=50,
              log=None):
    """Fits the stiffness k, the restitution e and the initial velocities v0 so
    that the positions after `steps` steps match `target` (loss = sum of
    squared distances), by Adam on (log k, logit e, v0) with gradients from
    the adjoint.  Adam's first steps move every coordinate by about its
    learning rate whatever the gradient's size, so v0 gets its own rate
    `lr_v` in m/s (default: `lr` times the mean speed of v0, at least 1e-3 m/s).
    Returns the history [(loss, k, e)] and the fitted k, e and v0."""
    lk = torch.tensor(math.log(float(k0)), dtype=torch.float64, requires_grad=True)
    le = torch.tensor(math.log(float(e0) / (1.0 - float(e0))), dtype=torch.float64, requires_grad=True)
    v = v0.clone().detach().requires_grad_(True)
    if lr_v is None:
        lr_v = lr * max(float(v0.norm(dim=-1).mean()), 1e-3 / lr)
    opt = torch.optim.Adam([{"params": [lk, le], "lr": lr}, {"params": [v], "lr": lr_v}])
    n = pos0.shape[0]
    hist = []
    for _ in range(iters):
        opt.zero_grad()
        k = lk.exp()
        e = torch.sigmoid(le)
        z = torch.zeros(n, 3, dtype=torch.float64)
        p = simulate(scene, k, e, pos0, quat0, v, z, z, steps, ckpt_every)[0]
        loss = ((p - target) ** 2).sum()
        loss.backward()
        hist.append((float(loss.detach()), float(k.detach()), float(e.detach())))
        if log:
            log(hist[-1])
        opt.step()
    return hist, float(lk.detach().exp()), float(torch.sigmoid(le.detach())), v.detach()
