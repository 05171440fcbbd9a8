"""CPU: the torch model of the time step with the minimal-energy force
(zdx.autograd.step, the forward that the adjoint differentiates) against the
oracle's time step with ``force_model minimal`` on 8 overlapping hulls; the
per-pair contact call is served by the oracle here (forward only), so this
pins the force, damping, lever arm and integrator of the torch model and of
the oracle's step against each other."""
import os

import numpy as np
import pytest

import golden_inputs as GI
import oracle as O
from zdx import synth

torch = pytest.importorskip("torch")


class OracleContacts:
    """Stands in for zdx.Context in zdx.autograd: meshes and contact loops from
    the oracle (no loop Jacobians)."""

    def __init__(self):
        self.meshes, self.mesh_props = {}, {}

    def register_mesh(self, tris, center=True):
        t, props, _ = O.build_mesh(np.asarray(tris, float), center)
        mid = len(self.meshes)
        self.meshes[mid] = (t, props)
        self.mesh_props[mid] = props
        return mid

    def contact_pairs(self, ma, mb, pa, qa, pb, qb, split=True, k=1.0, vertices=False, sensitivities=False):
        recs, lo = [], [0]
        for i in range(len(ma)):
            ta, pra = self.meshes[int(ma[i])]
            tb, _ = self.meshes[int(mb[i])]
            r = O.pair_contact(ta, tb, pa[i], qa[i], pb[i], qb[i], pra[5], pra[6], split=split, k=k)
            recs.append(r["loops"])
            lo.append(lo[-1] + len(r["loops"]))
        rec = np.concatenate(recs) if recs else np.zeros((0, 28))
        return dict(loops=rec, loop_offset=np.array(lo), loop_jac=np.zeros((len(rec), 72)))


def test_torch_step_matches_oracle_minimal_step(tmp_path):
    from zdx.autograd import Scene, body_inverse_inertia, simulate

    zc = OracleContacts()
    hull = synth.sphere_hull(128, 7) * 0.005
    mid = zc.register_mesh(hull)
    props = zc.mesh_props[mid]
    dens, e, k, n, steps = 2500.0, 0.5, 4e6, 8, 100
    rng = np.random.default_rng(11)
    pos = np.array([[(i & 1) * 0.0094, ((i >> 1) & 1) * 0.0094, (i >> 2) * 0.0094] for i in range(n)])
    pos += rng.uniform(-2e-4, 2e-4, size=pos.shape)
    quat = rng.normal(size=(n, 4))
    quat /= np.linalg.norm(quat, axis=1, keepdims=True)
    vel = rng.normal(scale=0.05, size=(n, 3))
    om = rng.normal(scale=2.0, size=(n, 3))
    parts = [GI._particle("h.stl", pos[i], vel[i], quat[i], om[i], density=dens, rest=e) for i in range(n)]
    g = dict(steps=steps, dt=1e-5, split_contacts=1, gravity=[0.0, 0.0, 0.0], center_mesh=1, vtk_prefix="h",
             output_dir="out", output_interval=steps, force_model="minimal", minimal_k=k)
    synth.write_stl_ascii(str(tmp_path / "h.stl"), hull)
    (tmp_path / "c.txt").write_text(synth.format_config(g, parts))
    osim = O.Sim(str(tmp_path / "c.txt"), str(tmp_path))
    s0 = osim.state()
    contacts = sum(int(osim.step(1)[0]) for _ in range(steps))
    so = osim.state()
    assert contacts > 50

    V, I = float(props[0]), np.asarray(props[7:16]).reshape(3, 3) * dens
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64))
    scene = Scene(zctx=zc, mesh=np.full(n, mid), radius=np.full(n, float(props[4])), mass=t(np.full(n, dens * V)),
                  Ibinv=t(np.broadcast_to(body_inverse_inertia(I), (n, 3, 3)).copy()),
                  req=t(np.full(n, np.cbrt(3.0 * V / (4.0 * np.pi)))), floor_mesh=None, floor_pos=None,
                  floor_top=0.0, gravity=t([0.0, 0.0, 0.0]), dt=1e-5, split=True)
    with torch.no_grad():
        st = simulate(scene, t(k), t(e), t(s0[:, 0:3]), t(s0[:, 3:7]), t(s0[:, 7:10]), t(s0[:, 10:13]),
                      t(s0[:, 13:16]), steps, 50)
    st = np.concatenate([x.numpy() for x in st], axis=1)
    err = np.abs(st - so) / np.maximum(np.abs(so), 1e-3)
    assert np.max(err) <= 1e-12, np.max(err)
    assert np.max(np.abs(so[:, 7:10] - s0[:, 7:10])) > 1e-3


def test_body_inverse_inertia_follows_the_reference_threshold():
    from zdx.autograd import body_inverse_inertia

    big = np.diag([2.0, 3.0, 4.0])
    assert np.allclose(body_inverse_inertia(big), np.diag([0.5, 1 / 3, 0.25]))
    small = np.diag([1e-8, 2e-8, 3e-8])                    # det 6e-24 < 1e-18: identity
    assert np.array_equal(body_inverse_inertia(small), np.eye(3))
