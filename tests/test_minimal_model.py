"""CPU: the differentiable model's forward (zdx.autograd: the time step with
the minimal-energy particle-particle force and the reference's wall force)
against the oracle's time step with ``force_model minimal`` on 8 hulls
dropped onto a plane wall, colliding with each other and the wall.  The
contact geometry (loops, wall sections) is served by the oracle here, so this
pins the model's force laws, summation order and integrator against the
oracle's step, whose wall force is pinned bit-exactly by the compiled
reference (the wall_drop golden simulation) and whose minimal force by the
box-on-box closed form."""
import numpy as np
import pytest

import golden_inputs as GI
import oracle as O
from zdx import synth

torch = pytest.importorskip("torch")


class OracleContacts:
    """Stands in for zdx.Context in zdx.autograd: meshes, contact loops and
    wall sections from the oracle (no derivatives)."""

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

    def wall_contacts(self, mesh, pos, quat, wall_tris, wall_normals, jac=False):
        L = O.lib()
        out = np.zeros((len(mesh), len(wall_tris), 32))
        for b in range(len(mesh)):
            t, props = self.meshes[int(mesh[b])]
            tol = props[5] * 0.1 if props[5] > 0 else (props[6] * 1e-2 if props[6] > 0 else 1e-6)
            for w in range(len(wall_tris)):
                a, xc, n = np.zeros(1), np.zeros(3), np.zeros(3)
                if L.orc_wall_contact(t.reshape(-1), len(t), np.asarray(pos[b], float), np.asarray(quat[b], float),
                                      np.ascontiguousarray(wall_tris[w], dtype=float).reshape(-1),
                                      np.asarray(wall_normals[w], float), tol, a, xc, n):
                    if np.dot(n, pos[b] - xc) < 0.0:
                        n = -n
                    out[b, w, 0], out[b, w, 1], out[b, w, 2:5], out[b, w, 5:8] = 1.0, a[0], xc, n
        return out


def _drop_case(tmp_path, n=8, steps=200, k=4e6, e=0.5):
    hull = synth.sphere_hull(128, 7) * 0.005
    floor = synth.plane_wall(0.2, 0.2)
    rng = np.random.default_rng(11)
    pos = np.array([[(i & 1) * 0.0101, ((i >> 1) & 1) * 0.0101, 0.0051 + (i >> 2) * 0.0101] for i in range(n)])
    pos += rng.uniform(-1e-4, 1e-4, size=pos.shape)
    quat = rng.normal(size=(n, 4))
    quat /= np.linalg.norm(quat, axis=1, keepdims=True)
    vel = rng.normal(scale=0.05, size=(n, 3))
    vel[:, 2] -= 0.3
    om = rng.normal(scale=2.0, size=(n, 3))
    parts = [GI._particle("h.stl", pos[i], vel[i], quat[i], om[i], density=2500.0, rest=e) for i in range(n)]
    walls = [dict(stl="floor.stl", pos=[0.005, 0.005, 0.0], quat=[1.0, 0.0, 0.0, 0.0], scale=1.0, mu=0.5,
                  restitution=e)]
    g = dict(steps=steps, dt=1e-5, split_contacts=1, gravity=[0.0, 0.0, -9.81], center_mesh=1, vtk_prefix="h",
             output_dir="out", output_interval=steps, force_model="minimal", minimal_k=k)
    synth.write_stl_ascii(str(tmp_path / "h.stl"), hull)
    synth.write_stl_ascii(str(tmp_path / "floor.stl"), floor)
    (tmp_path / "c.txt").write_text(synth.format_config(g, parts, walls))
    return hull, floor, str(tmp_path / "c.txt")


def test_model_forward_matches_oracle_minimal_step_with_wall(tmp_path):
    from zdx.autograd import Scene, body_inverse_inertia, make_wall, simulate

    n, steps, k, e = 8, 200, 4e6, 0.5
    hull, floor, cfg = _drop_case(tmp_path, n, steps, k, e)
    osim = O.Sim(cfg, str(tmp_path))
    s0 = osim.state()
    ctr = np.zeros(7, dtype=np.int64)
    pp = wall = 0
    for _ in range(steps):
        c = osim.step(1)
        pp += int(c[0])
    so = osim.state()
    assert pp > 20

    zc = OracleContacts()
    mid = zc.register_mesh(hull)
    props = zc.mesh_props[mid]
    dens = 2500.0
    V, I = float(props[0]), np.asarray(props[7:16]).reshape(3, 3) * dens
    W = make_wall(zc, floor, [0.005, 0.005, 0.0], [1.0, 0.0, 0.0, 0.0], 0.5)
    scene = Scene(zctx=zc, mesh=np.full(n, mid), radius=np.full(n, float(props[4])), mass=np.full(n, dens * V),
                  inv_mass=np.full(n, 1.0 / (dens * V)), Ibinv=np.broadcast_to(body_inverse_inertia(I), (n, 3, 3)).copy(),
                  eqr=np.full(n, np.cbrt(3.0 * V / (4.0 * 3.141592653589793))), young=np.full(n, 1e7),
                  poisson=np.full(n, 0.25), mu=np.full(n, 0.5), walls=[W], gravity=np.array([0.0, 0.0, -9.81]),
                  dt=1e-5, split=True)
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64))
    with torch.no_grad():
        st = simulate(scene, t(k), t(e), t(s0[:, 0:3]), t(s0[:, 3:7]), t(s0[:, 7:10]), t(s0[:, 10:13]),
                      t(s0[:, 13:16]), steps, 50)
    st = np.concatenate([x.numpy() for x in st], axis=1)
    # the bodies hit the wall: their vertical velocity turned round
    assert (so[:, 9] > s0[:, 9] + 0.1).sum() >= 4
    err = np.abs(st - so) / np.maximum(np.abs(so), 1e-3)
    assert np.max(err) <= 1e-12, np.max(err)


def test_body_inverse_inertia_follows_the_reference_threshold():
    from zdx.autograd import body_inverse_inertia

    big = np.diag([2.0, 3.0, 4.0])
    assert np.allclose(body_inverse_inertia(big), np.diag([0.5, 1 / 3, 0.25]))
    small = np.diag([1e-8, 2e-8, 3e-8])                    # det 6e-24 < 1e-18: identity
    assert np.array_equal(body_inverse_inertia(small), np.eye(3))
