"""Closed forms and invariants that pin the oracle where the reference's own
outputs leave room (and that a plausible slip -- a dropped term, a sign, a
transposed operand -- would fail):

* box-on-box: two boxes whose side faces cross: the contact loop is the
  rectangle where B's side faces cut A's face, so |Sn| equals the overlap
  area A, Fn = k/2 |Sn| = k/2 dOmega/d(delta) with Omega = delta A (the energy
  identity f = -dPsi/dlambda n of the minimal model, polyline.py:655-668), n
  is the face normal pointing from A to B, and xc0 = -2/3 of the rectangle
  centre's in-plane part (src/main.cpp:1042-1052);
* Sn of a closed loop is invariant under a common translation of both bodies
  and rotates with a common rotation;
* planar polygon: Sn = area * normal, Gn = -2/3 area * normal x centroid;
* mass properties of a box and of a tetrahedron (volume, centroid, inertia);
* torque-free rigid rotation keeps the angular momentum L unchanged.
"""
import numpy as np
import pytest

import oracle as O
from zdx import synth


def _box(sx, sy, sz):
    return synth.box_wall(sx, sy, sz)


def _rot(q):
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _pair(tA, tB, pA, qA, pB, qB, k=3.0):
    ta, pa, _ = O.build_mesh(tA)
    tb, _, _ = O.build_mesh(tB)
    return O.pair_contact(ta, tb, np.asarray(pA, float), np.asarray(qA, float), np.asarray(pB, float),
                          np.asarray(qB, float), pa[5], pa[6], split=True, k=k)


@pytest.mark.parametrize("depth,off,pa", [(0.2, (0.0, 0.0), (0.0, 0.0)), (0.3, (0.2, 0.2), (0.3, -0.7)),
                                           (0.45, (0.3, -0.3), (-1.1, 0.25))])
def test_box_on_box_area_force_and_contact_point(depth, off, pa):
    # A: 1 x 1.2 x 0.8 box centred at (0, pa); B: 2 x 2 x 2 box whose -x face
    # sits at x = 0.5 - depth inside A, so A's four side faces cross B's face:
    # the loop is A's 1.2 x 0.8 cross-section.  tol = mean edge of A / 10; the
    # depths and offsets keep every split point farther than tol from a corner
    # (closer ones are snapped into the corner, as the reference does).
    k = 3.0
    pA = np.array([0.0, pa[0], pa[1]])
    pB = pA + np.array([0.5 - depth + 1.0, off[0], off[1]])
    r = _pair(_box(1.0, 1.2, 0.8), _box(2, 2, 2), pA, [1, 0, 0, 0], pB, [1, 0, 0, 0], k=k)
    assert len(r["loops"]) == 1
    L = r["loops"][0]
    Sn, n, xc0, area, Fn, F = L[0:3], L[6:9], L[9:12], L[12], L[13], L[14:17]
    A = 1.2 * 0.8
    assert area == pytest.approx(A, rel=1e-13)
    assert np.linalg.norm(Sn) == pytest.approx(A, rel=1e-13)
    assert np.allclose(n, [1.0, 0.0, 0.0], atol=1e-13)                 # A -> B
    assert Fn == pytest.approx(0.5 * k * A, rel=1e-13)                 # = k/2 dOmega/ddelta, Omega = delta A
    assert np.allclose(F, Fn * n, atol=1e-13)
    c = np.array([0.5 - depth, pa[0], pa[1]])                          # centre of the contact rectangle
    c_perp = c - n * (n @ c)
    assert np.allclose(xc0, -2.0 / 3.0 * c_perp, atol=1e-13)
    # torques are r x F about each centre: A gets r_A x (-F), B gets r_B x F
    assert np.allclose(L[17:20], np.cross(xc0 - pA, -F), atol=1e-13)
    assert np.allclose(L[20:23], np.cross(xc0 - pB, F), atol=1e-13)


def test_sn_translation_invariant_and_rotation_covariant():
    rng = np.random.default_rng(3)
    tA = synth.sphere_hull(64, 11)
    tB = synth.star_mesh(7, 2)
    # the walk direction of a loop (hence the sign of Sn) is not geometric;
    # the canonical Sn is the one of the A->B normal: -area * n
    canon = lambda r: -r["loops"][:, 12:13] * r["loops"][:, 6:9]
    base = _pair(tA, tB, [0, 0, 0], [1, 0, 0, 0], [1.45, 0.1, -0.05], [0.9, 0.1, 0.3, -0.2])
    assert len(base["loops"]) >= 1
    Sn0 = canon(base)
    # translation of both bodies: same loops, Sn unchanged
    t = np.array([12.5, -3.25, 0.75])
    tr = _pair(tA, tB, t, [1, 0, 0, 0], t + [1.45, 0.1, -0.05], [0.9, 0.1, 0.3, -0.2])
    assert len(tr["loops"]) == len(base["loops"])
    assert np.allclose(canon(tr), Sn0, rtol=1e-9, atol=1e-12)
    # common rotation R about the origin: Sn -> R Sn
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    R = _rot(q)

    def qmul(a, b):
        w1, x1, y1, z1 = a
        w2, x2, y2, z2 = b
        return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                         w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])

    qB = np.array([0.9, 0.1, 0.3, -0.2])
    qB /= np.linalg.norm(qB)
    ro = _pair(tA, tB, [0, 0, 0], q, R @ np.array([1.45, 0.1, -0.05]), qmul(q, qB))
    assert len(ro["loops"]) == len(base["loops"])
    got = sorted(canon(ro).tolist())
    want = sorted((Sn0 @ R.T).tolist())
    assert np.allclose(got, want, rtol=1e-9, atol=1e-12)


def test_planar_polygon_sn_gn_xc0():
    L = O.lib()
    # regular hexagon of radius 0.7 in the plane z = 0.3, centred at (0.4, -0.2)
    ang = np.arange(6) * np.pi / 3
    P = np.stack([0.4 + 0.7 * np.cos(ang), -0.2 + 0.7 * np.sin(ang), np.full(6, 0.3)], axis=1)
    Sn, Gn = np.zeros(3), np.zeros(3)
    L.orc_sn_gn(np.ascontiguousarray(P.reshape(-1)), 6, Sn, Gn)
    area = 1.5 * np.sqrt(3.0) * 0.7 ** 2
    nh = np.array([0.0, 0.0, 1.0])                       # counter-clockwise seen from +z
    c = np.array([0.4, -0.2, 0.3])
    assert np.allclose(Sn, area * nh, rtol=1e-13, atol=1e-15)
    assert np.allclose(Gn, -2.0 / 3.0 * area * np.cross(nh, c), rtol=1e-12, atol=1e-14)
    xc0, nA, ar = np.zeros(3), np.zeros(3), np.zeros(1)
    assert L.orc_xc0(Sn, Gn, xc0, nA, ar) == 1
    assert ar[0] == pytest.approx(area, rel=1e-14)
    assert np.allclose(nA, -nh, atol=1e-15)              # nA = -Sn/|Sn|
    assert np.allclose(xc0, -2.0 / 3.0 * (c - nh * (nh @ c)), atol=1e-13)


def test_mass_properties_box_and_tetrahedron():
    sx, sy, sz = 0.3, 0.5, 0.7
    _, p, _ = O.build_mesh(_box(sx, sy, sz) + np.array([1.0, 2.0, -0.5]), center=False)
    V = sx * sy * sz
    assert p[0] == pytest.approx(V, rel=1e-13)
    assert np.allclose(p[1:4], [1.0, 2.0, -0.5], atol=1e-13)
    I = np.diag([sy * sy + sz * sz, sx * sx + sz * sz, sx * sx + sy * sy]) * V / 12.0
    assert np.allclose(p[7:16].reshape(3, 3), I, rtol=1e-12, atol=1e-15)
    # unit right tetrahedron (0,0,0),(1,0,0),(0,1,0),(0,0,1): V = 1/6, centroid 1/4
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    tris = v[np.array([(0, 2, 1), (0, 1, 3), (0, 3, 2), (1, 2, 3)])]
    _, p, _ = O.build_mesh(tris, center=False)
    assert p[0] == pytest.approx(1.0 / 6.0, rel=1e-14)
    assert np.allclose(p[1:4], 0.25, atol=1e-15)
    # about the centroid: diagonal 1/80, products of inertia -(-1/480) -> +1/480 off-diagonal
    It = np.full((3, 3), 1.0 / 480.0) + np.eye(3) * (1.0 / 80.0 - 1.0 / 480.0)
    assert np.allclose(p[7:16].reshape(3, 3), It, rtol=1e-12, atol=1e-16)


def test_torque_free_rotation_keeps_angular_momentum():
    L = O.lib()
    Ib = np.diag([1.0, 2.0, 3.5])
    Ibinv = np.linalg.inv(Ib)
    q = np.array([0.8, 0.2, -0.4, 0.4])
    q /= np.linalg.norm(q)
    R = _rot(q)
    omega = np.array([1.3, -0.7, 2.2])
    Lw = R @ Ib @ R.T @ omega
    s = np.concatenate([[0.0, 0.0, 0.0], q, [0.0, 0.0, 0.0], omega, Lw])
    for _ in range(1000):
        L.orc_integrate(s, np.ascontiguousarray(Ibinv.reshape(-1)), 1.0, np.zeros(3), np.zeros(3), np.zeros(3), 1e-3)
    assert np.array_equal(s[13:16], Lw)                  # L += T dt with T = 0: exactly unchanged
    assert abs(np.linalg.norm(s[3:7]) - 1.0) < 1e-14     # q renormalised every step


@pytest.mark.parametrize("depth,off,pa,moving", [(0.2, (0.0, 0.0), (0.0, 0.0), False),
                                                 (0.3, (0.2, 0.2), (0.3, -0.7), False),
                                                 (0.3, (0.2, 0.2), (0.3, -0.7), True)])
def test_minimal_force_model_in_the_step_box_on_box(depth, off, pa, moving, tmp_path):
    """The oracle's time step with "force_model minimal": two boxes get
    F_B = k/2 A n and -F_B, acting at the centre of the contact rectangle (the
    torque about each body centre is (c - x) x F), whatever the world position
    of the pair, at rest or approaching and spinning (polyline.py's force has
    no velocity term)."""
    import golden_inputs as GI

    k = 5.0
    text, files, pA, pB, c = GI.box_pair_case(depth, off, pa, k)
    if moving:
        text = text.replace("vel = 0 0 0", "vel = 1.5 0 0", 1).replace("vel = 0 0 0", "vel = -1.5 0 0", 1)
        text = text.replace("omega = 0 0 0", "omega = 0.3 -0.2 0.1", 1)
        assert "vel = 1.5 0 0" in text and "vel = -1.5 0 0" in text
    for fn, tris in files.items():
        synth.write_stl_ascii(str(tmp_path / fn), tris)
    cfg = tmp_path / "boxes.txt"
    cfg.write_text(text)
    s = O.Sim(str(cfg), str(tmp_path))
    ctr = s.step(1)
    assert ctr[0] == 1                                      # one contact loop
    F, T, cnt = s.forces()
    Fb = 0.5 * k * (1.2 * 0.8) * np.array([1.0, 0.0, 0.0])
    assert np.allclose(F[1], Fb, rtol=1e-12, atol=1e-12) and np.allclose(F[0], -Fb, rtol=1e-12, atol=1e-12)
    assert np.allclose(T[0], np.cross(c - pA, -Fb), atol=1e-12)
    assert np.allclose(T[1], np.cross(c - pB, Fb), atol=1e-12)
    assert list(cnt) == [1, 1]
