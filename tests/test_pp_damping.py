"""The reference's dormant particle-particle normal damping (src/main.cpp:
2244-2248 computes cn, 2258-2260 leaves ``fn += -cn vn`` commented out),
switched on by ``force_model zdem_damped``, pinned by mirror symmetry against
the reference's wall force (2363-2431, damping on), which the compiled
reference pins bit-exactly:

Two mirror-image bodies approaching head-on with speeds +v and -v touch on
their mirror plane, where a wall would cut the first body in the same
section.  The particle-particle law has kn_pp = 2 (E/R) A and m_eff = m/2, so
sqrt(kn m_eff) and hence cn are the wall's, while the approach speed is 2v:
every term of the damped pair force is twice the wall's, F_pp = 2 F_wall.
Without the damping the two differ by 2 cn v, and when the bodies separate
the damping is off (vn > 0), so the damped force equals the undamped one.

The bodies are bipyramids whose cap triangles run round the axis in order, so
the wall's section points (taken in triangle order, not sorted by angle)
form the true section polygon."""
import os

import numpy as np
import pytest

import oracle as O
from zdx import synth


def bipyramid(n=12, r=5e-3, h=4e-3):
    ang = 2 * np.pi * np.arange(n) / n
    ring = np.stack([r * np.cos(ang), r * np.sin(ang), np.zeros(n)], 1)
    top, bot = np.array([0.0, 0.0, h]), np.array([0.0, 0.0, -h])
    tris = [[top, ring[k], ring[(k + 1) % n]] for k in range(n)] + \
           [[bot, ring[(k + 1) % n], ring[k]] for k in range(n)]
    return np.array(tris)


# A's apex passes the mirror plane by 0.6 h, so the section (radius 0.6 r) is
# well above the snapping tolerance (a tenth of the mean edge, about 0.1 r)
H, DELTA = 4e-3, 2.4e-3


def write_cases(d, v, model):
    """Config files: the pair (A below, its mirror B above, speeds +v, -v) and
    A alone against a plane wall at z = 0."""
    synth.write_stl_ascii(os.path.join(d, "bp.stl"), bipyramid())
    synth.write_stl_ascii(os.path.join(d, "plane.stl"), synth.plane_wall(0.1, 0.1))
    mat = dict(density=2500.0, young=1e4, poisson=0.25, mu=0.5, restitution=0.5)
    zA, zB = -(H - DELTA), H - DELTA
    # off the plane wall's diagonal, so one wall triangle holds the whole section
    part = lambda z, vz: dict(stl=os.path.join(d, "bp.stl"), pos=[0.02, -0.02, z], vel=[0.0, 0.0, vz],
                              quat=[1.0, 0.0, 0.0, 0.0], omega=[0.0, 0.0, 0.0], scale=1.0, **mat)
    g = dict(steps=1, dt=1e-7, split_contacts=1, contact_debug=0, gravity=[0.0, 0.0, 0.0], center_mesh=1,
             vtk_prefix="m", output_dir=os.path.join(d, "out"), output_interval=1000, force_model=model)
    pair = os.path.join(d, "pair.txt")
    with open(pair, "w") as f:
        f.write(synth.format_config(g, [part(zA, v), part(zB, -v)]))
    wall = os.path.join(d, "wall.txt")
    walls = [dict(stl=os.path.join(d, "plane.stl"), pos=[0.0, 0.0, 0.0], quat=[1.0, 0.0, 0.0, 0.0], scale=1.0,
                  mu=0.5, restitution=0.5)]
    with open(wall, "w") as f:
        f.write(synth.format_config(g, [part(zA, v)], walls))
    return pair, wall


def forces_oracle(cfg, d):
    s = O.Sim(cfg, d)
    c = s.step(1)
    F, T, cnt = s.forces()
    return F, c


@pytest.mark.parametrize("v", [0.5, 2.0])
def test_damped_pair_force_is_twice_the_wall_force(tmp_path, v):
    d = str(tmp_path)
    pair, wall = write_cases(d, v, "zdem_damped")
    Fp, cp = forces_oracle(pair, d)
    Fw, cw = forces_oracle(wall, d)
    assert cp[0] == 1 and cw[0] == 1
    # the force on A (below) points down in both; the pair's is twice the wall's
    assert Fp[0, 2] < 0 and Fw[0, 2] < 0
    assert np.allclose(Fp[0], 2.0 * Fw[0], rtol=1e-9, atol=1e-12 * abs(Fw[0, 2])), (Fp[0], Fw[0])
    assert np.allclose(Fp[1], -Fp[0], rtol=0, atol=0)
    # without the damping the pair force misses 2 cn v (a sizeable part at
    # these speeds and a soft material, young 1e4)
    pair0, _ = write_cases(d, v, "zdem")
    Fu, _ = forces_oracle(pair0, d)
    assert abs(Fu[0, 2] - 2.0 * Fw[0, 2]) > 0.1 * abs(Fw[0, 2])
    assert abs(Fu[0, 2]) < abs(Fp[0, 2])


def test_damping_is_off_while_the_bodies_separate(tmp_path):
    d = str(tmp_path)
    pair, _ = write_cases(d, -1.0, "zdem_damped")
    pair0, _ = write_cases(d + "/", -1.0, "zdem")
    Fd, _ = forces_oracle(pair, d)
    Fu, _ = forces_oracle(pair0, d)
    assert Fd[0, 2] < 0 and np.array_equal(Fd, Fu)


@pytest.mark.gpu
@pytest.mark.parametrize("v", [0.5, 2.0])
def test_device_damped_pair_force(tmp_path, v):
    """The same identity on the device step, and the device forces equal to
    the oracle's (to the last bits of cbrt/pow/log)."""
    zdx = pytest.importorskip("zdx")
    d = str(tmp_path)
    ctx = zdx.Context(0)
    out = {}
    for model in ("zdem_damped", "zdem"):
        pair, wall = write_cases(d, v, model)
        for name, cfg in (("pair", pair), ("wall", wall)):
            sim = zdx.Sim.from_config(ctx, cfg, d)
            for mode in (zdx.MODE_PLAIN, zdx.MODE_BATCHED):
                sim2 = zdx.Sim.from_config(ctx, cfg, d)
                sim2.step(1, mode)
                out[(model, name, mode)] = sim2.forces()[0]
                sim2.close()
            sim.close()
            out[(model, name, "oracle")] = forces_oracle(cfg, d)[0]
    ctx.close()
    for mode in (zdx.MODE_PLAIN, zdx.MODE_BATCHED):
        Fp, Fw = out[("zdem_damped", "pair", mode)], out[("zdem_damped", "wall", mode)]
        assert np.allclose(Fp[0], 2.0 * Fw[0], rtol=1e-9, atol=1e-12 * abs(Fw[0, 2])), (Fp[0], Fw[0])
        assert abs(out[("zdem", "pair", mode)][0, 2] - 2.0 * Fw[0, 2]) > 0.1 * abs(Fw[0, 2])
        for key in ("zdem_damped", "zdem"):
            for name in ("pair", "wall"):
                assert np.allclose(out[(key, name, mode)], out[(key, name, "oracle")], rtol=1e-12, atol=1e-15)
