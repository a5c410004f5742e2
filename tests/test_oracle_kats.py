"""The reference's 24 unit-test known answers, restated against the FP64 oracle.

Each test names the reference test it restates (tests/snow_tests.cpp:line)
and uses the same inputs, expected values and the same absolute tolerance
(1e-10).  These pin the oracle's stage functions (oracle/snow_oracle.cpp),
which are then the parity source for the GPU path.
"""
import numpy as np
import pytest

import oracle as o

TOL = 1e-10


def test_spline():  # snow_tests.cpp:30-44
    assert abs(o.N(1, 1) - 2.0 / 3.0) < TOL
    assert abs(o.N(1, 2) - 1.0 / 6.0) < TOL
    assert o.N(1, 3) == 0
    assert o.N(1, 10) == 0
    assert abs(o.N(1.5, 0.75) - o.N(0.75, 1.5)) < TOL


def test_N_dash():  # snow_tests.cpp:46-60
    assert abs(o.N_dash(4, 4.3) + o.N_dash(4.3, 4)) < TOL
    assert abs(o.N_dash(1, 1)) < TOL
    assert abs(o.N_dash(0, 1) + 0.5) < TOL
    assert abs(o.N_dash(0, 2)) < TOL
    assert abs(o.N_dash(0, 0.5) + 0.625) < TOL


def _totals(g):
    d = g.dump()
    return d["mass"].sum(), d["momentum"].sum(axis=0)


def test_P2G():  # snow_tests.cpp:62-77
    g = o.Grid()
    p = o.make_particle((2.1, 3.5, 4.7), (3, -2, 0.5), 2, 1)
    o.lib.orc_p2g(g.h, p, 0.5)
    m, mom = _totals(g)
    assert abs(m - 2) < TOL
    assert np.linalg.norm(mom - np.array([6, -4, 1])) < TOL


def test_overlap():  # snow_tests.cpp:79-96
    g = o.Grid()
    a = o.make_particle((2.1, 3.5, 4.7), (3, -2, 0.5), 2, 1)
    b = o.make_particle((2.3, 3.7, 4.4), (-1, 4, 2), 1.5, 1)
    o.lib.orc_p2g(g.h, a, 0.5)
    o.lib.orc_p2g(g.h, b, 0.5)
    m, mom = _totals(g)
    assert abs(m - 3.5) < TOL
    assert np.linalg.norm(mom - np.array([4.5, 2, 4])) < TOL


def test_translation():  # snow_tests.cpp:98-122
    ga, gb = o.Grid(), o.Grid()
    o.lib.orc_p2g(ga.h, o.make_particle((2.15, 3.35, 4.65), (1, 2, 3), 1, 1), 0.5)
    o.lib.orc_p2g(gb.h, o.make_particle((2.65, 3.35, 4.65), (1, 2, 3), 1, 1), 0.5)
    assert len(ga) == len(gb)
    da = ga.dump()
    for n in range(len(ga)):
        i, j, k = da["idx"][n]
        nb = gb.get((i + 1, j, k))
        assert nb is not None
        assert abs(da["mass"][n] - nb["mass"]) < TOL
        assert np.linalg.norm(da["momentum"][n] - nb["momentum"]) < TOL


def test_volume():  # snow_tests.cpp:124-131
    g = o.Grid()
    p = o.make_particle((1, 2, 3), (1, 2, 3), 1, 7)
    o.lib.orc_p2g(g.h, p, 0.5)
    o.lib.orc_calculate_volume(g.h, p, 0.5)
    assert abs(p.V0 - 1) < TOL


def test_force_zero():  # snow_tests.cpp:133-144
    g = o.Grid()
    p = o.make_particle((1, 2, 3), (1, 2, 3), 1, 7)
    o.lib.orc_p2g(g.h, p, 0.5)
    o.lib.orc_calculate_volume(g.h, p, 0.5)
    o.lib.orc_force_calculation(g.h, p, 0.5, 1, 1, 0)
    d = g.dump()
    assert np.all(np.linalg.norm(d["force"], axis=1) < TOL)


def test_force_non_zero():  # snow_tests.cpp:146-162
    g = o.Grid()
    Fe = np.eye(3)
    Fe[0, 0] = 0.8  # particle.F_E.col(0) << 0.8, 0, 0
    p = o.make_particle((0.5, 0.5, 0.5), (0, 0, 0), 1, 1, Fe=Fe)
    o.lib.orc_p2g(g.h, p, 1)
    o.lib.orc_force_calculation(g.h, p, 1, 1, 1, 0)
    f = g.get((0, 0, 0))["force"]
    assert abs(f[0] - (-0.06888020833)) < TOL
    assert abs(f[1] - (-0.02296006944)) < TOL
    assert abs(f[2] - (-0.02296006944)) < TOL
    assert np.linalg.norm(g.dump()["force"].sum(axis=0)) < TOL


def test_velocity():  # snow_tests.cpp:164-180
    g = o.Grid()
    g.set((0, 0, 0), mass=2, momentum=(6, 4, 0), force=(2, -4, 1))
    g.set((0, 0, 1), mass=0, force=(2, -4, 1))
    o.lib.orc_grid_velocity(g.h, o.vec(0, -9.81, 0), 0.01)
    v0 = g.get((0, 0, 0))["velocity"]
    assert abs(v0[0] - 3.01) < TOL
    assert abs(v0[1] - 1.8819) < TOL
    assert abs(v0[2] - 0.005) < TOL
    assert np.all(np.abs(g.get((0, 0, 1))["velocity"]) < TOL)


def _grid_collide(node, vel, friction, dx):
    g = o.Grid()
    g.set(node, velocity=vel)
    arr, nc = o.colliders_array([o.make_plane((0, 0, 0), (0, 1, 0), friction)])
    o.lib.orc_grid_collisions(g.h, arr, nc, dx)
    return g.get(node)["velocity"]


def test_collision_outside_grid():  # snow_tests.cpp:182-192
    assert np.linalg.norm(_grid_collide((0, 1, 0), (2, -3, 0), 1, 0.1) - [2, -3, 0]) < TOL


def test_escaping_collision_grid():  # snow_tests.cpp:193-203
    assert np.linalg.norm(_grid_collide((0, -1, 0), (2, 3, 0), 1, 1) - [2, 3, 0]) < TOL


def test_sticking_grid():  # snow_tests.cpp:204-214
    assert np.linalg.norm(_grid_collide((0, -1, 0), (0.2, -3, 0), 0.6, 0.1)) < TOL


def test_sliding_grid():  # snow_tests.cpp:215-224
    assert np.linalg.norm(_grid_collide((0, -1, 0), (2, -3, 0), 0.2, 0.1) - [1.4, 0, 0]) < TOL


def _uniform_grid(vel_fn):
    g = o.Grid()
    for i in range(4):
        for j in range(4):
            for k in range(4):
                g.set((i, j, k), velocity=vel_fn(i, j, k))
    return g


def test_rigid_translation():  # snow_tests.cpp:226-246
    g = _uniform_grid(lambda i, j, k: (3, -2, 1))
    p = o.make_particle((0.15, 0.15, 0.15), (0, 0, 0), 0, 0)
    o.lib.orc_stress_update(g.h, p, 0.1, 0.1, 0.1, 0.1)
    assert np.linalg.norm(np.array(p.Fe[:]) - np.eye(3).ravel()) < TOL
    assert np.linalg.norm(np.array(p.Fp[:]) - np.eye(3).ravel()) < TOL


def test_elastic():  # snow_tests.cpp:248-272
    g = _uniform_grid(lambda i, j, k: (3, -2, 1))
    p = o.make_particle((0.15, 0.15, 0.15), (0, 0, 0), 0, 0, Fe=np.diag([0.99, 1, 1]))
    o.lib.orc_stress_update(g.h, p, 0.1, 0.1, 0.025, 0.1)
    assert np.linalg.norm(np.array(p.Fe[:]) - np.diag([0.99, 1, 1]).ravel()) < TOL
    assert np.linalg.norm(np.array(p.Fp[:]) - np.eye(3).ravel()) < TOL


def test_plastic():  # snow_tests.cpp:274-301
    g = _uniform_grid(lambda i, j, k: (3, -2, 1))
    p = o.make_particle((0.15, 0.15, 0.15), (0, 0, 0), 0, 0, Fe=np.diag([0.95, 1, 1]))
    o.lib.orc_stress_update(g.h, p, 0.1, 0.1, 0.025, 0.1)
    Fe = np.array(p.Fe[:]).reshape(3, 3)
    Fp = np.array(p.Fp[:]).reshape(3, 3)
    assert np.linalg.norm(Fe - np.diag([0.975, 1, 1])) < TOL
    assert np.linalg.norm(Fp - np.diag([0.95 / 0.975, 1, 1])) < TOL
    assert np.linalg.norm(Fe @ Fp - np.diag([0.95, 1, 1])) < TOL


def test_linear_velocity_field():  # snow_tests.cpp:303-327
    g = _uniform_grid(lambda i, j, k: (-0.1 * (0.1 * i), 0, 0))
    p = o.make_particle((0.15, 0.15, 0.15), (0, 0, 0), 0, 0)
    o.lib.orc_stress_update(g.h, p, 0.1, 0.1, 0.025, 0.1)
    assert np.linalg.norm(np.array(p.Fe[:]) - np.diag([0.99, 1, 1]).ravel()) < TOL
    assert np.linalg.norm(np.array(p.Fp[:]) - np.eye(3).ravel()) < TOL


def _g2p_grid(vp, node_p, node_v):
    g = o.Grid()
    p = o.make_particle((0.15, 0.15, 0.15), vp, 0, 0)
    o.lib.orc_p2g(g.h, p, 0.1)
    d = g.dump()
    for ijk in d["idx"]:
        g.set(ijk, mass=1, momentum=node_p, velocity=node_v)
    o.lib.orc_g2p(g.h, p, 0.1, 0.95)
    return np.array(p.v[:])


def test_G2P():  # snow_tests.cpp:329-343
    assert np.linalg.norm(_g2p_grid((2, 3, 4), (2, 3, 4), (2, 2, 4)) - [2, 2, 4]) < TOL


def test_FLIP_PIC():  # snow_tests.cpp:345-359
    assert np.linalg.norm(_g2p_grid((10, 0, 0), (1, 0, 0), (3, 0, 0)) - [11.55, 0, 0]) < TOL


def _particle_collide(x, v, friction):
    p = o.make_particle(x, v, 1, 1)
    arr, nc = o.colliders_array([o.make_plane((0, 0, 0), (0, 1, 0), friction)])
    o.lib.orc_particle_collisions(p, arr, nc)
    return np.array(p.v[:])


def test_collision_outside_particle():  # snow_tests.cpp:361-369
    assert np.linalg.norm(_particle_collide((0, 1, 0), (2, -3, 0), 1) - [2, -3, 0]) < TOL


def test_escaping_collision_particle():  # snow_tests.cpp:371-379
    assert np.linalg.norm(_particle_collide((0, -1, 0), (2, 3, 0), 1) - [2, 3, 0]) < TOL


def test_sticking_particle():  # snow_tests.cpp:381-389
    assert np.linalg.norm(_particle_collide((0, -1, 0), (0.2, -3, 0), 0.6)) < TOL


def test_sliding_particle():  # snow_tests.cpp:391-399
    assert np.linalg.norm(_particle_collide((0, -1, 0), (2, -3, 0), 0.2) - [1.4, 0, 0]) < TOL


def test_position_update():  # snow_tests.cpp:401-407
    p = o.make_particle((1, 2, 3), (2, -1, 0.5), 1, 1)
    o.lib.orc_position_update(p, 0.1)
    assert np.linalg.norm(np.array(p.x[:]) - [1.2, 1.9, 3.05]) < TOL


@pytest.mark.parametrize("planted", ["none"])
def test_kats_detect_planted_mistake(planted):
    """A plausible slip (gradient sign flipped) must break the force KAT.

    Computed on the CPU from the oracle's own stencil: negating N_dash's
    sign convention changes f(0,0,0) from the known answer.
    """
    idx = np.zeros(192, dtype=np.int32)
    w = np.zeros(64)
    grad = np.zeros(192)
    import ctypes as C
    p = o.make_particle((0.5, 0.5, 0.5), (0, 0, 0), 1, 1, Fe=np.diag([0.8, 1, 1]))
    o.lib.orc_build_stencil(p, 1.0, idx.ctypes.data_as(C.POINTER(C.c_int)),
                            w.ctypes.data_as(C.POINTER(C.c_double)),
                            grad.ctypes.data_as(C.POINTER(C.c_double)))
    P = np.zeros(9)
    o.lib.orc_force_constitutive(p, 1, 1, 0, P.ctypes.data_as(C.POINTER(C.c_double)))
    M = -1.0 * P.reshape(3, 3) @ np.diag([0.8, 1, 1]).T
    n0 = [n for n in range(64) if tuple(idx.reshape(64, 3)[n]) == (0, 0, 0)][0]
    g0 = grad.reshape(64, 3)[n0]
    f_good = M @ g0
    f_bad = M @ (-g0)
    assert np.linalg.norm(f_good - [-0.06888020833, -0.02296006944, -0.02296006944]) < TOL
    assert np.linalg.norm(f_bad - [-0.06888020833, -0.02296006944, -0.02296006944]) > 1e-3
