"""Contact on the GPU: the whole-step closed forms of the oracle tests
(tests/test_oracle_invariants.py) run through the C ABI, and moving colliders
(CollisionBody::velocity, collision.hpp:10, entering collision_helper as v_c,
grid.cpp:206-231) against the oracle.

Closed-form tolerances: velocities 2e-6 of the scene's speed scale and
positions 2e-7 m, a few FP32 roundings of |v| <= 3 m/s and |x| <= 0.2 m."""
import numpy as np
import pytest

import mpmref
import oracle as orc
import snowmpm
import test_oracle_invariants as closed

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _prm():
    return snowmpm.params(1e-4, 0.02, *snowmpm.lame(1.4e5, 0.2), 10, 0.025, 0.0075, 0.95)


def _one_step(x, v0, colliders):
    sc = dict(x=x, v=np.tile(v0, (len(x), 1)), m=np.full(len(x), 4e-4), params=_prm(), colliders=colliders)
    got = mpmref.run_gpu(sc, 1)
    assert got["status"][0] == 0
    return got


@pytest.mark.parametrize("vc", [(0.0, 0.0, 0.0), (0.7, 0.0, -0.3)])
def test_inside_sticky_halfspace_closed_form(vc):
    """Every particle inside a moving friction-1 half-space reaches v_c in one
    step and advances by dt v_c."""
    x = closed._block()
    got = _one_step(x, np.array([0.2, -3.0, 0.1]), [snowmpm.make_plane((0, 0.5, 0), (0, 1, 0), 1.0, velocity=vc)])
    assert np.abs(got["v"] - np.array(vc)).max() < 6e-6
    assert np.abs(got["x"] - (x + 1e-4 * np.array(vc))).max() < 2e-7
    assert np.abs(got["Fe"] - np.eye(3)).max() < 1e-6


def test_inside_frictionless_halfspace_closed_form():
    x = closed._block()
    v0, vc = np.array([0.2, -3.0, 0.1]), np.array([0.7, 0.05, -0.3])
    got = _one_step(x, v0, [snowmpm.make_plane((0, 0.5, 0), (0, 1, 0), 0.0, velocity=vc)])
    want = np.array([v0[0], vc[1], v0[2]])
    assert np.abs(got["v"] - want).max() < 6e-6
    assert np.abs(got["x"] - (x + 1e-4 * want)).max() < 2e-7


def test_particle_collision_order_closed_form():
    """Particle collisions at the pre-advection x, advection with the collided
    v, a moving box that no grid node lies in."""
    x, lab, v0, vc, (c, he, mu) = closed.thin_box_case()
    got = _one_step(x, v0, [snowmpm.make_box(c, he, mu, velocity=vc)])
    prm = _prm()
    xw, vw = closed.thin_box_expected(x, lab, v0, vc, prm.dt, prm.gravity[:])
    assert np.abs(got["v"] - vw).max() < 6e-6
    assert np.abs(got["x"] - xw).max() < 2e-7


def test_grid_contact_closed_form():
    """Grid nodes inside a moving sticky plane, particles outside: G2P and the
    deformation update read the post-collision grid velocity."""
    x, v0, vc, (pt, nrm, mu) = closed.grid_contact_case()
    got = _one_step(x, v0, [snowmpm.make_plane(pt, nrm, mu, velocity=vc)])
    prm = _prm()
    xw, vw, Fw, _ = closed.grid_contact_expected(x, v0, vc, prm.dt, prm.dx, prm.gravity[:], pt[1], orc.N,
                                                 orc.N_dash)
    assert np.abs(got["v"] - vw).max() < 6e-6
    assert np.abs(got["x"] - xw).max() < 2e-7
    assert np.abs(got["Fe"] - Fw).max() < 2e-6


@pytest.mark.parametrize("case", ["conveyor_plane", "moving_sphere", "moving_box"])
def test_moving_colliders_match_oracle(case):
    """The 1,000-particle cube on moving bodies for 20 steps, GPU against the
    oracle: a conveyor ground plane (v_c = (1.5, 0, -0.4), friction 0.5) under
    a cube dropped onto it, a sphere driven into the cube's side
    (v_c = (-2, -0.5, 0.3)), a box lifting it (v_c = (0.2, 1.5, 0))."""
    sc = snowmpm.generate_scene(0, seed=11, scale=0.25)
    lo, hi = sc["x"].min(0), sc["x"].max(0)
    sc["x"] = sc["x"] - np.array([0, lo[1] - 0.004, 0])
    lo, hi = sc["x"].min(0), sc["x"].max(0)
    mid = 0.5 * (lo + hi)
    sc["v"] = np.tile(np.array([0.0, -1.0, 0.0]), (len(sc["x"]), 1))
    if case == "conveyor_plane":
        sc["colliders"] = [snowmpm.make_plane((0, 0, 0), (0, 1, 0), 0.5, velocity=(1.5, 0, -0.4))]
    elif case == "moving_sphere":
        sc["colliders"] = [snowmpm.make_plane((0, 0, 0), (0, 1, 0), 0.2),
                           snowmpm.make_sphere((hi[0] + 0.045, mid[1], mid[2]), 0.05, 0.4,
                                               velocity=(-2.0, -0.5, 0.3))]
    else:
        sc["colliders"] = [snowmpm.make_box((mid[0], lo[1] - 0.1 + 0.003, mid[2]), (0.3, 0.1, 0.3), 0.6,
                                            velocity=(0.2, 1.5, 0.0))]
    nsteps = 20
    ref = mpmref.run_oracle(sc, nsteps)
    got = mpmref.run_gpu(sc, nsteps)
    plain = dict(sc, colliders=[orc_static(c) for c in sc["colliders"]])
    ref_static = mpmref.run_oracle(plain, nsteps)
    # the body's velocity changes the outcome
    assert np.abs(ref_static["v"] - ref["v"]).max() > 0.05
    ex, ev = mpmref.rel_errors(got, ref, sc, nsteps)
    print(f"{case}: rel err x {ex:.3e} v {ev:.3e}")
    assert ex < TOL and ev < TOL
    assert np.abs(got["Fe"] - ref["Fe"]).max() < 1e-4
    assert got["status"][0] == 0


def orc_static(c):
    s = snowmpm.Collider()
    for name, _ in snowmpm.Collider._fields_:
        setattr(s, name, getattr(c, name))
    s.velocity[:] = [0.0, 0.0, 0.0]
    return s
