"""Edge cases of the step against the FP64 oracle:
  * particles exactly on cell faces: the kernels bin with x * (1/dx) where
    the reference divides (grid.cpp:67), so a particle on a face may take the
    stencil one node lower; the node it gains carries weight 0 (the cubic
    B-spline vanishes at distance 2), so the step must still match;
  * a one-particle scene (a single-cell home block, every other node empty);
  * 200 particles in one cell (a cell spanning several P2G staging chunks);
  * an engine with no resident scene (a step is a no-op);
  * the semi-implicit solve with no unknowns (every node held by a
    collider: no iteration, the explicit step bit for bit) and with an
    iteration cap of 1.
Bound: the 1e-4 parity metric of mpmref.rel_errors after 10 steps."""
import numpy as np
import pytest

import mpmref
import snowmpm

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _scene(x, v, dx=0.02, dt=1e-4, E=1.4e5):
    prm = snowmpm.params(dt, dx, *snowmpm.lame(E, 0.2), 10.0, 0.025, 0.0075, 0.95)
    return dict(x=np.ascontiguousarray(x, dtype=np.float64), v=np.ascontiguousarray(v, dtype=np.float64),
                m=np.full(len(x), 4e-4), params=prm, colliders=[snowmpm.make_plane((0, 0, 0), (0, 1, 0), 0.5)])


def _check(sc, nsteps=10, label=""):
    ref = mpmref.run_oracle(sc, nsteps)
    got = mpmref.run_gpu(sc, nsteps)
    ex, ev = mpmref.rel_errors(got, ref, sc, nsteps)
    print(f"{label}: {len(sc['x'])} particles, rel err x {ex:.3e} v {ev:.3e}", flush=True)
    assert got["status"][0] == 0
    assert ex < TOL and ev < TOL
    assert np.abs(got["Fe"] - ref["Fe"]).max() < 1e-4
    return got, ref


def test_particles_on_cell_faces_match_oracle():
    dx = 0.02
    g = np.stack(np.meshgrid(*[np.arange(16)] * 3, indexing="ij"), -1).reshape(-1, 3)
    # every coordinate a multiple of dx/2: half the particles sit on cell faces
    # (x / dx an integer in float32), the rest at cell centres
    # (float32 values, so the GPU's stored positions are the oracle's inputs)
    x = ((g * 0.5 + np.array([10.0, 3.0, 10.0])) * dx).astype(np.float32).astype(np.float64)
    # where the oracle's FP64 x / dx and the kernels' float32 x * (float)(1/dx)
    # take different cells
    split = np.floor(x / dx) != np.floor(x.astype(np.float32) * np.float32(1.0 / dx))
    print(f"particles whose x / dx and x * (1/dx) floor differently: {int(split.any(axis=1).sum())}", flush=True)
    assert split.any(axis=1).sum() > 100
    rng = np.random.default_rng(7)
    v = np.tile([0.2, -1.5, 0.1], (len(x), 1)) + rng.normal(0, 0.1, x.shape)
    _check(_scene(x, v, dx), label="lattice-aligned")


def test_single_particle_matches_oracle():
    _check(_scene([[0.31, 0.05, 0.27]], [[0.5, -2.0, 0.3]]), label="one particle")


def test_crowded_cell_matches_oracle():
    rng = np.random.default_rng(11)
    x = np.array([0.41, 0.11, 0.37]) + rng.uniform(0.001, 0.019, (200, 3))  # one 0.02 m cell
    v = np.tile([0.0, -1.0, 0.0], (200, 1)) + rng.normal(0, 0.05, (200, 3))
    _check(_scene(x, v, E=2e4), label="200 particles in one cell")


def test_step_without_scenes_is_a_noop():
    with snowmpm.Engine(capacity=1024) as eng:
        eng.step(3)
        eng.step_graph(2)
        sc = _scene([[0.3, 0.2, 0.3]], [[0.0, 0.0, 0.0]])
        sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
        eng.remove_scene(sid)
        eng.step(2)


def _implicit(x, v, colliders, maxit=200):
    prm = snowmpm.params(4e-4, 0.02, *snowmpm.lame(1e6, 0.2), 10.0, 0.025, 0.0075, 0.95, beta=1.0,
                         cr_tolerance=1e-6, cr_max_iterations=maxit)
    x = np.ascontiguousarray(x, dtype=np.float64)
    return dict(x=x, v=np.ascontiguousarray(v, dtype=np.float64), m=np.full(len(x), 4e-4), params=prm,
                colliders=colliders)


def test_implicit_scene_with_every_node_held_by_a_collider():
    """A block of snow wholly inside a sticky half-space: every node with mass
    is changed by the collider, so the solve has no unknowns and takes no
    iteration, and the step is the explicit one bit for bit (the block stops
    in place: velocities 0 and positions unchanged to FP32 rounding, the
    tolerances of tests/test_gpu_contact.py)."""
    g = np.stack(np.meshgrid(*[np.arange(6)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = (g + 0.5) * 0.01 + np.array([0.2, -0.3, 0.2])
    out = {}
    for beta in (1.0, 0.0):
        sc = _implicit(x, np.tile([0.0, -1.0, 0.0], (len(x), 1)), [snowmpm.make_plane((0, 0, 0), (0, 1, 0), 1e6)])
        sc["params"].beta = beta
        with snowmpm.Engine(capacity=len(x) + 64) as eng:
            sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
            eng.step(1)
            out[beta] = (eng.cr_stats(sid)["iterations"], eng.download(sid))
            assert eng.status(sid)[0] == 0
    assert out[1.0][0] == 0
    for f in ("x", "v", "Fe"):
        assert np.array_equal(out[1.0][1][f], out[0.0][1][f]), f
    got = out[1.0][1]
    assert np.abs(got["v"]).max() < 6e-6 and np.abs(got["x"] - x).max() < 2e-7


def test_implicit_iteration_cap():
    """cr_max_iterations = 1: the solve stops after one iteration, reports it,
    and leaves a finite state; the residual history has both ends."""
    g = np.stack(np.meshgrid(*[np.arange(8)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = (g + 0.5) * 0.01 + np.array([0.1, 0.012, 0.1])
    sc = _implicit(x, np.tile([0.3, -2.0, 0.1], (len(x), 1)), [snowmpm.make_plane((0, 0, 0), (0, 1, 0), 0.5)],
                   maxit=1)
    with snowmpm.Engine(capacity=len(x) + 64) as eng:
        sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
        eng.step(3)
        st = eng.cr_stats(sid)
        got = eng.download(sid)
        assert eng.status(sid)[0] == 0
    assert st["iterations"] == 1 and len(st["residual"]) == 2
    assert np.all(np.isfinite(got["v"])) and np.all(np.isfinite(got["x"]))
