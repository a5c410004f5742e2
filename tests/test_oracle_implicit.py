"""The semi-implicit velocity solve of the FP64 oracle (the step the
reference declares and never defines: grid.hpp:77, 81, 89-92).  There is no
reference behaviour to match, so it is pinned by mathematics:
  * constitutive_differential equals the central finite difference of the
    corotated P(F) (force_constitutive, grid.cpp:162-172);
  * apply_A is self-adjoint and positive in the mass inner product (M A =
    M + beta dt^2 H with H a Hessian);
  * the solved velocities satisfy A v = v* for the explicit v* of the same
    step, to the solver's tolerance;
  * the conjugate-residual residual never grows;
  * beta = 0 is the explicit step; the solve keeps a mirror-symmetric scene
    mirror-symmetric."""
import numpy as np
import pytest

import oracle as o

MU, LA = o.lame(1e6, 0.2)


def _prm(beta, dt=4e-4, tol=1e-10):
    return o.params(dt, 0.02, MU, LA, 10, 0.025, 0.0075, 0.95, beta=beta, cr_tolerance=tol, cr_max_iterations=200)


def _P(F, Fp, mu0, la0, xi):
    q = o.make_particle((0, 0, 0), (0, 0, 0), 1, 1, Fe=F, Fp=Fp)
    out = (o.C.c_double * 9)()
    o.lib.orc_force_constitutive(o.C.byref(q), mu0, la0, xi, out)
    return np.array(out[:]).reshape(3, 3)


@pytest.mark.parametrize("seed", range(6))
def test_constitutive_differential_is_the_derivative(seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    F = Q @ np.diag(rng.uniform(0.9, 1.1, 3)) if seed % 2 else np.eye(3) + rng.normal(0, 0.05, (3, 3))
    Fp = np.diag(rng.uniform(0.95, 1.02, 3))
    dF = rng.normal(0, 1, (3, 3))
    p = o.make_particle((0, 0, 0), (0, 0, 0), 1, 1, Fe=F, Fp=Fp)
    dP = o.constitutive_differential(p, dF, 3e4, 2e4, 10)
    eps = 1e-6
    fd = (_P(F + eps * dF, Fp, 3e4, 2e4, 10) - _P(F - eps * dF, Fp, 3e4, 2e4, 10)) / (2 * eps)
    assert np.abs(fd - dP).max() <= 1e-7 * np.abs(dP).max()


def _scene(seed=0, n=6, v=(0.3, -2.0, 0.1), prestress=0.01):
    rng = np.random.default_rng(seed)
    gi = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = (gi + 0.5) * 0.01 + np.array([0.1, 0.012, 0.1]) + rng.uniform(-1e-3, 1e-3, (len(gi), 3))
    Fe = (np.eye(3) + rng.normal(0, prestress, (len(x), 3, 3))).reshape(-1, 9)
    st = o.State(x, np.tile(v, (len(x), 1)) + rng.normal(0, 0.2, (len(x), 3)), 4e-4, Fe=Fe)
    st.setup(0.02)
    return st, [o.make_plane((0, 0, 0), (0, 1, 0), 0.5)]


def _unknowns(st, cs, prm):
    """The step's grid and unknown nodes, with the pre-step particles."""
    pre = st.copy()
    g = o.Grid()
    st.step(cs, prm, grid=g)
    idx = st.krylov_nodes(g)
    d = g.dump()
    key = {tuple(k): i for i, k in enumerate(d["idx"].tolist())}
    rows = [key[tuple(k)] for k in idx.tolist()]
    return pre, g, idx, d["mass"][rows], d["velocity"][rows]


def test_apply_A_is_self_adjoint_and_positive():
    st, cs = _scene()
    prm = _prm(1.0)
    pre, g, idx, m, _ = _unknowns(st, cs, prm)
    assert len(idx) > 100
    rng = np.random.default_rng(3)
    for _ in range(3):
        a, b = rng.normal(size=(len(idx), 3)), rng.normal(size=(len(idx), 3))
        Aa, Ab = pre.apply_A(g, prm, a), pre.apply_A(g, prm, b)
        lhs, rhs = (m[:, None] * a * Ab).sum(), (m[:, None] * Aa * b).sum()
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs))
        assert (m[:, None] * a * Aa).sum() > 0
    # A differs from I by beta dt^2 M^-1 H: halving beta halves A - I
    a = rng.normal(size=(len(idx), 3))
    h1 = pre.apply_A(g, prm, a) - a
    h2 = pre.apply_A(g, _prm(0.5), a) - a
    assert np.abs(h1 - 2 * h2).max() <= 1e-12 * np.abs(h1).max() and np.abs(h1).max() > 1e-3


def test_solution_satisfies_A_v_equals_explicit_velocity():
    st, cs = _scene()
    st_exp = st.copy()
    prm = _prm(1.0, tol=1e-12)
    pre, g, idx, m, v = _unknowns(st, cs, prm)
    stats = o.last_cr_stats()
    ge = o.Grid()
    st_exp.step(cs, _prm(0.0), grid=ge)
    de = ge.dump()
    key = {tuple(k): i for i, k in enumerate(de["idx"].tolist())}
    vstar = de["velocity"][[key[tuple(k)] for k in idx.tolist()]]
    Av = pre.apply_A(g, prm, v)
    err = np.sqrt((m[:, None] * (Av - vstar) ** 2).sum() / (m[:, None] * vstar ** 2).sum())
    print(f"CR: {stats['iterations']} iterations, {stats['nodes']} unknowns, relative residual {err:.2e}")
    assert err <= 1e-11
    assert np.abs(v - vstar).max() > 1e-3  # the solve does change the velocities


def test_residual_never_grows():
    st, cs = _scene(prestress=0.02)
    prm = _prm(1.0, dt=8e-4, tol=1e-12)
    st.step(cs, prm)
    r = o.last_cr_stats()["residual"]
    assert len(r) > 5
    assert np.all(np.diff(r) <= 1e-12 * r[0]), r
    assert r[-1] <= 1e-12 * r[0] * 1e3


def test_beta_zero_is_the_explicit_step():
    a, cs = _scene()
    b = a.copy()
    prm0 = _prm(0.0)
    for _ in range(5):
        a.step(cs, prm0)
        b.step(cs, o.params(4e-4, 0.02, MU, LA, 10, 0.025, 0.0075, 0.95))
    assert np.array_equal(a.raw, b.raw)


def test_mirror_symmetric_scene_stays_symmetric():
    """Particles mirrored in x = 0 (velocities mirrored too) on a ground plane:
    five implicit steps keep the mirror symmetry to rounding."""
    rng = np.random.default_rng(9)
    gi = np.stack(np.meshgrid(np.arange(4), np.arange(6), np.arange(6), indexing="ij"), -1).reshape(-1, 3)
    half = (gi + 0.5) * 0.01 + np.array([0.0, 0.012, 0.1]) + rng.uniform(-1e-3, 1e-3, (len(gi), 3))
    half[:, 0] = np.abs(half[:, 0]) + 1e-4
    vh = np.array([0.5, -2.0, 0.1]) + rng.normal(0, 0.2, (len(half), 3))
    mir = np.array([-1.0, 1.0, 1.0])
    x = np.concatenate([half, half * mir])
    v = np.concatenate([vh, vh * mir])
    st = o.State(x, v, 4e-4)
    st.setup(0.02)
    cs = [o.make_plane((0, 0, 0), (0, 1, 0), 0.5)]
    prm = _prm(1.0, tol=1e-13)
    for _ in range(5):
        st.step(cs, prm)
    X, V = st.field("x"), st.field("v")
    k = len(half)
    assert np.abs(X[:k] * mir - X[k:]).max() <= 1e-12
    assert np.abs(V[:k] * mir - V[k:]).max() <= 1e-9 * np.abs(V).max()
