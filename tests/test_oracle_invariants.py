"""Pins for the oracle beyond the reference's known answers.

The reference's tests only use diagonal F_E and single particles, so these
checks close what they leave open, each chosen so that a plausible slip
fails: golden SVD/polar/clamp vectors from numpy/scipy (non-diagonal and
inverted F), B-spline partition of unity and linear reproduction, stress
rotation-equivariance, force balance, conservation in P2G, exact linear
kinematics in stress_update, det(F_E F_P) preservation, PIC/FLIP closed form,
the whole step's free-fall closed form and the lattice-interior volume.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as o

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "svd_vectors.npz"))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def stencil(x, dx):
    p = o.make_particle(x, (0, 0, 0), 1, 1)
    idx = np.zeros(192, dtype=np.int32)
    w = np.zeros(64)
    g = np.zeros(192)
    o.lib.orc_build_stencil(p, dx, idx.ctypes.data_as(C.POINTER(C.c_int)), _dp(w), _dp(g))
    return idx.reshape(64, 3), w, g.reshape(64, 3)


def test_svd_golden_vectors():
    for F, S, R, Cl in zip(GOLD["F"], GOLD["S"], GOLD["R"], GOLD["C"]):
        U, s, V = o.svd3(F)
        assert np.all(s >= 0) and s[0] >= s[1] >= s[2]
        assert np.allclose(s, S, atol=1e-12, rtol=1e-12)
        assert np.allclose(U @ np.diag(s) @ V.T, F, atol=1e-12)
        assert np.allclose(U.T @ U, np.eye(3), atol=1e-12) and np.allclose(V.T @ V, np.eye(3), atol=1e-12)
        assert np.allclose(U @ V.T, R, atol=1e-10)
        assert np.allclose(U @ np.diag(np.clip(s, 1 - 0.025, 1 + 0.0075)) @ V.T, Cl, atol=1e-10)


def test_golden_vectors_separate_signed_svd():
    """A 'rotation' SVD (det U = det V = +1, last sigma negative) gives another
    clamp result on inverted F: the fixtures would catch that convention slip."""
    inv = [k for k, F in enumerate(GOLD["F"]) if np.linalg.det(F) < 0]
    assert len(inv) >= 10
    differs = 0
    for k in inv:
        U, s, Vt = np.linalg.svd(GOLD["F"][k])
        if np.linalg.det(U) < 0:
            U[:, 2] *= -1
            s = s.copy()
            s[2] *= -1
        else:
            Vt[2] *= -1
            s = s.copy()
            s[2] *= -1
        signed = U @ np.diag(np.clip(s, 1 - 0.025, 1 + 0.0075)) @ Vt
        differs += not np.allclose(signed, GOLD["C"][k], atol=1e-6)
    assert differs == len(inv)


@pytest.mark.parametrize("seed", range(5))
def test_bspline_partition_and_linear_reproduction(seed):
    rng = np.random.default_rng(seed)
    dx = 0.07
    x = rng.uniform(-3, 3, size=3)
    idx, w, g = stencil(x, dx)
    assert abs(w.sum() - 1) < 1e-13                       # sum w = 1
    assert np.abs(g.sum(axis=0)).max() < 1e-11           # sum grad w = 0
    xi = idx * dx
    assert np.abs((w[:, None] * xi).sum(0) - x).max() < 1e-12          # sum w x_i = x_p
    assert np.abs(xi.T @ g - np.eye(3)).max() < 1e-11                 # sum x_i grad w^T = I


def test_stress_rotation_equivariance_and_golden_polar():
    rng = np.random.default_rng(1)
    mu, lam = 3.0, 2.0
    for F, R in zip(GOLD["F"][40:100], GOLD["R"][40:100]):
        Fp = np.eye(3) + 0.01 * rng.normal(size=(3, 3))
        xi = 10.0
        P = np.zeros(9)
        o.lib.orc_force_constitutive(o.make_particle((0, 0, 0), (0, 0, 0), 1, 1, Fe=F, Fp=Fp), mu, lam, xi, _dp(P))
        P = P.reshape(3, 3)
        h = np.exp(xi * (1 - np.linalg.det(Fp)))
        J = np.linalg.det(F)
        expect = 2 * mu * h * (F - R) + lam * h * (J - 1) * J * np.linalg.inv(F).T
        assert np.allclose(P, expect, atol=1e-9)
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        if np.linalg.det(Q) < 0:
            Q[:, 0] *= -1
        P2 = np.zeros(9)
        o.lib.orc_force_constitutive(o.make_particle((0, 0, 0), (0, 0, 0), 1, 1, Fe=Q @ F, Fp=Fp), mu, lam, xi,
                                     _dp(P2))
        assert np.allclose(P2.reshape(3, 3), Q @ P, atol=1e-9)


def test_force_balance_nondiagonal():
    rng = np.random.default_rng(2)
    for _ in range(10):
        F = np.eye(3) + 0.05 * rng.normal(size=(3, 3))
        g = o.Grid()
        p = o.make_particle(rng.uniform(0, 1, 3), (0, 0, 0), 1.0, 0.3, Fe=F)
        o.lib.orc_p2g(g.h, p, 0.1)
        o.lib.orc_force_calculation(g.h, p, 0.1, 5.0, 4.0, 10.0)
        f = g.dump()["force"]
        assert np.abs(f.sum(0)).max() < 1e-9 * max(1.0, np.abs(f).max())
        assert np.abs(f).max() > 1e-3


def test_p2g_conserves_many_particles():
    rng = np.random.default_rng(3)
    g = o.Grid()
    M, P = 0.0, np.zeros(3)
    for _ in range(200):
        x = rng.uniform(-1, 1, 3)
        v = rng.normal(size=3)
        m = rng.uniform(0.1, 2)
        o.lib.orc_p2g(g.h, o.make_particle(x, v, m, 1), 0.13)
        M += m
        P += m * v
    d = g.dump()
    assert abs(d["mass"].sum() - M) < 1e-10 * M
    assert np.abs(d["momentum"].sum(0) - P).max() < 1e-10 * max(1, np.abs(P).max())


def _affine_grid(A, b, base, dx):
    g = o.Grid()
    for i in range(base[0], base[0] + 4):
        for j in range(base[1], base[1] + 4):
            for k in range(base[2], base[2] + 4):
                xi = np.array([i, j, k]) * dx
                g.set((i, j, k), velocity=A @ xi + b)
    return g


def test_stress_update_linear_field_exact():
    """v_i = A x_i + b reproduces grad v = A exactly; without clamping
    F_E' = (I + dt A) F_E and F_P is unchanged."""
    rng = np.random.default_rng(4)
    dx, dt = 0.1, 0.01
    for _ in range(10):
        A = rng.normal(size=(3, 3))
        b = rng.normal(size=3)
        x = rng.uniform(0.5, 1.5, 3)
        base = np.floor(x / dx).astype(int) - 1
        g = _affine_grid(A, b, base, dx)
        Fe = np.eye(3) + 0.02 * rng.normal(size=(3, 3))
        p = o.make_particle(x, (0, 0, 0), 1, 1, Fe=Fe)
        o.lib.orc_stress_update(g.h, p, dx, dt, 0.9, 10.0)
        assert np.allclose(np.array(p.Fe[:]).reshape(3, 3), (np.eye(3) + dt * A) @ Fe, atol=1e-11)
        assert np.allclose(np.array(p.Fp[:]).reshape(3, 3), np.eye(3), atol=1e-11)


def test_plasticity_preserves_total_deformation():
    rng = np.random.default_rng(5)
    dx, dt = 0.1, 0.05
    plastic = 0
    for _ in range(20):
        A = 0.5 * rng.normal(size=(3, 3))
        x = rng.uniform(0.5, 1.5, 3)
        base = np.floor(x / dx).astype(int) - 1
        g = _affine_grid(A, np.zeros(3), base, dx)
        Fe = np.eye(3) + 0.01 * rng.normal(size=(3, 3))
        Fp = np.eye(3) + 0.01 * rng.normal(size=(3, 3))
        p = o.make_particle(x, (0, 0, 0), 1, 1, Fe=Fe, Fp=Fp)
        o.lib.orc_stress_update(g.h, p, dx, dt, 0.025, 0.0075)
        Fe2 = np.array(p.Fe[:]).reshape(3, 3)
        Fp2 = np.array(p.Fp[:]).reshape(3, 3)
        total = (np.eye(3) + dt * A) @ Fe @ Fp
        assert np.allclose(Fe2 @ Fp2, total, atol=1e-11)
        s = np.linalg.svd(Fe2, compute_uv=False)
        assert s.min() >= 1 - 0.025 - 1e-12 and s.max() <= 1 + 0.0075 + 1e-12
        plastic += not np.allclose(Fp2, Fp)
    assert plastic > 10


def test_g2p_pic_flip_closed_form():
    rng = np.random.default_rng(6)
    for _ in range(5):
        c, c_old = rng.normal(size=3), rng.normal(size=3)
        vp = rng.normal(size=3)
        m = rng.uniform(0.5, 2)
        x = rng.uniform(0.2, 0.8, 3)
        g = o.Grid()
        p = o.make_particle(x, vp, 1, 1)
        o.lib.orc_p2g(g.h, p, 0.1)
        for ijk in g.dump()["idx"]:
            g.set(ijk, mass=m, momentum=m * c_old, velocity=c)
        o.lib.orc_g2p(g.h, p, 0.1, 0.95)
        expect = 0.05 * c + 0.95 * (vp + c - c_old)
        assert np.allclose(np.array(p.v[:]), expect, atol=1e-12)


def test_step_freefall_closed_form():
    n = 6
    h = 0.01
    gi = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = (gi + 0.5) * h + np.array([0.11, 1.0, -0.3])
    v0 = np.array([0.4, -2.0, 1.1])
    st = o.State(x, np.tile(v0, (len(x), 1)), 4e-4)
    st.setup(0.02)
    prm = o.params(1e-4, 0.02, *o.lame(1.4e5, 0.2), 10, 0.025, 0.0075, 0.95)
    cs = [o.make_plane((0, 0, 0), (0, 1, 0), 0.2)]
    xk, vk = x.copy(), v0.copy()
    for _ in range(15):
        st.step(cs, prm)
        vk = vk + prm.dt * np.array([0, -9.81, 0])
        xk = xk + prm.dt * vk
    assert np.abs(st.field("v") - vk).max() < 1e-11
    assert np.abs(st.field("x") - xk).max() < 1e-11
    assert np.abs(st.field("Fe") - np.eye(3).reshape(9)).max() < 1e-11


def test_setup_volume_lattice_interior():
    """8 particles per cell on a regular lattice: interior rho_p = 8 m / dx^3
    (half-spacing sums of the cubic B-spline equal 2 per axis), so V_p0 = h^3."""
    n, h = 12, 0.01
    gi = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = (gi + 0.5) * h + 0.3
    st = o.State(x, np.zeros_like(x), 400 * h ** 3)
    st.setup(2 * h)
    V0 = st.field("V0")
    interior = np.all((gi >= 4) & (gi < n - 4), axis=1)
    assert interior.sum() > 0
    assert np.allclose(V0[interior], h ** 3, rtol=1e-12)
    assert np.all(V0[~interior] > h ** 3 * 0.99)


@pytest.mark.parametrize("seed", range(3))
def test_collision_rotation_equivariance(seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    n = np.array([0, 1.0, 0])
    for v, mu in [((2, -3, 0), 0.2), ((0.2, -3, 0), 0.6), ((2, 3, 0), 1.0), ((1.0, -1.0, 0.5), 0.3)]:
        v = np.array(v, dtype=float)
        x = np.array([0.3, -0.1, 0.2])
        a = o.make_plane((0, 0, 0), n, mu)
        b = o.make_plane((0, 0, 0), Q @ n * 3.0, mu)  # unnormalised on purpose
        out_a, out_b = o.vec(0, 0, 0), o.vec(0, 0, 0)
        o.lib.orc_collision_helper(o.vec(*x), o.vec(*v), a, out_a)
        o.lib.orc_collision_helper(o.vec(*(Q @ x)), o.vec(*(Q @ v)), b, out_b)
        assert np.allclose(Q @ np.array(out_a[:]), np.array(out_b[:]), atol=1e-12)


def test_box_and_sphere_sdf():
    box = o.make_box((0, -0.5, 0), (1, 0.5, 1), 0.2)
    sph = o.make_sphere((1, 2, 3), 0.5, 0.2)
    assert abs(o.lib.orc_collider_phi(box, o.vec(0.2, 0.3, 0.1)) - 0.3) < 1e-15
    assert abs(o.lib.orc_collider_phi(box, o.vec(0.2, -0.1, 0.1)) + 0.1) < 1e-15
    assert abs(o.lib.orc_collider_phi(box, o.vec(1.3, 0.4, 0.0)) - 0.5) < 1e-15
    assert abs(o.lib.orc_collider_phi(sph, o.vec(1, 2, 4)) - 0.5) < 1e-15
    # sliding on the box top behaves like the reference plane (snow_tests.cpp:215-224)
    out = o.vec(0, 0, 0)
    o.lib.orc_collision_helper(o.vec(0, -0.01, 0), o.vec(2, -3, 0), box, out)
    assert np.allclose(out[:], [1.4, 0, 0], atol=1e-12)


def test_position_projection():
    """The optional projection (an extension; the reference only advects,
    grid.cpp:335-338): flagged bodies move inside particles to their closest
    surface point (phi = 0 there), leave outside ones and velocities alone,
    and unflagged bodies change nothing."""
    plane = o.make_plane((0, 0.1, 0), (0, 2, 0), 0.2)
    box = o.make_box((0, -0.5, 0), (1, 0.5, 1), 0.2)
    sph = o.make_sphere((1, 2, 3), 0.5, 0.2)
    for c in (plane, box, sph):
        c.flags = 1
    cases = [(plane, (0.3, 0.05, -0.2), (0.3, 0.1, -0.2)),     # below the plane: up to it
             (box, (0.2, -0.1, 0.3), (0.2, 0.0, 0.3)),          # nearest face is the top
             (box, (0.95, -0.5, 0.1), (1.0, -0.5, 0.1)),        # nearest face is x = 1
             (sph, (1.0, 2.0, 3.2), (1.0, 2.0, 3.5))]           # radially out
    for c, x, want in cases:
        p = o.make_particle(x, (1.0, -2.0, 0.5), 1, 1)
        o.position_projection(p, [c])
        assert np.allclose(p.x[:], want, atol=1e-14), (c.type, p.x[:])
        assert p.v[:] == [1.0, -2.0, 0.5]
        assert abs(o.lib.orc_collider_phi(c, p.x)) < 1e-14
    p = o.make_particle((0.3, 0.5, 0.0), (0, 0, 0), 1, 1)  # outside: unchanged
    o.position_projection(p, [plane])
    assert p.x[:] == [0.3, 0.5, 0.0]
    plane.flags = 0
    p = o.make_particle((0.3, 0.05, 0.0), (0, 0, 0), 1, 1)
    o.position_projection(p, [plane])
    assert p.x[:] == [0.3, 0.05, 0.0]


# ---- whole-step closed forms under contact (pin the order of grid.cpp:402-446) ----

def _block(n=6, h=0.01, lo=(0.013, 0.101, -0.047)):
    gi = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    return (gi + 0.5) * h + np.array(lo)


@pytest.mark.parametrize("vc", [(0.0, 0.0, 0.0), (0.7, 0.0, -0.3)])
def test_step_inside_sticky_halfspace_closed_form(vc):
    """A block with uniform v0 wholly inside a friction-1 half-space (every
    particle and every stencil node below the plane), pressing into it.
    After one step: every node sticks to the body (v_i = v_c), so grad v = 0
    and F_E stays I; G2P gives PIC = v_c and FLIP = v0 + sum w (v_c - v0) =
    v_c; the particle collision sees v - v_c = 0; x advances by dt v_c.
    Reading the grid velocity before grid_collisions (S14) would leave
    v = v0 + dt g; advecting before collisions would move x by dt v0."""
    x = _block()
    v0 = np.array([0.2, -3.0, 0.1])
    st = o.State(x, np.tile(v0, (len(x), 1)), 4e-4)
    st.setup(0.02)
    prm = o.params(1e-4, 0.02, *o.lame(1.4e5, 0.2), 10, 0.025, 0.0075, 0.95)
    cs = [o.make_plane((0, 0.5, 0), (0, 1, 0), 1.0, velocity=vc)]
    st.step(cs, prm)
    assert np.abs(st.field("v") - np.array(vc)).max() < 1e-12
    assert np.abs(st.field("x") - (x + prm.dt * np.array(vc))).max() < 1e-14
    assert np.abs(st.field("Fe") - np.eye(3).reshape(9)).max() < 1e-12


def test_step_inside_frictionless_halfspace_closed_form():
    """The same block in a frictionless moving half-space: grid_collisions
    removes only the normal part of v - v_c (grid.cpp:213-230), so every
    particle leaves the step with v = v_c + tangential part of
    (v0 + dt g - v_c) = (v0_x, v_c,y, v0_z), and x advances by dt times that."""
    x = _block()
    v0 = np.array([0.2, -3.0, 0.1])
    vc = np.array([0.7, 0.05, -0.3])
    st = o.State(x, np.tile(v0, (len(x), 1)), 4e-4)
    st.setup(0.02)
    prm = o.params(1e-4, 0.02, *o.lame(1.4e5, 0.2), 10, 0.025, 0.0075, 0.95)
    st.step([o.make_plane((0, 0.5, 0), (0, 1, 0), 0.0, velocity=vc)], prm)
    want = np.array([v0[0], vc[1], v0[2]])
    assert np.abs(st.field("v") - want).max() < 1e-12
    assert np.abs(st.field("x") - (x + prm.dt * want)).max() < 1e-14


def thin_box_case():
    """Particles around a moving thin box (y in [0.106, 0.114]) that no grid
    node (y = 0.02 j) lies in, so only particle_collisions sees it.  Groups:
    A inside near the top face (outward normal +y), B inside near the bottom
    face (normal -y), C 0.1 mm above the top face (outside before advection,
    inside after), D well outside.  Returns positions, group labels, v0, v_c
    and the collider arguments."""
    xs = np.array([0.013, 0.021, 0.034, 0.046])
    zs = np.array([-0.031, -0.012, 0.007])
    ys = {"A": 0.1135, "B": 0.1066, "C": 0.1141, "D": 0.131}
    pts, lab = [], []
    for g, y in ys.items():
        for a in xs:
            for c in zs:
                pts.append((a, y, c))
                lab.append(g)
    return (np.array(pts), np.array(lab), np.array([0.3, -3.0, 0.2]), np.array([0.5, 0.0, 0.0]),
            ((0.0, 0.11, 0.0), (1.0, 0.004, 1.0), 0.3))


def thin_box_expected(x, lab, v0, vc, dt, g):
    """Closed form of one step for thin_box_case: the grid sees no collider,
    so G2P returns v0 + dt g for every particle; particle_collisions at the
    pre-advection x (grid.cpp:327-333) sticks group A to the box (|v_t| =
    0.28 <= 0.3 * 3) and leaves B (separating from the bottom face), C (not
    yet inside) and D alone; advection uses the collided v (grid.cpp:335-338)."""
    v1 = np.tile(v0 + dt * np.asarray(g), (len(x), 1))
    v1[lab == "A"] = vc
    return x + dt * v1, v1


def test_step_particle_collision_order_closed_form():
    x, lab, v0, vc, (c, he, mu) = thin_box_case()
    st = o.State(x, np.tile(v0, (len(x), 1)), 4e-4)
    st.setup(0.02)
    prm = o.params(1e-4, 0.02, *o.lame(1.4e5, 0.2), 10, 0.025, 0.0075, 0.95)
    st.step([o.make_box(c, he, mu, velocity=vc)], prm)
    xw, vw = thin_box_expected(x, lab, v0, vc, prm.dt, (0, -9.81, 0))
    assert np.abs(st.field("v") - vw).max() < 1e-12
    assert np.abs(st.field("x") - xw).max() < 1e-14
    # group C ends inside the box: colliding after advection would have caught it
    assert (np.abs(st.field("x")[lab == "C"][:, 1] - 0.11) < 0.004).all()


def grid_contact_case():
    """A block whose particles all lie above a sticky moving plane y = 0.105
    while the two lowest node layers of its stencils (y = 0.08, 0.10) lie in
    it: only grid_collisions sees the plane.  Returns x, v0, v_c and the
    plane's arguments."""
    return _block(), np.array([0.2, -1.0, 0.1]), np.array([0.7, 0.0, -0.3]), ((0, 0.105, 0), (0, 1, 0), 1.0)


def grid_contact_expected(x, v0, vc, dt, dx, g, Y0, N, N_dash):
    """Closed form of one step for grid_contact_case.  Every node has
    p/m = v0 and zero force (F_E = I), so after grid_velocity v_i = v0 + dt g;
    grid_collisions sticks the nodes with y_i <= Y0 to the plane (v_i = v_c).
    With D = v_c - v0 - dt g and the x/z weights summing to 1:
      G2P: PIC = FLIP = sum_i w v_i = v0 + dt g + W D,  W = sum_{y_i <= Y0} wy
      stress_update: grad v = D (0, G, 0)^T,  G = sum_{y_i <= Y0} N'_y / dx
    (sum_x dwx = sum_z dwz = 0), so F_E = I + dt D (0, G, 0)^T (no clamp:
    checked by the caller).  The particles lie above the plane, so
    particle_collisions leaves v, and x advances by dt v."""
    gy = x[:, 1] / dx
    base = np.floor(gy).astype(int) - 1
    W = np.zeros(len(x))
    G = np.zeros(len(x))
    for k in range(4):
        j = base + k
        inside = j * dx <= Y0
        W += np.where(inside, [N(a, b) for a, b in zip(j, gy)], 0.0)
        G += np.where(inside, [N_dash(a, b) for a, b in zip(j, gy)], 0.0) / dx
    D = vc - v0 - dt * np.asarray(g)
    v1 = v0 + dt * np.asarray(g) + W[:, None] * D
    F = np.tile(np.eye(3), (len(x), 1, 1))
    F[:, :, 1] += dt * G[:, None] * D
    return x + dt * v1, v1, F, W


def test_step_grid_contact_closed_form():
    """Pins S14 (stress_update and G2P read the post-collision grid velocity):
    reading it before grid_collisions would give v = v0 + dt g and F_E = I."""
    x, v0, vc, (pt, nrm, mu) = grid_contact_case()
    prm = o.params(1e-4, 0.02, *o.lame(1.4e5, 0.2), 10, 0.025, 0.0075, 0.95)
    st = o.State(x, np.tile(v0, (len(x), 1)), 4e-4)
    st.setup(prm.dx)
    st.step([o.make_plane(pt, nrm, mu, velocity=vc)], prm)
    xw, vw, Fw, W = grid_contact_expected(x, v0, vc, prm.dt, prm.dx, (0, -9.81, 0), pt[1], o.N, o.N_dash)
    assert (W > 0.01).any() and (W == 0).any()       # contact reaches some particles, not all
    s = np.linalg.svd(Fw, compute_uv=False)
    assert s.max() < 1 + 0.0075 and s.min() > 1 - 0.025  # no clamp: the closed form applies
    assert np.abs(st.field("v") - vw).max() < 1e-12
    assert np.abs(st.field("x") - xw).max() < 1e-14
    assert np.abs(st.field("Fe").reshape(-1, 3, 3) - Fw).max() < 1e-12
