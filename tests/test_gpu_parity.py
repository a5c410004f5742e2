"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Tolerances: 1e-4 (FP32 GPU vs FP64 oracle) on the metrics of
mpmref.rel_errors: velocity error relative to the scene's largest speed, and
position error beyond FP32 storage rounding relative to the scene's largest
displacement; after 10 steps, and after 20-30 steps for the contact cases.
Bitwise equality where the design guarantees it (repeat runs, permuted
input order, a scene alone vs in a batch).
"""
import os

import numpy as np
import pytest

import mpmref
import snowmpm

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _scene(kind, scale, seed=11, index=0):
    return snowmpm.generate_scene(kind, seed=seed, index=index, scale=scale)


def test_setup_volume_matches_oracle():
    sc = _scene(0, 0.3)
    ref = mpmref.run_oracle(sc, 0)
    got = mpmref.run_gpu(sc, 0)
    rel = np.abs(got["V0"] - ref["V0"]) / np.abs(ref["V0"])
    assert rel.max() < 1e-5, rel.max()


def test_freefall_closed_form():
    """Uniform v, F=I, no contact: each step v += dt g, x += dt v (exactly rigid)."""
    n_side = 8
    h = 0.01
    g = np.stack(np.meshgrid(*[np.arange(n_side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = (g + 0.5) * h + np.array([0.3, 2.0, -0.2])
    v0 = np.array([0.7, -1.5, 0.25])
    v = np.tile(v0, (len(x), 1))
    prm = snowmpm.params(1e-4, 0.02, *snowmpm.lame(1.4e5, 0.2), 10, 0.025, 0.0075, 0.95)
    sc = dict(x=x, v=v, m=np.full(len(x), 4e-4), params=prm,
              colliders=[snowmpm.make_plane((0, 0, 0), (0, 1, 0), 0.2)])
    nsteps = 20
    got = mpmref.run_gpu(sc, nsteps)
    gvec = np.array(prm.gravity[:])
    vk = v0.copy()
    xk = x.copy()
    for _ in range(nsteps):
        vk = vk + prm.dt * gvec
        xk = xk + prm.dt * vk
    assert np.abs(got["v"] - vk).max() < 1e-5
    assert np.abs(got["x"] - xk).max() < 1e-5
    assert np.abs(got["Fe"] - np.eye(3)).max() < 1e-5


@pytest.mark.parametrize("kind,scale,index", [(0, 1.0, 0), (1, 0.15, 0), (2, 0.05, 0), (3, 1.0, 0), (3, 1.0, 1),
                                               (3, 1.0, 2), (3, 0.1, 0)])
def test_ten_steps_match_oracle(kind, scale, index):
    """The 64,000-particle cube and three scenes of the batched distribution
    (46,930, 84,643 and 3,561 particles; box, sphere and torus shapes, random
    velocities up to 8 m/s) at full size, and the two tori, the slab
    (per-particle E, box collider) and a batched scene at reduced size."""
    sc = _scene(kind, scale, index=index)
    ref = mpmref.run_oracle(sc, 10)
    got = mpmref.run_gpu(sc, 10)
    ex, ev = mpmref.rel_errors(got, ref, sc, 10)
    used = mpmref.x_rounding_used(got, ref, sc, 10)
    print(f"kind {kind} n={len(sc['x'])}: rel err x {ex:.3e} v {ev:.3e} (x error {used:.2f} of its FP32 allowance)")
    assert ex < TOL and ev < TOL
    jp_ref = np.linalg.det(ref["Fp"])
    assert np.abs(got["Jp"] - jp_ref).max() < 1e-4
    assert np.abs(got["Fe"] - ref["Fe"]).max() < 1e-4


def test_full_size_tori_ten_steps():
    """The two-tori workload at full size (1,365,824 particles, +-5 m/s):
    10 steps against the oracle, the same bound as the reduced cases (the
    oracle takes about 15 s on 16 cores)."""
    sc = _scene(1, 1.0)
    assert len(sc["x"]) == 1365824
    ref = mpmref.run_oracle(sc, 10)
    got = mpmref.run_gpu(sc, 10)
    ex, ev = mpmref.rel_errors(got, ref, sc, 10)
    print(f"tori n={len(sc['x'])}: rel err x {ex:.3e} v {ev:.3e}")
    assert ex < TOL and ev < TOL
    assert got["status"][0] == 0
    assert np.abs(got["Fe"] - ref["Fe"]).max() < 1e-4


@pytest.mark.timeout(400)  # the oracle needs about two minutes and 32 GB of host memory
def test_full_size_slab_ten_steps():
    """The 8,000,000-particle slab (per-particle E, box collider) at full size:
    10 steps against the oracle."""
    sc = _scene(2, 1.0)
    assert len(sc["x"]) == 8000000
    ref = mpmref.run_oracle(sc, 10)
    got = mpmref.run_gpu(sc, 10)
    ex, ev = mpmref.rel_errors(got, ref, sc, 10)
    print(f"slab n={len(sc['x'])}: rel err x {ex:.3e} v {ev:.3e}")
    assert ex < TOL and ev < TOL
    assert got["status"][0] == 0
    assert np.abs(got["Fe"] - ref["Fe"]).max() < 1e-4


@pytest.mark.parametrize("seed,n,side", [(5, 3000, 0.1), (6, 900, 0.05)])
def test_random_density_matches_oracle(seed, n, side):
    """Uniformly random positions, so cells hold anywhere from 0 to ~50
    particles: P2G's cells of more than 8 particles, cells split across its
    64-particle staging chunks, empty cells, partial x rows and blocks of many
    chunks are all exercised (a jittered 8-per-cell lattice hits none of
    them).  Setup + 10 steps against the oracle."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, side, size=(n, 3)) + np.array([0.31, 0.047, -0.23])
    v = rng.normal(0.0, 0.5, size=(n, 3))
    prm = snowmpm.params(1e-4, 0.02, *snowmpm.lame(1.4e5, 0.2), 10.0, 0.025, 0.0075, 0.95)
    sc = dict(x=x, v=v, m=np.full(n, 4e-4), params=prm, colliders=[snowmpm.make_plane((0, 0, 0), (0, 1, 0), 0.2)])
    cells = np.floor(x / prm.dx).astype(int)
    _, counts = np.unique(cells, axis=0, return_counts=True)
    assert counts.max() > 16, counts.max()  # the long-cell path is reached
    ref = mpmref.run_oracle(sc, 10)
    got = mpmref.run_gpu(sc, 10)
    ex, ev = mpmref.rel_errors(got, ref, sc, 10)
    print(f"random density n={n}, max {counts.max()} per cell: rel err x {ex:.3e} v {ev:.3e}")
    assert ex < TOL and ev < TOL
    assert np.abs(got["V0"] - ref["V0"]).max() / ref["V0"].max() < 1e-5
    assert got["status"][0] == 0


def test_contact_plasticity_matches_oracle():
    """Cube hitting the ground fast enough to plastify: F_E, J_P and v vs oracle."""
    sc = _scene(0, 0.3)
    sc["v"] = np.tile(np.array([1.0, -6.0, 0.5]), (len(sc["x"]), 1))
    sc["x"] = sc["x"] - np.array([0, 0.015, 0])  # bottom layer starts 5 mm above the plane
    nsteps = 30
    ref = mpmref.run_oracle(sc, nsteps)
    got = mpmref.run_gpu(sc, nsteps)
    jp_ref = np.linalg.det(ref["Fp"])
    assert (jp_ref < 0.999).any(), "scenario must plastify"
    ex, ev = mpmref.rel_errors(got, ref, sc, nsteps)
    print(f"contact: rel err x {ex:.3e} v {ev:.3e}, min Jp {jp_ref.min():.4f}")
    assert ex < TOL and ev < TOL
    assert np.abs(got["Jp"] - jp_ref).max() < 1e-3


def test_eight_colliders_last_one_hit():
    """The contact case with the ground plane as the last of eight colliders
    (seven far-away spheres and boxes first): every collider of the scene
    record reaches the particle and grid collision loops, in list order."""
    sc = _scene(0, 0.3)
    sc["v"] = np.tile(np.array([1.0, -6.0, 0.5]), (len(sc["x"]), 1))
    sc["x"] = sc["x"] - np.array([0, 0.015, 0])
    far = [snowmpm.make_sphere((5.0 + k, 3.0, -4.0), 0.2 + 0.1 * k, 0.1 * k) for k in range(4)]
    far += [snowmpm.make_box((-6.0, 2.0 + k, 7.0), (0.3, 0.2, 0.1 * (k + 1)), 0.3) for k in range(3)]
    sc["colliders"] = far + list(sc["colliders"])
    assert len(sc["colliders"]) == 8
    nsteps = 30
    ref = mpmref.run_oracle(sc, nsteps)
    got = mpmref.run_gpu(sc, nsteps)
    jp_ref = np.linalg.det(ref["Fp"])
    assert (jp_ref < 0.999).any(), "scenario must plastify"
    ex, ev = mpmref.rel_errors(got, ref, sc, nsteps)
    print(f"eight colliders: rel err x {ex:.3e} v {ev:.3e}")
    assert ex < TOL and ev < TOL
    assert np.abs(got["Jp"] - jp_ref).max() < 1e-3


def test_sparse_blocks_exceed_the_grid():
    """21,952 particles 0.1 m apart, each in its own home block: more home
    blocks than the P2G and G2P grids hold workgroups (the persistent-loop
    path), random velocities and a ground plane, against the oracle."""
    rng = np.random.default_rng(8)
    g = np.stack(np.meshgrid(*[np.arange(28)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = g * 0.1 + np.array([-1.4, 0.05, -1.4]) + rng.uniform(-0.01, 0.01, (len(g), 3))
    v = rng.uniform(-3, 3, x.shape)
    prm = snowmpm.params(1e-4, 0.02, *snowmpm.lame(1.4e5, 0.2), 10, 0.025, 0.0075, 0.95)
    sc = dict(x=x, v=v, m=np.full(len(x), 4e-4), params=prm,
              colliders=[snowmpm.make_plane((0, 0, 0), (0, 1, 0), 0.2)])
    ref = mpmref.run_oracle(sc, 5)
    got = mpmref.run_gpu(sc, 5)
    ex, ev = mpmref.rel_errors(got, ref, sc, 5)
    print(f"sparse: {len(x)} particles, rel err x {ex:.3e} v {ev:.3e}")
    assert ex < TOL and ev < TOL
    assert got["status"][0] == 0


def test_slab_contact_matches_oracle():
    """The slab workload (per-particle E, box collider with friction 0.2)
    driven into the box top: contact, friction and plasticity against the
    oracle."""
    sc = _scene(2, 0.05)
    sc["x"] = sc["x"] - np.array([0, sc["x"][:, 1].min() - 0.005, 0])
    sc["v"] = np.tile(np.array([0.4, -4.0, -0.3]), (len(sc["x"]), 1))
    nsteps = 30
    ref = mpmref.run_oracle(sc, nsteps)
    got = mpmref.run_gpu(sc, nsteps)
    jp_ref = np.linalg.det(ref["Fp"])
    assert (jp_ref < 0.999).any(), "scenario must plastify"
    ex, ev = mpmref.rel_errors(got, ref, sc, nsteps)
    print(f"slab contact: rel err x {ex:.3e} v {ev:.3e}, min Jp {jp_ref.min():.4f}")
    assert ex < TOL and ev < TOL
    assert np.abs(got["Jp"] - jp_ref).max() < 1e-3


def test_position_projection_matches_oracle():
    """Colliders flagged for position projection (an extension; the reference
    only changes velocities): a cube that starts 4 mm inside the ground and
    7 mm inside a sphere cap, GPU vs oracle, and no particle left inside a
    flagged body.  (The scene is well conditioned: a 1e-7 relative
    perturbation of the input moves the oracle's result by 2.5e-7.  A
    projection of many particles onto one small patch is not, so it is no
    parity case.)"""
    sc = _scene(0, 0.3)
    sc["v"] = np.tile(np.array([0.5, -1.0, 0.3]), (len(sc["x"]), 1))
    sc["x"] = sc["x"] - np.array([0, sc["x"][:, 1].min() + 0.004, 0])
    c0 = sc["x"].mean(axis=0)
    centre, radius = (c0[0] + 0.03, 0.004 - 0.05, c0[2] - 0.02), 0.05
    sc["colliders"] = [snowmpm.make_plane((0, 0, 0), (0, 1, 0), 0.2, project=True),
                       snowmpm.make_sphere(centre, radius, 0.3, project=True)]
    nsteps = 20
    ref = mpmref.run_oracle(sc, nsteps)
    got = mpmref.run_gpu(sc, nsteps)
    plain = dict(sc, colliders=[snowmpm.make_plane((0, 0, 0), (0, 1, 0), 0.2),
                                snowmpm.make_sphere(centre, radius, 0.3)])
    ref_plain = mpmref.run_oracle(plain, nsteps)

    def depth(x):
        return np.minimum(x[:, 1], np.linalg.norm(x - np.array(centre), axis=1) - radius)

    assert depth(ref_plain["x"]).min() < -3e-3, "without projection particles must end inside"
    assert depth(ref["x"]).min() > 0.0
    assert depth(got["x"]).min() > -1e-6
    ex, ev = mpmref.rel_errors(got, ref, sc, nsteps)
    print(f"projection: rel err x {ex:.3e} v {ev:.3e}")
    assert ex < TOL and ev < TOL
    assert got["status"][0] == 0


def test_track_fp_matches_oracle():
    sc = _scene(0, 0.3)
    sc["v"] = np.tile(np.array([0.0, -6.0, 0.0]), (len(sc["x"]), 1))
    sc["x"] = sc["x"] - np.array([0, 0.015, 0])
    ref = mpmref.run_oracle(sc, 20)
    got = mpmref.run_gpu(sc, 20, track_fp=True)
    assert np.abs(got["Fp"] - ref["Fp"]).max() < 1e-3
    assert np.abs(np.linalg.det(got["Fp"]) - got["Jp"]).max() < 1e-4


def test_deterministic_bitwise():
    sc = _scene(1, 0.12)
    a = mpmref.run_gpu(sc, 8)
    b = mpmref.run_gpu(sc, 8)
    for k in ("x", "v", "Fe", "Jp"):
        assert np.array_equal(a[k], b[k]), k


def test_permutation_invariant():
    """Input order barely matters: cells are summed in particle-id order within
    a block and the block partials in a fixed order; only the particle ids
    (the tie-break inside a cell) change with the input order."""
    sc = _scene(0, 0.3)
    a = mpmref.run_gpu(sc, 6)
    rng = np.random.default_rng(3)
    perm = rng.permutation(len(sc["x"]))
    sc2 = dict(sc, x=sc["x"][perm], v=sc["v"][perm], m=sc["m"][perm])
    b = mpmref.run_gpu(sc2, 6)
    inv = np.argsort(perm)
    # particle ids differ under permutation, so compare within tolerance
    assert np.abs(b["x"][inv] - a["x"]).max() < 1e-6
    assert np.abs(b["v"][inv] - a["v"]).max() < 1e-5


def test_batch_equals_alone_bitwise():
    s1 = _scene(3, 0.1, index=1)
    s2 = _scene(3, 0.1, index=2)
    a1 = mpmref.run_gpu(s1, 7)
    a2 = mpmref.run_gpu(s2, 7)
    with snowmpm.Engine(capacity=len(s1["x"]) + len(s2["x"]) + 64) as eng:
        i1 = eng.add_scene(s1["params"], s1["colliders"], s1["x"], s1["v"], s1["m"])
        i2 = eng.add_scene(s2["params"], s2["colliders"], s2["x"], s2["v"], s2["m"])
        eng.step(7)
        b1, b2 = eng.download(i1), eng.download(i2)
    for k in ("x", "v", "Fe", "Jp"):
        assert np.array_equal(a1[k], b1[k]), k
        assert np.array_equal(a2[k], b2[k]), k


def test_conservation_sums():
    """P2G/G2P conserve mass; free-flight momentum changes only by gravity."""
    sc = _scene(0, 0.3)
    sc["x"] = sc["x"] + np.array([0, 1.0, 0])  # far from the ground
    with snowmpm.Engine(capacity=len(sc["x"]) + 64) as eng:
        sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
        eng.setup()
        s0 = eng.sums(sid)
        eng.step(10)
        s1 = eng.sums(sid)
    M = sc["m"].sum()
    assert abs(s0[0] - M) / M < 1e-6
    dp = (s1[4:7] - s0[4:7]) / M
    expect = 10 * sc["params"].dt * np.array(sc["params"].gravity[:])
    assert np.abs(dp - expect).max() < 1e-4


def test_nan_is_reported():
    sc = _scene(0, 0.2)
    sc["v"] = sc["v"].copy()
    sc["v"][5] = np.nan
    with snowmpm.Engine(capacity=len(sc["x"]) + 64) as eng:
        sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
        rc = eng.step_rc(2)
        st, _ = eng.status(sid)
    assert rc == snowmpm.ERR_DEVICE
    assert st & snowmpm.STATUS_NAN


def _block_lattice(n, dx=0.02):
    """One particle per 4^3-node block of an n x n x n arrangement, 1.04 m (13
    blocks) up: its stencil base sits at node 1 of its block, so the scene has
    n^3 home blocks and (n + 1)^3 blocks in all."""
    i = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = (4 * i + 2.5) * dx + np.array([0.0, 1.04, 0.0])
    prm = snowmpm.params(1e-4, dx, *snowmpm.lame(1.4e5, 0.2), 10, 0.025, 0.0075, 0.95)
    return dict(x=x, v=np.zeros_like(x), m=np.full(len(x), 4e-4), params=prm,
                colliders=[snowmpm.make_plane((0, 0, 0), (0, 1, 0), 0.2)])


def test_hash_size_does_not_change_bits():
    """Block hash slots are assigned by insertion races, but nothing summed
    depends on them: a 128-slot table (the cube's 64 blocks fill half of it,
    long probe chains) gives the same bits as the default 65,536 slots."""
    sc = _scene(0, 0.5)
    out = []
    for hl in (7, 0):
        with snowmpm.Engine(capacity=len(sc["x"]) + 64, hash_log2=hl) as eng:
            sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
            eng.step(10)
            out.append(eng.download(sid))
            assert eng.status(sid)[0] == 0
    for k in ("x", "v", "Fe", "Jp"):
        assert np.array_equal(out[0][k], out[1][k]), k


@pytest.mark.parametrize("hash_log2,expect,absent", [(9, snowmpm.STATUS_OVERFLOW, snowmpm.STATUS_HASH_FULL),
                                                     (8, snowmpm.STATUS_HASH_FULL, 0)])
def test_small_hash_is_reported(hash_log2, expect, absent):
    """343 home blocks, 512 blocks in all.  With 512 slots every block is
    inserted but the P2G partial sums hold only 256 home blocks (overflow);
    with 256 slots the table fills.  Either way the step fails with
    ERR_DEVICE and the scene's status names the cause."""
    sc = _block_lattice(7)
    with snowmpm.Engine(capacity=len(sc["x"]) + 64, hash_log2=hash_log2) as eng:
        sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
        rc = eng.step_rc(1)
        st, _ = eng.status(sid)
    assert rc == snowmpm.ERR_DEVICE
    assert st & expect
    assert not st & absent


def test_far_particle_is_reported():
    """A particle beyond the block-coordinate range (|x| > ~5 km at dx 0.02)
    is left out of the grid and the step reports MPM_STATUS_RANGE."""
    sc = _scene(0, 0.2)
    sc["x"] = sc["x"].copy()
    sc["x"][7] = (1.0e4, 1.0, 0.0)
    with snowmpm.Engine(capacity=len(sc["x"]) + 64) as eng:
        sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
        rc = eng.step_rc(1)
        st, _ = eng.status(sid)
    assert rc == snowmpm.ERR_DEVICE
    assert st & snowmpm.STATUS_RANGE


def test_capacity_is_enforced():
    sc = _scene(0, 0.2)
    with snowmpm.Engine(capacity=len(sc["x"]) - 1) as eng:
        with pytest.raises(snowmpm.MpmError) as e:
            eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
        assert e.value.code == snowmpm.ERR_CAPACITY
        half = len(sc["x"]) // 2
        sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"][:half], sc["v"][:half], sc["m"][:half])
        eng.step(2)
        assert eng.status(sid)[0] == 0


def test_fp64_fallback_matches_oracle():
    """cond(trial F_E) = 1.3/0.8 > 1.5 sends every particle's first plasticity
    update (and the upload's stress) through the FP64 SVD; the clamp then
    brings F_E back into range, and the result still matches the oracle."""
    sc = _scene(0, 0.25)
    x = sc["x"]
    n = len(x)
    sc["x"] = x + np.array([0, 1.0, 0])
    sc["v"] = np.stack([40.0 * (x[:, 1] - x[:, 1].mean()), np.zeros(n), np.zeros(n)], 1)
    Fe = np.tile(np.diag([1.3, 0.8, 1.0]).reshape(9), (n, 1))
    ref = mpmref.run_oracle(sc, 3, Fe=Fe)
    got = mpmref.run_gpu(sc, 3, Fe=Fe)
    _, fb = got["status"]
    assert fb >= n
    ex, ev = mpmref.rel_errors(got, ref, sc, 3)
    print(f"fallbacks {fb}: rel err x {ex:.3e} v {ev:.3e}")
    assert ex < TOL and ev < TOL
    assert np.abs(got["Fe"] - ref["Fe"]).max() < 1e-4
    assert np.abs(got["Jp"] - np.linalg.det(ref["Fp"])).max() < 1e-4


def test_prestressed_elastic_matches_oracle():
    """Rotated, stretched F_E inside the clamp range (elastic fast path on the
    GPU, SVD every step in the oracle): the boundary accelerates from the
    stress alone; x, v and F_E must match."""
    sc = _scene(0, 0.3)
    n = len(sc["x"])
    sc["x"] = sc["x"] + np.array([0, 1.0, 0])
    sc["v"] = np.zeros((n, 3))
    th = 0.4
    Q = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    Fe = Q @ np.diag([1.006, 0.985, 0.992])
    Fe = np.tile(Fe.reshape(9), (n, 1))
    ref = mpmref.run_oracle(sc, 10, Fe=Fe)
    got = mpmref.run_gpu(sc, 10, Fe=Fe)
    vref = ref["v"]
    assert np.abs(vref[:, 0]).max() > 1e-3  # the stress does accelerate the boundary
    dv = np.abs(got["v"] - vref).max() / np.abs(vref).max()
    print(f"prestressed: max |dv|/max|v| {dv:.3e}")
    assert dv < 1e-3
    assert np.abs(got["Fe"] - ref["Fe"]).max() < 1e-5
    assert np.abs(got["x"] - ref["x"]).max() < 1e-6


@pytest.mark.parametrize("track_fp", [False, True])
def test_golden_trajectory(track_fp):
    """The committed oracle trajectory (tests/golden/oracle_trajectory.npz):
    1,000 particles in a ground/tilted-plane wedge, setup + 10 steps."""
    import os
    ref = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_trajectory.npz"))
    s = 2 ** -0.5
    prm = snowmpm.params(1e-4, 0.02, *snowmpm.lame(1.4e5, 0.2), 10.0, 0.025, 0.0075, 0.95)
    sc = dict(x=ref["x0"], v=ref["v0"], m=ref["m"], params=prm,
              colliders=[snowmpm.make_plane((0, 0, 0), (0, 1, 0), 0.2),
                         snowmpm.make_plane((0.09, 0, 0), (-s, s, 0), 0.5)])
    got = mpmref.run_gpu(sc, 10, track_fp=track_fp)
    ex, ev = mpmref.rel_errors(got, dict(x=ref["x_10"], v=ref["v_10"]), sc, 10)
    print(f"golden trajectory (track_fp={track_fp}): rel err x {ex:.3e} v {ev:.3e}")
    assert ex < TOL and ev < TOL
    assert np.abs(got["V0"] - ref["V0"]).max() / ref["V0"].max() < 1e-5
    assert np.abs(got["Fe"] - ref["Fe_10"].reshape(-1, 3, 3)).max() < 1e-4
    assert np.abs(got["Jp"] - np.linalg.det(ref["Fp_10"].reshape(-1, 3, 3))).max() < 1e-4
    if track_fp:
        assert np.abs(got["Fp"] - ref["Fp_10"].reshape(-1, 3, 3)).max() < 1e-4
    assert got["status"][0] == 0


@pytest.mark.parametrize("track_fp", [False, True])
def test_permute_period_does_not_change_bits(track_fp):
    """Steps between full re-sorts update each particle at its buffer index;
    P2G order comes from the sort, not the buffer layout, so every re-sort
    period (1 = every step, 7, the default 16) gives the same bits, through
    contact and plastic flow and across the re-sort steps."""
    sc = _scene(0, 0.3)
    sc["v"] = np.tile(np.array([1.0, -6.0, 0.5]), (len(sc["x"]), 1))
    sc["x"] = sc["x"] - np.array([0, 0.015, 0])
    out = []
    for period in (1, 7, 0):
        with snowmpm.Engine(capacity=len(sc["x"]) + 64, track_fp=track_fp, permute_period=period) as eng:
            sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
            eng.step(11)
            eng.step_graph(12)
            eng.step(17)
            out.append(eng.download(sid))
            assert eng.status(sid)[0] == 0
    keys = ("x", "v", "Fe", "Jp", "V0") + (("Fp",) if track_fp else ())
    for o in out[1:]:
        for k in keys:
            assert np.array_equal(out[0][k], o[k]), k
