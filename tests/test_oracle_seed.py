"""The seeding oracle (oracle/seed_oracle.cpp) pinned without Blender.

The reference's seeder is a Blender add-on (blender/seed_export.py:43-131)
whose ray caster is Blender's own; bpy is not available here, so the
restatement is pinned by what the mathematics fixes:
  * lattice coordinates against the add-on's own loop, restated in plain
    Python (float32 bounds and spacing, double accumulation);
  * inside/outside against closed forms (box, box with a cavity, octahedron)
    and against the generalised winding number (icosphere, torus), which
    classifies points without rays;
  * ray hit counts against constructions with known answers: one plane,
    a box from inside, duplicate faces and faces closer than the re-cast
    epsilon (counted once), faces farther apart (counted twice), a ray along
    a shared edge (counted exactly once), and the 256-cast cap.
Points within 1e-5 of the surface are set aside when comparing against the
float64 closed forms; the GPU is compared with this oracle bit for bit.
"""
import numpy as np
import pytest

import meshes
import oracle as orc


def test_axis_matches_addon_loop():
    rng = np.random.default_rng(3)
    for _ in range(200):
        lo = float(np.float32(rng.uniform(-3, 3)))
        hi = float(np.float32(lo + rng.uniform(0, 2)))
        s = float(rng.choice([0.005, 0.01, 0.02, 0.05, 0.1, 0.3, 0.0137]))
        got = orc.seed_axis(lo, hi, s)
        s32, ref, x = float(np.float32(s)), [], lo + float(np.float32(s)) * 0.5
        while x < hi:
            ref.append(float(np.float32(x)))
            x += s32
        assert got.tolist() == ref


def test_directions_are_the_addons_unit_vectors():
    raw = np.array([[1.0, 0.371, 0.529], [-0.427, 1.0, 0.233], [0.191, -0.463, 1.0]])
    for k in range(3):
        d = orc.seed_direction(k).astype(float)
        ref = raw[k] / np.linalg.norm(raw[k])
        assert np.abs(d - ref).max() < 2e-7
        assert abs(np.linalg.norm(d) - 1.0) < 2e-7


def _plane_square(z, half=10.0, diag_first=True):
    v = np.array([[-half, -half, z], [half, -half, z], [half, half, z], [-half, half, z]])
    f = np.array([[0, 1, 2], [0, 2, 3]] if diag_first else [[0, 1, 3], [1, 2, 3]], dtype=np.int32)
    return v, f


def test_plane_hits():
    v, f = _plane_square(0.3)
    for k in range(3):  # every direction points up (positive z component)
        assert orc.seed_ray_hits(v, f, (0.1, -0.2, 0.0), k) == 1
        assert orc.seed_ray_hits(v, f, (0.1, -0.2, 0.5), k) == 0


def test_box_from_inside_and_outside():
    v, f = meshes.box((0, 0, 0), (1, 1, 1))
    for k in range(3):
        assert orc.seed_ray_hits(v, f, (0.5, 0.4, 0.3), k) == 1
        assert orc.seed_ray_hits(v, f, (-2.0, 0.5, 0.5), k) in (0, 2)
        assert orc.seed_ray_hits(v, f, (5.0, 5.0, 5.0), k) == 0


def test_duplicate_and_close_faces_count_once():
    v1, f1 = _plane_square(0.3)
    dup = meshes.merge((v1, f1), (v1, f1))
    assert all(orc.seed_ray_hits(*dup, (0.0, 0.0, 0.0), k) == 1 for k in range(3))
    # 1e-6 apart along z (at most 4.8e-6 along the rays, whose z components
    # are >= 0.2): closer than the re-cast epsilon
    close = meshes.merge(_plane_square(0.3), _plane_square(0.3 + 1e-6))
    assert all(orc.seed_ray_hits(*close, (0.0, 0.0, 0.0), k) == 1 for k in range(3))
    # 1e-3 apart: two crossings
    far = meshes.merge(_plane_square(0.3), _plane_square(0.301))
    assert all(orc.seed_ray_hits(*far, (0.0, 0.0, 0.0), k) == 2 for k in range(3))


def test_shared_edge_is_watertight_and_counted_once():
    # points whose rays pass on, or within rounding of, the shared diagonal
    for diag_first in (True, False):
        v, f = _plane_square(0.25, half=1.0, diag_first=diag_first)
        rng = np.random.default_rng(11)
        for _ in range(300):
            k = int(rng.integers(3))
            d = orc.seed_direction(k).astype(float)
            # the point whose ray meets the plane at (s, s) or (s, -s)
            s = float(rng.uniform(-0.9, 0.9))
            target = np.array([s, s if diag_first else -s, 0.25])
            p = np.float32(target - d * (0.25 / d[2]) + rng.normal(0, 1e-7, 3) * (rng.random() < 0.5))
            assert orc.seed_ray_hits(v, f, p.astype(float), k) == 1


def test_cast_cap_of_256():
    slabs = [meshes.box((-1, -1, 0.01 * i), (1, 1, 0.01 * i + 0.004)) for i in range(150)]
    m = meshes.merge(*slabs)  # 300 faces crossed straight up
    assert orc.seed_ray_hits(*m, (0.0, 0.0, -0.5), 2) == 256


def _lattice_inside(verts, tris, spacing, inside_fn):
    _, pts = meshes.lattice(verts, spacing)
    return pts, inside_fn(pts)


def test_box_seeds_its_whole_lattice():
    for flip in (False, True):
        v, f = meshes.box((-0.2, 0.02, -0.2), (0.2, 0.42, 0.2), flip=flip)
        got = orc.seed_mesh(v, f, 0.01)
        _, pts = meshes.lattice(v, 0.01)
        assert got.shape == (64000, 3)  # the 40^3 cube of the 64k-particle configuration
        assert np.array_equal(got, pts)


def test_cavity_box():
    # cavity faces at x, y, z = 0.3 / 0.7 fall between lattice points (0.05 + 0.1 k)
    v, f = meshes.cavity_box((0, 0, 0), (1, 1, 1), (0.3, 0.3, 0.3), (0.7, 0.7, 0.7))
    got = orc.seed_mesh(v, f, 0.1)
    _, pts = meshes.lattice(v, 0.1)
    in_cav = np.all((pts > 0.3) & (pts < 0.7), axis=1)
    assert np.array_equal(got, pts[~in_cav])
    assert got.shape[0] == 1000 - 64


def _compare_with_classifier(v, f, spacing, inside, tol=1e-5):
    got = orc.seed_mesh(v, f, spacing)
    _, pts = meshes.lattice(v, spacing)
    ref = inside(pts)
    got_set = {tuple(p) for p in got.tolist()}
    g = np.array([tuple(p) in got_set for p in pts.tolist()])
    bad = np.nonzero(g != ref)[0]
    for i in bad:  # only points within rounding of the surface may differ
        assert meshes.surface_dist(v, f, pts[i]) < tol, (pts[i], g[i], ref[i])
    assert ref.sum() > 100
    return got


def test_octahedron_closed_form():
    v, f = meshes.octahedron((0.01, -0.02, 0.03), 0.5)
    c = np.array([0.01, -0.02, 0.03])
    _compare_with_classifier(v, f, 0.03, lambda p: np.abs(p - c).sum(axis=1) < 0.5)


def test_icosphere_matches_winding_number():
    v, f = meshes.icosphere((0.1, 0.2, -0.1), 0.4, 2)
    _compare_with_classifier(v, f, 0.04, lambda p: np.abs(meshes.winding(v, f, p)) > 0.5)


def test_torus_matches_winding_number():
    v, f = meshes.torus((0.0, 0.0, 0.0), 0.3, 0.12, 24, 12)
    got = _compare_with_classifier(v, f, 0.03, lambda p: np.abs(meshes.winding(v, f, p)) > 0.5)
    # the hole stays empty
    assert not np.any(np.hypot(got[:, 0], got[:, 1]) < 0.3 - 0.12 - 0.01)


def test_empty_and_bad_input():
    v, f = meshes.box((0, 0, 0), (1, 1, 1))
    assert orc.seed_mesh(v, f, 5.0).shape == (0, 3)  # spacing/2 beyond the bounds
    assert orc.seed_mesh(np.zeros((0, 3)), np.zeros((0, 3), np.int32), 0.1).shape == (0, 3)
    with pytest.raises(ValueError):
        orc.seed_mesh(v, np.array([[0, 1, 8]]), 0.1)
    with pytest.raises(ValueError):
        orc.seed_mesh(v, f, 0.0)
