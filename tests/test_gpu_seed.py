"""GPU mesh seeding (mpm_seed_mesh through the C ABI) against the oracle.

The seeding kernel replaces the reference's Blender add-on seeding
(blender/seed_export.py:43-131).  At parity sizes its point set must equal
the CPU restatement's (oracle/seed_oracle.cpp) bit for bit: same lattice,
same float32 coordinates, same inside decisions, same order (x outer, z
inner).  Cases: the 64,000-particle cube, a box with a cavity whose faces
pass through lattice points, convex and non-convex meshes, a jittered
(generic) sphere, an open triangle soup, more stacked faces than the
256-cast cap, and empty results.  At a large size (a 9,216-triangle torus
on a 1.3M-point lattice) the result is checked by properties: particle count
against the polygonal torus's volume, and a sample of points against the
generalised winding number.
"""
import numpy as np
import pytest

import meshes
import oracle as orc
import snowmpm
import snowmpm.seed as seed

pytestmark = pytest.mark.gpu


def _same(v, f, spacing):
    got = snowmpm.seed_mesh(v, f, spacing)
    ref = orc.seed_mesh(v, f, spacing)
    print(f"seed: {f.shape[0]} triangles, spacing {spacing}: {got.shape[0]} points", flush=True)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)
    return got


def test_cube_64k():
    v, f = meshes.box((-0.2, 0.02, -0.2), (0.2, 0.42, 0.2))
    assert _same(v, f, 0.01).shape[0] == 64000


def test_cavity_faces_through_lattice_points():
    v, f = meshes.cavity_box((0, 0, 0), (1, 1, 1), (0.35, 0.25, 0.45), (0.65, 0.75, 0.55))
    _same(v, f, 0.1)
    _same(v, f, 0.05)


@pytest.mark.parametrize("name", ["octahedron", "icosphere2", "icosphere3", "torus"])
def test_closed_meshes(name):
    v, f = {"octahedron": lambda: meshes.octahedron((0.01, -0.02, 0.03), 0.5),
            "icosphere2": lambda: meshes.icosphere((0.1, 0.2, -0.1), 0.4, 2),
            "icosphere3": lambda: meshes.icosphere((0.0, 0.0, 0.0), 0.35, 3),
            "torus": lambda: meshes.torus((0.0, 0.1, 0.0), 0.3, 0.12, 24, 12)}[name]()
    _same(v, f, 0.03)


def test_jittered_sphere_and_triangle_soup():
    rng = np.random.default_rng(5)
    v, f = meshes.icosphere((0.0, 0.0, 0.0), 0.3, 2)
    v = v + rng.normal(0, 0.01, v.shape)
    _same(v, f, 0.025)
    # an open soup of random triangles: the same counting rule, no inside
    # notion needed for parity
    soup = rng.uniform(-0.5, 0.5, (60, 3))
    _same(soup, rng.integers(0, 60, (40, 3)).astype(np.int32), 0.05)


def test_more_faces_than_the_cast_cap():
    # 130 wide slabs: rays from the lower points cross more than 256 faces
    slabs = [meshes.box((-1, -1, 0.01 * i), (1, 1, 0.01 * i + 0.004)) for i in range(130)]
    m = meshes.merge(*slabs)
    assert orc.seed_ray_hits(*m, (0.0, 0.0, 0.002), 2) == 256
    _same(*m, 0.4)


def test_empty_results_and_bad_arguments():
    v, f = meshes.box((0, 0, 0), (1, 1, 1))
    assert snowmpm.seed_mesh(v, f, 5.0).shape == (0, 3)
    assert snowmpm.seed_mesh(v, np.zeros((0, 3), np.int32), 0.1).shape == (0, 3)
    with pytest.raises(snowmpm.MpmError) as e:
        snowmpm.seed_mesh(v, np.array([[0, 1, 8]]), 0.1)
    assert e.value.code == snowmpm.ERR_ARG
    with pytest.raises(snowmpm.MpmError):
        snowmpm.seed_mesh(v, f, -0.1)


def test_large_torus_properties():
    R, r = 0.3, 0.12
    v, f = meshes.torus((0.0, 0.0, 0.0), R, r, 96, 48)
    s = 0.005
    got = snowmpm.seed_mesh(v, f, s)
    # volume of the polygonal torus: 48-gon tube section swept by a 96-gon
    sec = 0.5 * 48 * r * r * np.sin(2 * np.pi / 48)
    vol = sec * 2 * np.pi * R * (96 / (2 * np.pi)) * np.sin(2 * np.pi / 96)
    print(f"seed: large torus {got.shape[0]} points, expected ~{vol / s ** 3:.0f}", flush=True)
    assert abs(got.shape[0] * s ** 3 / vol - 1.0) < 0.01
    rng = np.random.default_rng(2)
    _, pts = meshes.lattice(v, s)
    sample = pts[rng.choice(pts.shape[0], 1500, replace=False)]
    got_set = {tuple(p) for p in got.tolist()}
    inside = np.abs(meshes.winding(v, f, sample)) > 0.5
    for p, w in zip(sample, inside):
        if (tuple(p) in got_set) != w:
            assert meshes.surface_dist(v, f, p) < 1e-5


def test_cli_end_to_end(tmp_path):
    # a box in Blender's z-up frame; the input files are in the solver frame
    v, f = meshes.box((-0.1, -0.1, 0.0), (0.1, 0.1, 0.15))
    obj = tmp_path / "Snow.obj"
    obj.write_text("".join(f"v {a} {b} {c}\n" for a, b, c in v) + "".join(f"f {a + 1} {b + 1} {c + 1}\n" for a, b, c in f))
    out = tmp_path / "input"
    rc = seed.main([str(out), "--group", str(obj), "velocity=0,0,-2", "mass=4e-4", "spacing=0.02"])
    assert rc == 0
    x, vel, m = snowmpm.import_particles(str(out / "groups.csv"), str(out / "particles.csv"))
    ref = orc.seed_mesh(v, f, 0.02)
    assert x.shape == ref.shape and ref.shape[0] >= 700
    assert np.array_equal(x, np.stack([ref[:, 0], ref[:, 2], -ref[:, 1]], axis=1))
    assert np.all(vel == np.array([0.0, -2.0, -0.0]))
    assert np.all(m == float(np.float32(4e-4)))
