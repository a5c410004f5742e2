"""Host side of the seeding tool (snowmpm.seed): OBJ reading and the input
files it writes, checked without a GPU.

The writer follows the reference's Blender add-on (blender/seed_export.py):
groups sorted by object name (148-163), groups.csv "group_id,vx,vy,vz,mass"
and particles.csv "group_id,x,y,z" through csv.writer with "\\n" line ends
(165-255), float32 settings, and the Blender -> solver frame change
(x, y, z) -> (x, z, -y) (142-147).  The seeding itself is replaced here by a
stub returning fixed points; the GPU seeding is tested in test_gpu_seed.py.
The written files are read back with the product's importer, which restates
src/import.cpp.
"""
import numpy as np
import pytest

import snowmpm
import snowmpm.seed as seed


def _obj(path, text):
    path.write_text(text)
    return str(path)


def test_read_obj_forms(tmp_path):
    p = _obj(tmp_path / "m.obj", """# comment
mtllib m.mtl
o Snowball
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0 1.0
vt 0 0
vn 0 0 1
s off
f 1/1/1 2/1/1 3/1/1 4/1/1
f -4//1 -2//1 -1//1
g part
f 1 3 2  # trailing comment
""")
    v, t, name = seed.read_obj(p)
    assert name == "Snowball"
    assert v.shape == (4, 3) and v[3].tolist() == [0, 1, 0]
    assert t.tolist() == [[0, 1, 2], [0, 2, 3], [0, 2, 3], [0, 2, 1]]


def test_read_obj_name_defaults_to_stem(tmp_path):
    p = _obj(tmp_path / "Cube.001.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    assert seed.read_obj(p)[2] == "Cube.001"


@pytest.mark.parametrize("text", ["v 0 0\n", "v 0 0 x\n", "v 0 0 0\nf 1 2\n", "v 0 0 0\nf 1 1 a\n",
                                  "v 0 0 0\nf 1 1 2\n", "v 0 0 0\nf 0 1 1\n", "v 0 0 0\nf -2 1 1\n"])
def test_read_obj_errors(tmp_path, text):
    with pytest.raises(ValueError):
        seed.read_obj(_obj(tmp_path / "bad.obj", text))


def test_settings_are_float32_and_clamped():
    g = seed.Group(np.zeros((3, 3)), [[0, 1, 2]], "a", velocity=(0.1, -2, 0), mass=0.05, spacing=1e-6)
    assert g.mass == 0.05000000074505806
    assert g.velocity == (float(np.float32(0.1)), -2.0, 0.0)
    assert g.spacing == float(np.float32(0.0001))
    assert seed.Group(np.zeros((3, 3)), [[0, 1, 2]], "a", mass=0.0).mass == float(np.float32(1e-6))
    with pytest.raises(ValueError):
        seed.Group(np.zeros((3, 3)), [[0, 1, 2]], "a", mass=float("nan"))


def test_writer_format_and_roundtrip(tmp_path, monkeypatch):
    x01 = float(np.float32(0.1))  # the library returns float32 values
    pts = {"b": np.array([[0.125, 0.25, 0.0]]), "a": np.array([[1.0, 2.0, 3.0], [x01, 0.0, -0.5]])}
    calls = []

    def fake_seed(verts, tris, spacing, device=0):
        name = "a" if verts[0, 0] == 1.0 else "b"
        calls.append((name, spacing, device))
        return pts[name]

    monkeypatch.setattr(seed, "seed_mesh", fake_seed)
    ga = seed.Group(np.ones((3, 3)), [[0, 1, 2]], "a", velocity=(1, 2, 3), mass=0.25, spacing=0.02)
    gb = seed.Group(np.zeros((3, 3)), [[0, 1, 2]], "b", mass=0.05)
    log = []
    out = tmp_path / "deep" / "input"
    total = seed.seed_groups([gb, ga], out, log=log.append)  # sorted by name: a, b
    assert total == 3
    assert calls == [("a", float(np.float32(0.02)), 0), ("b", float(np.float32(0.05)), 0)]
    assert log == ["Group 0 'a': 2 particles", "Group 1 'b': 1 particles", "Exported 2 groups with 3 particles"]
    assert (out / "groups.csv").read_bytes() == (
        b"group_id,vx,vy,vz,mass\n0,1.0,3.0,-2.0,0.25\n1,0.0,0.0,-0.0,0.05000000074505806\n")
    assert (out / "particles.csv").read_bytes() == (
        "group_id,x,y,z\n0,1.0,3.0,-2.0\n0,%r,-0.5,-0.0\n1,0.125,0.0,-0.25\n" % x01).encode()
    x, v, m = snowmpm.import_particles(str(out / "groups.csv"), str(out / "particles.csv"))
    assert x.tolist() == [[1.0, 3.0, -2.0], [x01, -0.5, 0.0], [0.125, 0.0, -0.25]]
    assert v.tolist() == [[1.0, 3.0, -2.0]] * 2 + [[0.0, 0.0, 0.0]]
    assert m.tolist() == [0.25, 0.25, 0.05000000074505806]


def test_solver_frame_keeps_coordinates(tmp_path, monkeypatch):
    monkeypatch.setattr(seed, "seed_mesh", lambda v, t, s, device=0: np.array([[1.0, 2.0, 3.0]]))
    g = seed.Group(np.zeros((3, 3)), [[0, 1, 2]], "a", velocity=(4, 5, 6))
    seed.seed_groups([g], tmp_path, frame="solver", log=lambda s: None)
    assert (tmp_path / "particles.csv").read_text() == "group_id,x,y,z\n0,1.0,2.0,3.0\n"
    assert (tmp_path / "groups.csv").read_text().splitlines()[1] == "0,4.0,5.0,6.0,0.05000000074505806"


def test_no_groups_and_cli_errors(tmp_path, capsys):
    with pytest.raises(ValueError, match="No enabled snow mesh objects selected"):
        seed.seed_groups([], tmp_path)
    assert seed.main([str(tmp_path), "--group", str(tmp_path / "missing.obj")]) == 1
    p = _obj(tmp_path / "t.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    assert seed.main([str(tmp_path), "--group", p, "colour=red"]) == 1
    assert seed.main([str(tmp_path), "--group", p, "velocity=1,2"]) == 1
    assert "error:" in capsys.readouterr().err
