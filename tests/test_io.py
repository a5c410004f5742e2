"""The reference's on-disk formats through the product library (host code only).

Reader semantics restated from src/import.cpp:11-137: the header line is
skipped unchecked, empty lines are skipped, a trailing '\\r' is stripped from
the last field, extra fields are ignored, and a missing or empty field, a
duplicate group id, a particle naming an undefined group, a missing or empty
file all fail with the reference's message text.  Writer format from
src/export.cpp:10-22: frame_%06d.csv, header "x,y,z", rows "x,-z,y" with the
default stream formatting (6 significant digits).  The inputs are small files
written by the tests themselves.
"""
import os

import numpy as np
import pytest

import snowmpm


def _write(path, text):
    with open(path, "w", newline="") as f:
        f.write(text)
    return str(path)


GROUPS = "group_id,vx,vy,vz,mass\n1,0.5,-2,0,0.05\n2,0,0,1.25,0.1\n"
SEEDS = "group_id,x,y,z\n1,0.1,0.2,0.3\n2,-1.5,0,2e-3\n1,0,0,0\n"


def test_roundtrip_values(tmp_path):
    g = _write(tmp_path / "g.csv", GROUPS)
    p = _write(tmp_path / "p.csv", SEEDS)
    x, v, m = snowmpm.import_particles(g, p)
    assert x.shape == (3, 3)
    np.testing.assert_array_equal(x, [[0.1, 0.2, 0.3], [-1.5, 0, 2e-3], [0, 0, 0]])
    np.testing.assert_array_equal(v, [[0.5, -2, 0], [0, 0, 1.25], [0.5, -2, 0]])
    np.testing.assert_array_equal(m.ravel(), [0.05, 0.1, 0.05])


def test_header_unchecked_blank_lines_crlf_extra_fields(tmp_path):
    g = _write(tmp_path / "g.csv", "anything at all\r\n1,1,2,3,4\r\n\n2,5,6,7,8,ignored\r\n")
    p = _write(tmp_path / "p.csv", "h\n\n2,1,2,3\r\n\n1,4,5,6,extra\n")
    x, v, m = snowmpm.import_particles(g, p)
    np.testing.assert_array_equal(x, [[1, 2, 3], [4, 5, 6]])
    np.testing.assert_array_equal(v, [[5, 6, 7], [1, 2, 3]])
    np.testing.assert_array_equal(m.ravel(), [8, 4])


def test_numeric_prefix_parsing_like_stod(tmp_path):
    # std::stoi / std::stod read the longest numeric prefix (import.cpp:58-62)
    g = _write(tmp_path / "g.csv", "h\n 3abc,1e1x,  -2,+0.5,7.\n")
    p = _write(tmp_path / "p.csv", "h\n3,1,2,3\n")
    x, v, m = snowmpm.import_particles(g, p)
    np.testing.assert_array_equal(v, [[10, -2, 0.5]])
    assert m.ravel()[0] == 7.0


def test_header_only_gives_no_particles(tmp_path):
    g = _write(tmp_path / "g.csv", GROUPS)
    p = _write(tmp_path / "p.csv", "group_id,x,y,z\n")
    x, v, m = snowmpm.import_particles(g, p)
    assert x.shape == (0, 3) and m.size == 0


@pytest.mark.parametrize("groups,seeds,message", [
    (None, SEEDS, "File not open: "),
    (GROUPS, None, "File not open: "),
    ("", SEEDS, "Group data file is empty: "),
    (GROUPS, "", "Particle data file is empty: "),
    ("h\n1,0,0,0\n", SEEDS, "Missing CSV field: m"),
    ("h\n1,0,,0,1\n", SEEDS, "Empty CSV field: v_y"),
    ("h\n1,0,0,0,\r\n", SEEDS, "Empty CSV field: m"),
    ("h\n\r\n", SEEDS, "Empty CSV field: group_ID"),
    ("h\n3,0,0,0,1\n3,1,1,1,1\n", SEEDS, "Duplicate group id: 3"),
    (GROUPS, "h\n1,0,0\n", "Missing CSV field: z"),
    (GROUPS, "h\n7,0,0,0\n", "Particle references undefined group id: 7"),
    ("h\nx,0,0,0,1\n", SEEDS, "stoi"),
    (GROUPS, "h\n1,0,y,0\n", "stod"),
])
def test_reader_errors(tmp_path, groups, seeds, message):
    g = str(tmp_path / "missing_g.csv") if groups is None else _write(tmp_path / "g.csv", groups)
    p = str(tmp_path / "missing_p.csv") if seeds is None else _write(tmp_path / "p.csv", seeds)
    with pytest.raises(snowmpm.MpmError) as e:
        snowmpm.import_particles(g, p)
    assert e.value.code == snowmpm.ERR_IO
    assert message in str(e.value)


def test_frame_writer_format(tmp_path):
    out = tmp_path / "output"
    x = np.array([[0.1234567, 2.0, -3.5e-7], [1e-5, -0.0, 0.0], [123456789.0, 1.0 / 3.0, 2.5]])
    snowmpm.export_frame(str(out), 7, x)
    text = open(out / "frame_000007.csv").read()
    # x, -z, y with 6 significant digits; -(+0.0) prints as -0
    assert text == "x,y,z\n0.123457,3.5e-07,2\n1e-05,-0,-0\n1.23457e+08,-2.5,0.333333\n"


def test_frame_writer_numbering_and_empty(tmp_path):
    out = tmp_path / "o"
    snowmpm.export_frame(str(out), 123456, np.zeros((0, 3)))
    assert open(out / "frame_123456.csv").read() == "x,y,z\n"
    snowmpm.export_frame(str(out), 1234567, np.ones((1, 3)))
    assert open(out / "frame_1234567.csv").read() == "x,y,z\n1,-1,1\n"


def test_frame_writer_fails_loudly(tmp_path):
    blocker = _write(tmp_path / "file", "not a directory")
    with pytest.raises(snowmpm.MpmError) as e:
        snowmpm.export_frame(os.path.join(blocker, "sub"), 0, np.ones((1, 3)))
    assert "cannot" in str(e.value)
