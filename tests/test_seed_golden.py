"""The seeding tool against the reference add-on's own output.

tests/golden/seed_addon/ holds the input the reference ships
(input/groups.csv, input/particles.csv), written by its Blender add-on
(blender/seed_export.py) from a mesh that is not in the repository.  Without
the mesh the inside decisions cannot be replayed, but everything around them
is pinned by these files:
  * the lattice: on each axis the particles' coordinates are values of the
    add-on's loop x = lo + s/2 + k s (seed_export.py:109-127) accumulated in
    double and stored as float32, for s = float32(0.03) and exactly one
    float32 lo; a float32 accumulation reproduces none of them;
  * the order: x outer, y, z inner in Blender's frame (seed_export.py:111-127);
  * the files themselves: the tool's writer, given the add-on's points and
    the group's settings (velocity (0, 0, -2) in Blender's frame, mass
    0.05), writes both files byte for byte as the add-on did (float32
    values in repr, (x, y, z) -> (x, z, -y), "\\n" line ends).
"""
import os

import numpy as np

import oracle as orc
import snowmpm.seed as seed

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "seed_addon")
S = 0.03


def _points():
    rows = [l.split(",") for l in open(os.path.join(GOLD, "particles.csv")).read().splitlines()[1:] if l]
    solver = np.array([[float(v) for v in r[1:]] for r in rows])
    return np.stack([solver[:, 0], -solver[:, 2], solver[:, 1]], axis=1)  # back to Blender's frame


def _f32_near(v, k):
    out, a, b = [np.float32(v)], np.float32(v), np.float32(v)
    for _ in range(k):
        a, b = np.nextafter(a, np.float32(-np.inf)), np.nextafter(b, np.float32(np.inf))
        out += [a, b]
    return out


def test_lattice_is_the_addons_double_accumulation():
    pts = _points()
    s32 = float(np.float32(S))
    for e in range(3):
        vals = set(np.unique(pts[:, e]).tolist())
        found, found32 = [], []
        for k0 in range(6):  # the first lattice values may all lie outside the mesh
            for lo in _f32_near(min(vals) - s32 * 0.5 - k0 * s32, 200):
                seq = orc.seed_axis(float(lo), float(lo) + (len(vals) + k0 + 1) * s32, S).tolist()
                if vals <= set(seq):
                    found.append(float(lo))
                x, seq32 = np.float32(lo) + np.float32(s32) * np.float32(0.5), []
                for _ in range(len(vals) + k0 + 1):
                    seq32.append(float(x))
                    x = np.float32(x + np.float32(s32))
                if vals <= set(seq32):
                    found32.append(float(lo))
        assert len(set(found)) == 1, (e, found)
        assert not found32, e


def test_order_is_x_outer_z_inner():
    pts = _points()
    keys = [tuple(p) for p in pts.tolist()]
    assert keys == sorted(keys)


def test_writer_reproduces_the_addons_files(tmp_path, monkeypatch):
    pts = _points()
    monkeypatch.setattr(seed, "seed_mesh", lambda v, t, s, device=0: pts)
    g = seed.Group(np.zeros((3, 3)), [[0, 1, 2]], "Snow", velocity=(0.0, 0.0, -2.0), mass=0.05, spacing=S)
    seed.seed_groups([g], tmp_path, log=lambda s: None)
    for name in ("groups.csv", "particles.csv"):
        assert (tmp_path / name).read_bytes() == open(os.path.join(GOLD, name), "rb").read(), name
