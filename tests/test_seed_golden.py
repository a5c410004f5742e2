"""The seeding tool against the reference add-on's own output.

The reference ships input/groups.csv and input/particles.csv, written by its
Blender add-on (blender/seed_export.py) from a mesh that is not in the
repository.  No row of them is kept here: tests/golden/seed_addon_facts.json
(made by tests/golden/make_seed_fixture.py) holds facts derived from them:
their SHA-256 and sizes, the row count, the group's settings, and per
Blender axis the float32 lattice origin whose double-accumulated loop
x = lo + s/2 + k s (seed_export.py:109-127) reproduces every coordinate, and
the lattice indices used.  Without the mesh the inside decisions cannot be
replayed, but what surrounds them is pinned:
  * the lattice: the recorded origins give lattices that a float32
    accumulation does not reproduce (so the add-on's double loop matters);
    where the reference is checked out, every coordinate lies on exactly one
    such lattice per axis, and that origin is the recorded one;
  * the order: x outer, y, z inner in Blender's frame (seed_export.py:111-127);
  * the files themselves: the tool's writer, given the add-on's points and the
    group's settings (velocity (0, 0, -2) in Blender's frame, mass 0.05),
    writes files whose SHA-256 equal the reference's (float32 values in repr,
    (x, y, z) -> (x, z, -y), "\\n" line ends).
The last two need the points, so they run where the reference is checked out
and are skipped elsewhere.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as orc
import snowmpm.seed as seed

HERE = os.path.dirname(os.path.abspath(__file__))
FACTS = json.load(open(os.path.join(HERE, "golden", "seed_addon_facts.json")))
REF = "/root/reference/input"
have_ref = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "particles.csv")),
                              reason="needs the reference's input files (only the derived facts are kept)")


def _lattice(e):
    ax = FACTS["axes"][e]
    s32 = float(np.float32(FACTS["spacing"]))
    seq = orc.seed_axis(ax["origin_f32"], ax["origin_f32"] + (ax["k_max"] + 2) * s32, FACTS["spacing"])
    return seq, ax


def test_recorded_lattices_need_double_accumulation():
    s32 = np.float32(FACTS["spacing"])
    for e in range(3):
        seq, ax = _lattice(e)
        used = seq[ax["k_min"]:ax["k_max"] + 1]
        assert len(used) == ax["k_max"] - ax["k_min"] + 1 >= ax["distinct"]
        assert np.all(used == used.astype(np.float32))  # stored as float32
        x = np.float32(ax["origin_f32"]) + s32 * np.float32(0.5)
        f32 = []
        for _ in range(ax["k_max"] + 1):
            f32.append(float(x))
            x = np.float32(x + s32)
        assert f32[ax["k_min"]:] != used.tolist(), e  # a float32 loop drifts off the add-on's values


def _points():
    import make_seed_fixture as mk  # noqa: E402 (tests/golden)
    return mk.points(os.path.join(REF, "particles.csv"))


@pytest.fixture
def golden_path(monkeypatch):
    monkeypatch.syspath_prepend(os.path.join(HERE, "golden"))


@have_ref
def test_facts_match_the_reference_files(golden_path):
    import make_seed_fixture as mk
    assert mk.facts(REF) == FACTS


@have_ref
def test_order_is_x_outer_z_inner(golden_path):
    pts = _points()
    assert len(pts) == FACTS["rows"]
    keys = [tuple(p) for p in pts.tolist()]
    assert keys == sorted(keys)
    for e in range(3):
        seq, ax = _lattice(e)
        assert set(np.unique(pts[:, e]).tolist()) <= set(seq.tolist())


@have_ref
def test_writer_reproduces_the_addons_files(tmp_path, monkeypatch, golden_path):
    pts = _points()
    monkeypatch.setattr(seed, "seed_mesh", lambda v, t, s, device=0: pts)
    g = seed.Group(np.zeros((3, 3)), [[0, 1, 2]], "Snow", velocity=(0.0, 0.0, -2.0), mass=0.05,
                   spacing=FACTS["spacing"])
    seed.seed_groups([g], tmp_path, log=lambda s: None)
    for name in ("groups.csv", "particles.csv"):
        data = (tmp_path / name).read_bytes()
        assert len(data) == FACTS[name]["bytes"], name
        assert hashlib.sha256(data).hexdigest() == FACTS[name]["sha256"], name
