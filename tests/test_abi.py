"""CPU-side checks of the product library: it loads, exports every symbol
include/snowmpm.h declares, and its host-only entry points (workload
generators, CSV I/O, collider factories) behave.  No GPU compute here."""
import ctypes
import os
import re

import numpy as np
import pytest

import snowmpm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "snowmpm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 15
    lib = ctypes.CDLL(snowmpm._LIB)
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(snowmpm.EXPORTED)


def test_make_plane_normalises():
    c = snowmpm.make_plane((0, 1, 0), (0, 2, 0), 0.3)
    assert c.type == snowmpm.PLANE
    assert list(c.normal) == [0.0, 1.0, 0.0]
    assert c.friction == 0.3 and list(c.velocity) == [0, 0, 0]


@pytest.mark.parametrize("kind,n", [(0, 64000), (2, None)])
def test_generate_sizes(kind, n):
    sc = snowmpm.generate_scene(kind, seed=1, scale=1.0 if kind == 0 else 0.05)
    if n:
        assert sc["x"].shape == (n, 3)
    assert np.allclose(sc["m"], 400 * 0.01 ** 3)
    p = sc["params"]
    assert p.dx == (0.02)
    mu, la = snowmpm.lame(1.4e5, 0.2)
    assert abs(p.mu0 - mu) < 1e-9 and abs(p.lambda0 - la) < 1e-9


def test_full_slab_is_eight_million():
    sc = snowmpm.generate_scene(2, seed=5, scale=1.0)
    assert sc["x"].shape == (8_000_000, 3)
    assert sc["E"].min() >= 1e5 and sc["E"].max() <= 2e5
    # 8 particles per cell: lattice spacing dx/2, jitter +-0.1 spacing
    assert sc["x"][:, 1].min() > 0.2 - 0.0011 and sc["x"][:, 1].max() < 0.7 + 0.0011


def test_tori_count_and_velocities():
    sc = snowmpm.generate_scene(1, seed=5, scale=1.0)
    n = sc["x"].shape[0]
    assert 1_300_000 < n < 1_400_000  # two tori R=0.3 r=0.12 at spacing 0.005
    assert set(np.unique(sc["v"][:, 0])) == {-5.0, 5.0}


def test_generator_is_seeded():
    a = snowmpm.generate_scene(1, seed=9, scale=0.1)
    b = snowmpm.generate_scene(1, seed=9, scale=0.1)
    c = snowmpm.generate_scene(1, seed=10, scale=0.1)
    assert np.array_equal(a["x"], b["x"])
    assert not np.array_equal(a["x"], c["x"])


def test_batched_distribution():
    ns = [snowmpm.generate_scene(3, seed=3, index=i)["x"].shape[0] for i in range(64)]
    assert min(ns) >= 300 and max(ns) <= 260_000
    assert np.median(ns) > 3000


def test_probe_rejects_bad_arguments_without_a_gpu():
    lib = ctypes.CDLL(snowmpm._LIB)
    out = (ctypes.c_double * 6)()
    assert lib.mpm_probe_issue(0, 0, 4, 16, out) == snowmpm.ERR_ARG      # no waves
    assert lib.mpm_probe_issue(0, 1, 3, 16, out) == snowmpm.ERR_ARG      # K not in 1, 2, 4, 8, 16
    assert lib.mpm_probe_issue(0, 1, 4, 0, out) == snowmpm.ERR_ARG       # no iterations
    assert lib.mpm_probe_issue(0, 1, 4, 16, None) == snowmpm.ERR_ARG
    used, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert lib.mpm_slots(None, ctypes.byref(used), ctypes.byref(total)) == snowmpm.ERR_ARG  # no engine
    out3 = (ctypes.c_double * 3)()
    assert lib.mpm_probe_stream(0, -1, out3) == snowmpm.ERR_ARG          # no such mode
    assert lib.mpm_probe_stream(0, 3, out3) == snowmpm.ERR_ARG
    assert lib.mpm_probe_stream(0, 0, None) == snowmpm.ERR_ARG


def test_multi_step_rejects_bad_arguments_without_a_gpu():
    lib = ctypes.CDLL(snowmpm._LIB)
    lib.mpm_step_graph_multi.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int]
    none = (ctypes.c_void_p * 2)(None, None)
    assert lib.mpm_step_graph_multi(None, 1, 1) == snowmpm.ERR_ARG         # no list
    assert lib.mpm_step_graph_multi(none, 0, 1) == snowmpm.ERR_ARG         # no engines
    assert lib.mpm_step_graph_multi(none, 65, 1) == snowmpm.ERR_ARG        # too many
    assert lib.mpm_step_graph_multi(none, 2, 1) == snowmpm.ERR_ARG         # null engines
    assert "engine" in snowmpm.last_error()
