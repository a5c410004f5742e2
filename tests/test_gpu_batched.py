"""GPU tests of the batched and served forms: graph-replayed steps, the
autotuner and the continuous-batching request loop give the same bits as the
plain per-scene step (launch shape and batch composition never enter the
arithmetic)."""
import numpy as np
import pytest

import mpmref
import snowmpm

pytestmark = pytest.mark.gpu


def _scenes(k, scale=0.08, seed=5):
    return [snowmpm.generate_scene(3, seed=seed, index=i, scale=scale) for i in range(k)]


def test_graph_steps_equal_plain_steps():
    sc = snowmpm.generate_scene(1, seed=3, scale=0.12)
    a = mpmref.run_gpu(sc, 9)
    with snowmpm.Engine(capacity=len(sc["x"]) + 64) as eng:
        sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
        eng.step_graph(4)
        eng.step(1)
        eng.step_graph(4)
        b = eng.download(sid)
    for k in ("x", "v", "Fe", "Jp"):
        assert np.array_equal(a[k], b[k]), k


def test_autotune_keeps_results():
    """The autotuner times ordinary steps (launch-shape candidates, then 4
    re-sort periods x 16 steps); the scene's state after them equals, bit for
    bit, the same number of plain steps."""
    sc = snowmpm.generate_scene(0, seed=3, scale=0.3)
    with snowmpm.Engine(capacity=len(sc["x"]) + 64) as eng:
        sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
        eng.setup()
        chosen = eng.autotune(1)
        assert all(v > 0 for v in chosen.values())
        nsteps = eng.steps(sid)
        got = eng.download(sid)
    assert chosen["permute_period"] in (2, 4, 8, 16)
    assert nsteps > 64
    ref = mpmref.run_gpu(sc, nsteps)
    for k in ("x", "v", "Fe", "Jp"):
        assert np.array_equal(ref[k], got[k]), k


def test_served_results_equal_alone():
    scs = _scenes(6)
    reqs = []
    for i, sc in enumerate(scs):
        reqs.append(dict(sc, arrival_s=0.0 if i < 3 else 0.001 * i, steps=7 + 3 * i))
    cap = sum(len(s["x"]) for s in scs) + 1024
    with snowmpm.Engine(capacity=cap) as eng:
        res = eng.serve(reqs, mode=0, max_chunk=4, keep_outputs=True)
    assert res["completed"] == len(reqs)
    assert res["max_resident"] >= 2
    for q, (ox, ov) in zip(reqs, res["outputs"]):
        alone = mpmref.run_gpu(q, q["steps"])
        assert np.array_equal(alone["x"], ox)
        assert np.array_equal(alone["v"], ov)
    assert res["p99_s"] >= res["p50_s"] > 0


def test_serve_rejects_a_negative_step_count():
    sc = _scenes(1)[0]
    with snowmpm.Engine(capacity=len(sc["x"]) + 64) as eng:
        with pytest.raises(snowmpm.MpmError) as e:
            eng.serve([dict(sc, arrival_s=0.0, steps=-1)])
        assert e.value.code == snowmpm.ERR_ARG
        # the engine stays usable
        res = eng.serve([dict(sc, arrival_s=0.0, steps=2)])
    assert res["completed"] == 1


def test_plain_serving_mode():
    scs = _scenes(3)
    reqs = [dict(sc, arrival_s=0.0, steps=5) for sc in scs]
    with snowmpm.Engine(capacity=max(len(s["x"]) for s in scs) + 64) as eng:
        res = eng.serve(reqs, mode=1, keep_outputs=True)
    assert res["completed"] == 3 and res["max_resident"] == 1
    for q, (ox, ov) in zip(reqs, res["outputs"]):
        assert np.array_equal(mpmref.run_gpu(q, 5)["x"], ox)


def test_admission_waits_for_capacity():
    scs = _scenes(4)
    n_max = max(len(s["x"]) for s in scs)
    reqs = [dict(sc, arrival_s=0.0, steps=6) for sc in scs]
    cap = n_max + 8
    assert sum(len(s["x"]) for s in scs) > cap  # not all four fit at once
    with snowmpm.Engine(capacity=cap) as eng:
        res = eng.serve(reqs, mode=0, max_chunk=3, keep_outputs=True)
    assert res["completed"] == 4 and 1 <= res["max_resident"] < 4
    for q, (ox, _) in zip(reqs, res["outputs"]):
        assert np.array_equal(mpmref.run_gpu(q, 6)["x"], ox)


def test_remove_and_add_between_steps_keeps_results():
    """Scenes removed mid-run are compacted away at the next (re-sorting) step
    and new scenes are appended; the survivors' bits equal their solo runs."""
    s0, s1, s2 = _scenes(3, seed=9)
    with snowmpm.Engine(capacity=sum(len(s["x"]) for s in (s0, s1, s2)) + 1024) as eng:
        i0 = eng.add_scene(s0["params"], s0["colliders"], s0["x"], s0["v"], s0["m"])
        i1 = eng.add_scene(s1["params"], s1["colliders"], s1["x"], s1["v"], s1["m"])
        eng.step(5)
        eng.remove_scene(i1)
        eng.step(3)
        i2 = eng.add_scene(s2["params"], s2["colliders"], s2["x"], s2["v"], s2["m"])
        eng.step(20)
        eng.step_graph(6)
        g0, g2 = eng.download(i0), eng.download(i2)
    a0 = mpmref.run_gpu(s0, 34)
    a2 = mpmref.run_gpu(s2, 26)
    for k in ("x", "v", "Fe", "Jp"):
        assert np.array_equal(a0[k], g0[k]), k
        assert np.array_equal(a2[k], g2[k]), k


def test_upload_download_round_trip():
    """Scene upload and download convert on the GPU between the reference's
    host layout (doubles, per-particle x/v/F rows) and the FP32 buffers: the
    state comes back as the float32 rounding of what went in (J_P = det F_P
    formed in double), in the input order, and a download of some fields
    returns exactly those."""
    rng = np.random.default_rng(4)
    n = 1500
    x = rng.uniform(-1, 1, (n, 3))
    v = rng.uniform(-2, 2, (n, 3))
    m = rng.uniform(1e-4, 1e-3, n)
    V0 = rng.uniform(1e-7, 1e-6, n)
    E = rng.uniform(1e5, 2e5, n)
    Fe = (np.eye(3) + rng.normal(0, 0.05, (n, 3, 3))).reshape(n, 9)
    Fp = (np.eye(3) + rng.normal(0, 0.05, (n, 3, 3))).reshape(n, 9)
    prm = snowmpm.params(1e-4, 0.02, *snowmpm.lame(1.4e5, 0.2), 10, 0.025, 0.0075, 0.95)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    with snowmpm.Engine(capacity=n + 64, track_fp=True) as eng:
        eng.add_scene(prm, [], x[:7], v[:7], m[:7])  # another scene first: the second lands at an offset
        sid = eng.add_scene(prm, [snowmpm.make_plane((0, 0, 0), (0, 1, 0), 0.2)], x, v, m, V0=V0, Fe=Fe, Fp=Fp,
                            E=E)
        full = eng.download(sid)
        part = eng.download(sid, fields=("x", "Jp"))
    assert np.array_equal(full["x"], f32(x)) and np.array_equal(full["v"], f32(v))
    assert np.array_equal(full["Fe"].reshape(n, 9), f32(Fe)) and np.array_equal(full["Fp"].reshape(n, 9), f32(Fp))
    assert np.array_equal(full["V0"], f32(V0))
    P = Fp.reshape(n, 3, 3)
    det = (P[:, 0, 0] * (P[:, 1, 1] * P[:, 2, 2] - P[:, 1, 2] * P[:, 2, 1])
           - P[:, 0, 1] * (P[:, 1, 0] * P[:, 2, 2] - P[:, 1, 2] * P[:, 2, 0])
           + P[:, 0, 2] * (P[:, 1, 0] * P[:, 2, 1] - P[:, 1, 1] * P[:, 2, 0]))
    assert np.array_equal(full["Jp"], f32(det))
    assert set(part) == {"x", "Jp"}
    assert np.array_equal(part["x"], full["x"]) and np.array_equal(part["Jp"], full["Jp"])


def test_tile_sizes_do_not_change_bits():
    """P2G staging chunks of 16, 32, 48 and 64 particles and G2P workgroups of 64,
    128 and 256 threads give bitwise the same state: on uniformly random
    positions (cells of 0 to ~35 particles, cells longer than a chunk) and on
    a scene of the batched distribution, through contact."""
    rng = np.random.default_rng(12)
    n = 3000
    x = rng.uniform(0.0, 0.1, size=(n, 3)) + np.array([0.31, 0.047, -0.23])
    v = rng.normal(0.0, 0.5, size=(n, 3)) + np.array([0.0, -2.0, 0.0])
    prm = snowmpm.params(1e-4, 0.02, *snowmpm.lame(1.4e5, 0.2), 10.0, 0.025, 0.0075, 0.95)
    rand = dict(x=x, v=v, m=np.full(n, 4e-4), params=prm, colliders=[snowmpm.make_plane((0, 0, 0), (0, 1, 0), 0.2)])
    for sc in (rand, snowmpm.generate_scene(3, seed=5, index=7, scale=0.5)):
        out = []
        for chunk, thr in ((48, 128), (16, 128), (32, 128), (64, 64), (48, 256)):
            with snowmpm.Engine(capacity=len(sc["x"]) + 64, p2g_chunk=chunk, g2p_threads=thr) as eng:
                sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
                assert eng.tile_sizes() == dict(p2g_chunk=chunk, g2p_threads=thr)
                eng.step(25)
                out.append(eng.download(sid))
                assert eng.status(sid)[0] == 0
        for o in out[1:]:
            for k in ("x", "v", "Fe", "Jp"):
                assert np.array_equal(out[0][k], o[k]), k


def test_multi_engine_steps_equal_alone():
    """Scenes split over two engines and stepped together
    (mpm_step_graph_multi, each engine on its own stream), beside a third
    engine holding a semi-implicit scene (host-driven steps in the same
    call): every scene equals its solo run bit for bit."""
    scs = _scenes(4, seed=9)
    imp = snowmpm.generate_scene(0, seed=4, scale=0.2)
    imp["params"].beta = 1.0
    imp["params"].cr_tolerance = 1e-6
    groups = [scs[0::2], scs[1::2], [imp]]
    engines, sids = [], []
    try:
        for g in groups:
            e = snowmpm.Engine(capacity=sum(len(s["x"]) for s in g) + 1024)
            engines.append(e)
            sids.append([e.add_scene(s["params"], s["colliders"], s["x"], s["v"], s["m"]) for s in g])
        snowmpm.step_graph_multi(engines, 6)
        snowmpm.step_graph_multi(engines[:2], 3)
        got = [[e.download(sid) for sid in ss] for e, ss in zip(engines, sids)]
        assert engines[2].cr_stats(sids[2][0])["iterations"] > 0
    finally:
        for e in engines:
            e.close()
    for g, gg, nsteps in zip(groups, got, (9, 9, 6)):
        for s, b in zip(g, gg):
            a = mpmref.run_gpu(s, nsteps)
            for k in ("x", "v", "Fe", "Jp"):
                assert np.array_equal(a[k], b[k]), k


def test_multi_engine_reports_the_failing_engine():
    """The second engine's scene has a lone particle 1 m beside its body
    moving at 2e8 m/s, which leaves the block-coordinate range in its first
    step (as in the served-failure test): the call returns MPM_ERR_DEVICE
    naming engine 1, and the first engine's scene has still taken its
    steps."""
    good = _scenes(1, seed=2)[0]
    bad = dict(_scenes(1, seed=3)[0])
    bad["x"], bad["v"] = bad["x"].copy(), bad["v"].copy()
    bad["x"][3] = (bad["x"][:, 0].max() + 1.0, bad["x"][3, 1], bad["x"][3, 2])
    bad["v"][3] = (2.0e8, 0.0, 0.0)
    with snowmpm.Engine(capacity=len(good["x"]) + 64) as e0, snowmpm.Engine(capacity=len(bad["x"]) + 64) as e1:
        s0 = e0.add_scene(good["params"], good["colliders"], good["x"], good["v"], good["m"])
        e1.add_scene(bad["params"], bad["colliders"], bad["x"], bad["v"], bad["m"])
        with pytest.raises(snowmpm.MpmError) as err:
            snowmpm.step_graph_multi([e0, e1], 2)
        assert err.value.code == snowmpm.ERR_DEVICE
        assert "engine 1" in str(err.value)
        assert e0.steps(s0) == 2


def test_served_mix_with_a_semi_implicit_request():
    """A served trace in which one request has beta = 1: chunks that hold it
    take the host-driven steps of the solve, and every request, implicit or
    explicit, ends with the bits of its solo run."""
    scs = _scenes(3, seed=13)
    imp = snowmpm.generate_scene(0, seed=6, scale=0.2)
    imp["params"].beta, imp["params"].cr_tolerance = 1.0, 1e-6
    reqs = [dict(scs[0], arrival_s=0.0, steps=9), dict(imp, arrival_s=0.0, steps=6),
            dict(scs[1], arrival_s=0.002, steps=5), dict(scs[2], arrival_s=0.004, steps=8)]
    cap = sum(len(q["x"]) for q in reqs) + 1024
    with snowmpm.Engine(capacity=cap) as eng:
        res = eng.serve(reqs, mode=0, max_chunk=4, keep_outputs=True)
    assert res["completed"] == len(reqs) and res["failed"] == 0
    for q, (ox, ov) in zip(reqs, res["outputs"]):
        alone = mpmref.run_gpu(q, q["steps"])
        assert np.array_equal(alone["x"], ox)
        assert np.array_equal(alone["v"], ov)
