"""The batched and served workloads at their full size.

Batched: the 256 scenes of the batched distribution (9.5M particles, 1k-195k
per scene) resident in one engine, autotuned, then graph-replayed.  Sixteen
scenes (the smallest, the largest and 14 between them by size) must equal,
bit for bit, the same scene stepped alone with plain launches for the same
number of steps, and the five smallest must match the FP64 oracle.

Served: a Poisson trace (100 requests/s) of 24 full-size scenes of that
distribution, 100-300 steps each, through the continuous-batching loop;
every request must equal its solo run bit for bit.  Admission limited by the
block pool (bytes), not by particle slots, and a request that fails
mid-serve (retired alone, the others completed), are covered too.
"""
import numpy as np
import pytest

import mpmref
import snowmpm

pytestmark = pytest.mark.gpu

SEED = 20261203
TOL = 1e-4


def _solo(sc, nsteps):
    with snowmpm.Engine(capacity=len(sc["x"]) + 64) as eng:
        sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
        eng.step(nsteps)
        out = eng.download(sid, fields=("x", "v", "Fe", "Jp"))
        out["status"] = eng.status(sid)
    return out


def test_batched_256_scenes_equal_solo_and_oracle():
    scenes = [snowmpm.generate_scene(3, seed=SEED, index=i) for i in range(256)]
    ns = np.array([len(s["x"]) for s in scenes])
    print(f"batched: 256 scenes, {ns.sum()} particles ({ns.min()}-{ns.max()})", flush=True)
    got, steps = {}, {}
    order = np.argsort(ns, kind="stable")
    sample = sorted(set(order[np.linspace(0, 255, 16).astype(int)].tolist()) | set(order[:5].tolist()))
    with snowmpm.Engine(capacity=int(ns.sum()) + 4096) as eng:
        sids = [eng.add_scene(s["params"], s["colliders"], s["x"], s["v"], s["m"]) for s in scenes]
        eng.setup()
        chosen = eng.autotune(1)
        eng.step_graph(10)
        for i, sid in enumerate(sids):
            st = eng.status(sid)
            assert st == (0, 0), (i, st)
        for i in sample:
            got[i] = eng.download(sids[i], fields=("x", "v", "Fe", "Jp"))
            steps[i] = eng.steps(sids[i])
    print(f"autotune {chosen}, {steps[sample[0]]} steps per scene", flush=True)
    assert len(set(steps.values())) == 1 and steps[sample[0]] > 10
    for i in sample:
        alone = _solo(scenes[i], steps[i])
        assert alone["status"] == (0, 0)
        for k in ("x", "v", "Fe", "Jp"):
            assert np.array_equal(alone[k], got[i][k]), (i, k)
    for i in order[:5]:
        ref = mpmref.run_oracle(scenes[i], steps[i])
        ex, ev = mpmref.rel_errors(got[i], ref, scenes[i], steps[i])
        print(f"scene {i} (n={ns[i]}): {steps[i]} steps, rel err x {ex:.2e} v {ev:.2e}", flush=True)
        assert ex < TOL and ev < TOL
        assert np.abs(got[i]["Fe"] - ref["Fe"]).max() < 1e-4


def _trace(k, rate, lo, hi, seed):
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=k)
    arr = np.cumsum(gaps) - gaps[0]
    idx = rng.choice(256, size=k, replace=False)
    return [dict(snowmpm.generate_scene(3, seed=SEED, index=int(i)), arrival_s=float(a),
                 steps=int(rng.integers(lo, hi + 1))) for i, a in zip(idx, arr)]


def test_served_poisson_trace_equals_solo():
    reqs = _trace(24, 100.0, 100, 300, seed=7)
    ntot = sum(len(q["x"]) for q in reqs)
    print(f"served: 24 requests, {ntot} particles", flush=True)
    with snowmpm.Engine(capacity=ntot // 2 + 200_000) as eng:
        res = eng.serve(reqs, mode=0, max_chunk=25, keep_outputs=True)
    print(f"served: {res['scenes_per_s']:.1f} scenes/s, p50 {res['p50_s']:.3f} s, p99 {res['p99_s']:.3f} s, "
          f"max resident {res['max_resident']}", flush=True)
    assert res["completed"] == 24 and res["failed"] == 0
    assert res["max_resident"] >= 4
    assert all(s == 0 for s in res["status"])
    for q, (ox, ov) in zip(reqs, res["outputs"]):
        alone = _solo(q, q["steps"])
        assert np.array_equal(alone["x"], ox)
        assert np.array_equal(alone["v"], ov)


def test_admission_limited_by_block_bytes():
    """Twelve scenes of the batched distribution (0.3 scale, 271-38,619
    particles) in an engine with particle slots for all of them but a block
    pool of 2^10 hash slots (512 blocks at the 50% load limit): the byte
    footprint of their blocks, not the particle count, bounds how many are
    resident, no hash or partial-sum overflow occurs, and every result equals
    its solo run."""
    reqs = [dict(snowmpm.generate_scene(3, seed=SEED, index=i, scale=0.3), arrival_s=0.0, steps=12)
            for i in range(12)]
    ns = [len(q["x"]) for q in reqs]
    cap = sum(ns) + 1024
    with snowmpm.Engine(capacity=cap, hash_log2=10) as eng:
        fps = [eng.footprint(q) for q in reqs]
        pools = fps[0][1]
        res = eng.serve(reqs, mode=0, max_chunk=4, keep_outputs=True)
    per_particle = pools[0] / cap
    grid_bytes = [f - n * per_particle for (f, _), n in zip(fps, ns)]
    assert max(grid_bytes) <= pools[1] + pools[2]        # each fits alone
    assert sum(grid_bytes) > 2 * (pools[1] + pools[2])   # all together do not
    assert res["completed"] == 12 and res["failed"] == 0
    assert 1 <= res["max_resident"] < 12
    assert res["max_resident_bytes"] <= sum(pools)
    for q, (ox, _) in zip(reqs, res["outputs"]):
        assert np.array_equal(_solo(q, 12)["x"], ox)


def test_failing_request_is_retired_alone():
    """One request carries a lone particle 1 m beside its body moving at
    2e8 m/s: it moves rigidly (its stencil nodes hold only its own mass) and
    leaves the block-coordinate range (|x| > ~5 km at dx 0.02) in its first
    step.  The request is retired with MPM_STATUS_RANGE at the end of that
    chunk; the other requests complete with their solo results, and the
    engine serves again afterwards."""
    reqs = [dict(snowmpm.generate_scene(3, seed=SEED, index=i, scale=0.2), arrival_s=0.0, steps=9)
            for i in range(4)]
    bad = dict(reqs[1])
    bad["x"], bad["v"] = bad["x"].copy(), bad["v"].copy()
    bad["x"][3] = (bad["x"][:, 0].max() + 1.0, bad["x"][3, 1], bad["x"][3, 2])
    bad["v"][3] = (2.0e8, 0.0, 0.0)
    reqs[1] = bad
    with snowmpm.Engine(capacity=sum(len(q["x"]) for q in reqs) + 1024) as eng:
        res = eng.serve(reqs, mode=0, max_chunk=3, keep_outputs=True)
        assert res["failed"] == 1 and res["completed"] == 3
        assert res["status"][1] & snowmpm.STATUS_RANGE
        assert not res["status"][1] & snowmpm.STATUS_NAN
        for k in (0, 2, 3):
            assert res["status"][k] == 0
            assert np.array_equal(_solo(reqs[k], 9)["x"], res["outputs"][k][0])
        again = eng.serve([reqs[0]], mode=0, keep_outputs=True)
        assert again["completed"] == 1
        assert np.array_equal(again["outputs"][0][0], res["outputs"][0][0])


def test_request_with_eight_colliders():
    """A request carries up to eight colliders (the engine's limit): the
    ground plane as the last of eight, served result equal to the solo run."""
    sc = snowmpm.generate_scene(3, seed=SEED, index=2, scale=0.3)
    far = [snowmpm.make_sphere((5.0 + k, 3.0, -4.0), 0.2, 0.1) for k in range(7)]
    q = dict(sc, colliders=far + list(sc["colliders"]), arrival_s=0.0, steps=8)
    assert len(q["colliders"]) == 8
    with snowmpm.Engine(capacity=len(sc["x"]) + 64) as eng:
        res = eng.serve([q], keep_outputs=True)
    assert res["completed"] == 1
    assert np.array_equal(_solo(q, 8)["x"], res["outputs"][0][0])
