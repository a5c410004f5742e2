"""GPU tests of the block-aligned particle layout: home blocks' sorted ranges
start on 32-particle (128-byte) boundaries, so a re-sort leaves holes in the
buffers.  The holes never enter the arithmetic: scenes stepped, removed and
added in an engine with aligned ranges give the same bits as in one with
ranges packed end to end (dense_blocks), and the capacity checks count
particles, not slots."""
import numpy as np
import pytest

import snowmpm

pytestmark = pytest.mark.gpu


def _thin(sc, every):
    """The scene with every `every`-th particle dropped, so that block counts
    are not multiples of 32."""
    keep = (np.arange(sc["x"].shape[0]) % every) != 3
    out = dict(sc)
    for k in ("x", "v", "m", "E"):
        if sc.get(k) is not None:
            out[k] = sc[k][keep]
    return out


def _add(eng, sc):
    return eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"], E=sc.get("E"))


def _run(dense, scenes, late):
    """Add scenes, step 6 (re-sort every 4 steps), remove the second, step 3,
    add `late`, step 7; the downloads of the scenes still resident."""
    cap = sum(s["x"].shape[0] for s in scenes) + late["x"].shape[0] + 64
    with snowmpm.Engine(capacity=cap, permute_period=4, dense_blocks=dense) as eng:
        sids = [_add(eng, s) for s in scenes]
        eng.setup()
        eng.step(6)
        eng.remove_scene(sids[1])
        eng.step(3)
        sl = _add(eng, late)
        eng.setup()
        eng.step(7)
        keep = [sids[0]] + sids[2:] + [sl]
        out = [eng.download(s) for s in keep]
        st = [eng.status(s) for s in keep]
        eng.step(4)  # includes a re-sort (the period is 4): no removed particles left
        used, total = eng.slots()
        live = sum(o["x"].shape[0] for o in out)
    return out, st, (used, total, live)


def test_aligned_ranges_equal_dense_ranges():
    scenes = [_thin(snowmpm.generate_scene(3, seed=11, index=i, scale=0.12), 97) for i in range(4)]
    scenes.append(_thin(snowmpm.generate_scene(2, seed=11, scale=0.06), 101))
    late = _thin(snowmpm.generate_scene(0, seed=11, scale=0.3), 89)
    a, sa, ua = _run(False, scenes, late)
    b, sb, ub = _run(True, scenes, late)
    assert sa == sb and all(s == (0, 0) for s in sa)
    # the aligned engine holds holes (its blocks' counts are not multiples of
    # 32), within the quarter of the capacity kept for them; the dense one
    # holds the particles end to end
    assert ua[2] < ua[0] <= ua[1] and ua[0] - ua[2] <= ua[1] - ub[1]
    assert ub[0] == ub[2] and ub[1] < ua[1]
    for da, db in zip(a, b):
        for k in ("x", "v", "Fe", "Jp", "V0"):
            assert np.array_equal(da[k], db[k]), k


def test_capacity_counts_particles_not_slots():
    """An engine whose capacity is exactly its scenes' particles: after a
    re-sort with holes the scenes still fit, a scene may be removed and a
    smaller one added, and one more particle than the capacity is refused."""
    sc = _thin(snowmpm.generate_scene(3, seed=12, index=3, scale=0.2), 53)
    sc2 = _thin(snowmpm.generate_scene(3, seed=12, index=4, scale=0.1), 53)
    n = sc["x"].shape[0]
    with snowmpm.Engine(capacity=n + sc2["x"].shape[0]) as eng:
        s1 = _add(eng, sc)
        eng.step(2)
        s2 = _add(eng, sc2)
        eng.step(2)
        part = {k: (v[: v.shape[0] * 9 // 10] if isinstance(v, np.ndarray) else v) for k, v in sc2.items()}
        with pytest.raises(snowmpm.MpmError) as ei:
            _add(eng, part)
        assert ei.value.code == snowmpm.ERR_CAPACITY
        eng.remove_scene(s2)
        eng.step(1)
        s3 = _add(eng, part)
        eng.step(2)
        assert eng.status(s1) == (0, 0) and eng.status(s3) == (0, 0)
