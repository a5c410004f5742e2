"""mpm_scene_download_noalloc (no device allocation: staging through the
library's static 32 MiB device buffer) against mpm_scene_download, through
the C ABI: two calls in a row, the second for a smaller scene, bit for bit;
and a request larger than the buffer is refused with MPM_ERR_CAPACITY
while a smaller request for the same scene still works."""
import numpy as np
import pytest

import snowmpm

pytestmark = pytest.mark.gpu


def test_noalloc_download_equals_download():
    big = snowmpm.generate_scene(0, seed=5, scale=1.0)         # 64,000 particles
    small = snowmpm.generate_scene(3, seed=5, index=4, scale=0.3)
    print(f"no-alloc download: scenes of {len(big['x'])} and {len(small['x'])} particles", flush=True)
    assert len(small["x"]) < len(big["x"])
    with snowmpm.Engine(capacity=len(big["x"]) + len(small["x"]) + 64, track_fp=True) as eng:
        sb = eng.add_scene(big["params"], big["colliders"], big["x"], big["v"], big["m"])
        ss = eng.add_scene(small["params"], small["colliders"], small["x"], small["v"], small["m"])
        eng.step(5)
        for sid in (sb, ss):
            a = eng.download_noalloc(sid)
            b = eng.download(sid)
            assert set(a) == set(b) == {"x", "v", "Fe", "Fp", "Jp", "V0"}
            for k in a:
                assert np.array_equal(a[k], b[k]), (sid, k)
        # a partial request, then the full one again: nothing stale survives
        part = eng.download_noalloc(ss, fields=("v",))
        assert set(part) == {"v"} and np.array_equal(part["v"], eng.download(ss)["v"])
        again = eng.download_noalloc(sb)
        assert np.array_equal(again["x"], eng.download(sb)["x"])


def test_noalloc_refuses_more_than_its_buffer():
    sc = snowmpm.generate_scene(1, seed=5, scale=0.5)          # ~171k particles
    n = len(sc["x"])
    assert n * 26 * 8 > 32 << 20 > n * 3 * 8
    with snowmpm.Engine(capacity=n + 64, track_fp=True) as eng:
        sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"])
        eng.step(2)
        with pytest.raises(snowmpm.MpmError) as e:
            eng.download_noalloc(sid)
        assert e.value.code == snowmpm.ERR_CAPACITY
        x = eng.download_noalloc(sid, fields=("x",))["x"]
        assert np.array_equal(x, eng.download(sid, fields=("x",))["x"])
