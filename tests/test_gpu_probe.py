"""The probes behind the kernel bounds run and report consistent figures:
every issue-rate loop takes time and the interleaved loop is not faster than
the slower of its two parts; the streaming probe moves the bytes it names
at a rate below the HBM peak."""
import pytest

import snowmpm

pytestmark = pytest.mark.gpu


def test_issue_probe_runs():
    r = snowmpm.probe_issue(2, 4, iterations=256)
    assert r["mfma"] > 0 and r["valu"] > 0 and r["mixed"] > 0
    assert r["mixed"] >= 0.9 * max(r["mfma"], r["valu"])
    assert all(u > 0 for u in r["us"])


def test_stream_probe_runs():
    for mode, nbytes in ((0, 2 << 30), (1, 41 * 4 * 8_000_000), (2, 41 * 4 * 8_000_000)):
        r = snowmpm.probe_stream(mode)
        assert r["bytes"] == nbytes
        assert r["us"] > 0 and 0.5 < r["tb_per_s"] < 8.0
