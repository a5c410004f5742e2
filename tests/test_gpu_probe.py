"""The issue-rate probe behind the P2G bound runs and reports consistent
figures: every loop takes time, and the interleaved loop is not faster than
the slower of its two parts."""
import pytest

import snowmpm

pytestmark = pytest.mark.gpu


def test_issue_probe_runs():
    r = snowmpm.probe_issue(2, 4, iterations=256)
    assert r["mfma"] > 0 and r["valu"] > 0 and r["mixed"] > 0
    assert r["mixed"] >= 0.9 * max(r["mfma"], r["valu"])
    assert all(u > 0 for u in r["us"])
