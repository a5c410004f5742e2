"""The oracle reproduces its committed 10-step trajectory
(tests/golden/oracle_trajectory.npz, written by make_oracle_trajectory.py):
a regression pin on the oracle itself, independent of the GPU."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_oracle_trajectory as mot  # noqa: E402

FIX = os.path.join(HERE, "golden", "oracle_trajectory.npz")


def test_oracle_reproduces_committed_trajectory():
    ref = np.load(FIX)
    x, v, m, prm, planes = mot.scene()
    np.testing.assert_array_equal(x, ref["x0"])
    got = mot.run(x, v, m, prm, planes, (1, 10))
    np.testing.assert_allclose(got["V0"], ref["V0"], rtol=1e-13, atol=0)
    for k in (1, 10):
        for f, tol in (("x", 1e-13), ("v", 1e-11), ("Fe", 1e-12), ("Fp", 1e-12)):
            np.testing.assert_allclose(got[f"{f}_{k}"], ref[f"{f}_{k}"], rtol=0, atol=tol, err_msg=f"{f}_{k}")


def test_trajectory_exercises_contact_and_plasticity():
    ref = np.load(FIX)
    Jp = np.linalg.det(ref["Fp_10"].reshape(-1, 3, 3))
    assert (np.abs(Jp - 1) > 1e-6).sum() > 100          # yielded particles
    assert ref["v_10"][:, 1].max() > -3.0 + 1e-3          # the ground stopped some of the fall
    s = 2 ** -0.5
    assert ((-(ref["x_10"][:, 0] - 0.09) + ref["x_10"][:, 1]) * s <= 0).sum() > 0   # tilted plane reached
