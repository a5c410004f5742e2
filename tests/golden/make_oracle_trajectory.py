"""Writes oracle_trajectory.npz: a seeded 1,000-particle scene and the FP64
oracle's state after setup, 1 step and 10 steps.

The scene is a jittered 10x10x10 lattice (spacing 0.01, dx = 0.02, 8 ppc)
resting 2 mm above the ground plane y = 0 (friction 0.2), moving at
(0.5, -3, 0.2) m/s into the wedge it forms with a second, tilted plane
through (0.09, 0, 0) with normal (-1, 1, 0)/sqrt(2) (friction 0.5), so within
10 steps particles and nodes touch both colliders (applied in list order) and
yield plastically.  Material as the
64k cube workload: E = 1.4e5, nu = 0.2, theta_c = 0.025, theta_s = 0.0075,
xi = 10, alpha = 0.95, dt = 1e-4, g = (0, -9.81, 0).

The fixture pins the oracle against later changes to itself (tests check it
reproduces these states) and is the CI-size parity target for the GPU.
Run: python tests/golden/make_oracle_trajectory.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as orc  # noqa: E402


def scene():
    rng = np.random.default_rng(20261203)
    h = 0.01
    gi = np.stack(np.meshgrid(*[np.arange(10)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = (gi + 0.5) * h + rng.uniform(-0.1 * h, 0.1 * h, size=gi.shape)
    x += np.array([0.0, 0.002 - 0.5 * h + 0.1 * h, 0.0])
    v = np.tile([0.5, -3.0, 0.2], (len(x), 1))
    m = np.full(len(x), 400 * h ** 3)
    E, nu = 1.4e5, 0.2
    mu0, la0 = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    prm = dict(dt=1e-4, dx=0.02, mu0=mu0, lambda0=la0, hardening=10.0, max_compression=0.025,
               max_stretch=0.0075, alpha=0.95, gravity=(0.0, -9.81, 0.0))
    s = 2 ** -0.5
    planes = [((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.2), ((0.09, 0.0, 0.0), (-s, s, 0.0), 0.5)]
    return x, v, m, prm, planes


def run(x, v, m, prm, planes, steps):
    st = orc.State(x, v, m)
    st.setup(prm["dx"])
    V0 = st.field("V0").copy()
    cs = [orc.make_plane(*p) for p in planes]
    op = orc.params(prm["dt"], prm["dx"], prm["mu0"], prm["lambda0"], prm["hardening"], prm["max_compression"],
                    prm["max_stretch"], prm["alpha"], gravity=prm["gravity"])
    out = {"V0": V0}
    done = 0
    for k in steps:
        st.run(cs, op, k - done, nthreads=1)
        done = k
        for f in ("x", "v", "Fe", "Fp"):
            out[f"{f}_{k}"] = st.field(f).copy()
    return out


if __name__ == "__main__":
    x, v, m, prm, planes = scene()
    out = run(x, v, m, prm, planes, (1, 10))
    np.savez_compressed(os.path.join(HERE, "oracle_trajectory.npz"), x0=x, v0=v, m=m, **out)
    s = 2 ** -0.5
    wedge = int(((-(out["x_10"][:, 0] - 0.09) + out["x_10"][:, 1]) * s <= 0.0).sum())
    print("wrote", len(x), "particles; plastic at step 10:",
          int((np.abs(np.linalg.det(out["Fp_10"].reshape(-1, 3, 3)) - 1) > 1e-9).sum()),
          "; past the tilted plane at step 10:", wedge)
