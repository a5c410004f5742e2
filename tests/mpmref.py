"""Shared helpers: run the same seeded scene through the GPU engine (C ABI)
and the FP64 oracle, and compare them.

Error metrics (rel_errors), per scene after nsteps steps:
  ev = max_p |v_p - v_ref,p| / max_p |v_ref,p|
       (relative to the scene's largest speed; no floor)
  ex = max_p max(0, |x_p - x_ref,p| - a_p) / max_p |x_ref,p - x0,p|
       (relative to the scene's largest displacement, after an allowance
       a_p for storing positions in FP32)
The allowance: the engine keeps x in FP32, so the upload and each step's
x += dt v round x to half an ulp, at most u |x| per component with
u = 2^-24; over nsteps steps that is a_p = sqrt(3) (nsteps + 1) u
(max_c |x_ref,p,c| + D), D the largest displacement.  For a slab at rest
the 10-step displacement (1.2e-6 m) is itself about 5 ulps of x = 2 m, so
no FP32 position store can follow it to 1e-4 of the displacement; whatever
error exceeds that rounding must.  The velocity, from which the
displacement is integrated, carries no such allowance.
scene_errors keeps the scene-size metric |x - x_ref| / max(|x_ref|, L)."""
import numpy as np

import oracle as orc
import snowmpm

U32 = 2.0 ** -24  # FP32 unit roundoff


def orc_colliders(cs):
    out = []
    for c in cs:
        oc = orc.Collider()
        for name, _ in orc.Collider._fields_:
            setattr(oc, name, getattr(c, name))
        out.append(oc)
    return out


def orc_params(p, cr_tolerance=None):
    """The oracle's parameters for GPU parameters p; the semi-implicit solve
    converges the oracle to cr_tolerance (default: 1e-4 of the GPU's)."""
    tol = p.cr_tolerance if p.cr_tolerance > 0 else 1e-6
    return orc.params(p.dt, p.dx, p.mu0, p.lambda0, p.hardening, p.max_compression, p.max_stretch,
                      p.alpha, gravity=tuple(p.gravity), beta=p.beta,
                      cr_tolerance=cr_tolerance if cr_tolerance is not None else tol * 1e-4,
                      cr_max_iterations=256)


def run_oracle(sc, nsteps, V0=None, Fe=None, Fp=None, nthreads=0):
    prm = sc["params"]
    mu0 = la0 = None
    if sc.get("E") is not None:
        mu0, la0 = snowmpm.lame(sc["E"], prm.nu)
    st = orc.State(sc["x"], sc["v"], sc["m"], V0=V0, Fe=Fe, Fp=Fp, mu0=mu0, lambda0=la0)
    if V0 is None:
        st.setup(prm.dx)
    st.run(orc_colliders(sc["colliders"]), orc_params(prm), nsteps, nthreads=nthreads)
    return dict(x=st.field("x"), v=st.field("v"), Fe=st.field("Fe").reshape(-1, 3, 3),
                Fp=st.field("Fp").reshape(-1, 3, 3), V0=st.field("V0"))


def run_gpu(sc, nsteps, track_fp=False, V0=None, Fe=None, Fp=None, engine=None):
    n = sc["x"].shape[0]
    own = engine is None
    eng = engine or snowmpm.Engine(capacity=n + 64, track_fp=track_fp)
    try:
        sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"], V0=V0, Fe=Fe, Fp=Fp,
                            E=sc.get("E"))
        if nsteps:
            eng.step(nsteps)
        else:
            eng.setup()
        out = eng.download(sid)
        out["status"] = eng.status(sid)
        return out
    finally:
        if own:
            eng.close()


def x_allowance(ref_x, x0, nsteps):
    """Per-particle bound on what FP32 storage of positions alone loses over
    nsteps steps (see the module docstring)."""
    D = float(np.linalg.norm(ref_x - x0, axis=1).max())
    return np.sqrt(3.0) * (nsteps + 1) * U32 * (np.abs(ref_x).max(axis=1) + D), D


def rel_errors(got, ref, sc, nsteps):
    """(ex, ev) as defined in the module docstring."""
    a, D = x_allowance(ref["x"], sc["x"], nsteps)
    dx = np.linalg.norm(got["x"] - ref["x"], axis=1)
    ex = float(np.maximum(dx - a, 0.0).max() / D) if D > 0 else float(dx.max() > a.max())
    vmax = float(np.linalg.norm(ref["v"], axis=1).max())
    ev = float(np.linalg.norm(got["v"] - ref["v"], axis=1).max() / vmax)
    return ex, ev


def x_rounding_used(got, ref, sc, nsteps):
    """Largest position error as a fraction of its FP32 storage allowance
    (<= 1: the position error is what storing x in FP32 accounts for)."""
    a, _ = x_allowance(ref["x"], sc["x"], nsteps)
    return float((np.linalg.norm(got["x"] - ref["x"], axis=1) / a).max())


def scene_errors(got, ref, sc):
    """Scene-size metric: |x - x_ref| / max(|x_ref|, L), |v - v_ref| / max|v_ref|."""
    L = float(np.ptp(sc["x"], axis=0).max())
    ex = np.linalg.norm(got["x"] - ref["x"], axis=1) / np.maximum(np.linalg.norm(ref["x"], axis=1), L)
    vmax = float(np.linalg.norm(ref["v"], axis=1).max())
    return float(ex.max()), float(np.linalg.norm(got["v"] - ref["v"], axis=1).max() / vmax)
