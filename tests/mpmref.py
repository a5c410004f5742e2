"""Shared helpers: run the same seeded scene through the GPU engine (C ABI)
and the FP64 oracle, and compare them with the error metrics used throughout:
|x - x_ref| / max(|x_ref|, L_scene) and |v - v_ref| / max(|v_ref|, v_scale)."""
import numpy as np

import oracle as orc
import snowmpm


def orc_colliders(cs):
    out = []
    for c in cs:
        oc = orc.Collider()
        for name, _ in orc.Collider._fields_:
            setattr(oc, name, getattr(c, name))
        out.append(oc)
    return out


def orc_params(p):
    return orc.params(p.dt, p.dx, p.mu0, p.lambda0, p.hardening, p.max_compression, p.max_stretch,
                      p.alpha, gravity=tuple(p.gravity))


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


def rel_errors(got, ref, sc):
    L = float(np.ptp(sc["x"], axis=0).max())
    vscale = max(float(np.linalg.norm(sc["v"], axis=1).max()), 1.0)
    ex = np.linalg.norm(got["x"] - ref["x"], axis=1) / np.maximum(np.linalg.norm(ref["x"], axis=1), L)
    ev = np.linalg.norm(got["v"] - ref["v"], axis=1) / np.maximum(np.linalg.norm(ref["v"], axis=1), vscale)
    return float(ex.max()), float(ev.max())
