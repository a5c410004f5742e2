"""Benchmark: particle-steps/s of the MPM step on the 8M-particle slab.

One step = bin/sort + P2G + grid update + fused G2P over all particles of the
slab (4 x 0.5 x 4 m, 8,000,000 particles, 8 ppc, dx = 0.02, per-particle E,
box collider, dt = 5e-5), inputs resident in HBM before the timed region,
which by default is the workload's 500 steps after 5 warm-up steps.
The roofline entry is the dominant kernel's algorithmic bytes per launch
(84.1 B/particle for P2G, 131.5 B/particle for G2P, FP32 J_P layout) over its
average launch time measured with HIP events on the engine's stream.
The CPU baseline is the FP64 oracle (the reference algorithm restated, same
serial/OpenMP split) timed on a bounded sample of the same slab.

The line also carries the same slab driven into its box (contact, friction
and plastic yielding during the timed steps), the batched form (256 uneven scenes, graph-replayed,
autotuned) and the served form (Poisson arrivals at 10-200 requests/s,
continuous batching) with p50/p99 latency, beside the plain per-scene GPU step.

The semi-implicit entry (the 64k cube with the conjugate-residual velocity
solve) and the seeding entry (mesh -> lattice particles, the reference's
Blender add-on restated on the GPU) are reported beside them, each with its
own CPU baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-serving] [--no-seeding]
                       [--no-implicit]
       python bench.py --batched-only --steps K [--no-multi]
                                                   (the batched form alone, for profiling; --no-multi:
                                                   one engine only, without the two-engine leg)
       python bench.py --issue-probe               (MFMA/VALU issue rates behind the P2G bound)
       python bench.py --stream-probe              (HBM rate of a copy and of G2P's particle streams)
       python bench.py --config-baselines          (configs 0, 1, 3: plain GPU step beside the FP64 oracle)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "stomakhin-style-snow-simulator_amd"))

HBM_PEAK_GBS = 8000.0
MFMA_FP32_PEAK_TFLOPS = 157.3
ALG_BYTES = {"p2g": 84.1, "g2p": 131.5, "grid": 11.5}
# Useful multiply-adds per particle in P2G: per stencil node, mass w m (1),
# momentum w m v (3) and force M grad w (3 x 3): 64 x 13 = 832.
P2G_MACS_PER_PARTICLE = 832


def _kernels_sha():
    import hashlib
    path = os.path.join(ROOT, "stomakhin-style-snow-simulator_amd", "csrc", "mpm_kernels.hip")
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


STEP_KERNELS = {"k_bin": "bin", "k_scatter": "scatter", "k_cellsort": "cellsort", "k_blocks": "blocks",
                "k_p2g": "p2g", "k_grid": "grid", "k_end": "end"}


def write_traffic(fetch_csv, write_csv, steps):
    """profiles/traffic.json from two rocprofv3 counter runs of
    `bench.py --steps K --no-cpu --no-serving --no-seeding --no-implicit --copy-calibration`
    (one with --pmc FETCH_SIZE, one with --pmc WRITE_SIZE): HBM bytes per launch
    of each step kernel, the mean over its last `steps` dispatches (the
    re-sort G2P: its last dispatch), each counter scaled by what it reads for
    the 1 GiB device copy that ends the run."""
    import csv

    def load(path):
        rows = list(csv.DictReader(open(path)))
        out = []
        for r in rows:
            name = r.get("Kernel_Name") or r.get("Kernel-Name") or ""
            val = float(r.get("Counter_Value") or r.get("Counter-Value") or 0.0)
            did = int(r.get("Dispatch_Id") or r.get("Dispatch-Id") or 0)
            out.append((did, name, val))
        out.sort()
        return out

    def per_kernel(rows, calib_bytes):
        copies = [v for _, name, v in rows if "copyBuffer" in name]
        factor = calib_bytes / (copies[-1] * 1024.0)
        got = {}
        for _, name, v in rows:
            base = name.split("(")[0].split("<")[0].split()[-1].replace("snowmpm::", "")
            key = STEP_KERNELS.get(base)
            if base == "k_g2p":  # k_g2p<threads, track, permute, project>
                targs = name.split("<", 1)[1].split(">", 1)[0].split(",")
                key = "g2p_permute" if targs[2].strip() == "true" else "g2p"
            if base == "k_p2g" and "true" in name.split("<", 1)[1].split(">", 1)[0]:
                key = None  # k_p2g<chunk, true>: the semi-implicit solve's force differential
            if key:
                got.setdefault(key, []).append(v * 1024.0 * factor)
        # A re-sort G2P runs every 16th step: its last dispatch is the
        # steady-state one (the first gathers the particles from input order).
        last = {k: (v[-1:] if k == "g2p_permute" else v[-steps:]) for k, v in got.items()}
        return {k: sum(v) / len(v) for k, v in last.items()}, factor

    f, ff = per_kernel(load(fetch_csv), 1 << 30)
    w, wf = per_kernel(load(write_csv), 1 << 30)
    res = {k: int(f.get(k, 0) + w.get(k, 0)) for k in sorted(set(f) | set(w))}
    res["kernels_sha256"] = _kernels_sha()
    res["note"] = (f"HBM bytes per launch of each step kernel on the 8M-particle slab: FETCH_SIZE x {ff:.3f} + "
                   f"WRITE_SIZE x {wf:.3f} (KiB, scaled by the 1 GiB device copy), mean of the last {steps} "
                   "dispatches (g2p_permute: the last re-sort dispatch, a steady-state one)")
    with open(os.path.join(ROOT, "profiles", "traffic.json"), "w") as fo:
        json.dump(res, fo, indent=1)
    print(json.dumps(res))


def _host_info():
    """What the CPU baseline ran on: logical CPUs, the ones this process may
    use, OMP_NUM_THREADS and the CPU model (lscpu)."""
    import subprocess
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}


def cpu_baseline(sc, seed, scale=1.0, steps=3):
    """The FP64 oracle (the reference step restated: std::unordered_map grid
    rebuilt every step, OpenMP on the stencil build, stress update and G2P,
    the other stages serial) on the slab: 1 warm-up + `steps` timed steps on
    all the host threads this process may use, capped by OMP_NUM_THREADS (the
    GPU box's CPU share).  scale < 1 times a smaller slab of the same spacing."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import snowmpm
    import mpmref  # the oracle driver, used here only as the timed CPU baseline
    import oracle as orc

    if scale != 1.0:
        sc = snowmpm.generate_scene(2, seed=seed, scale=scale)
    n = sc["x"].shape[0]
    info = _host_info()
    threads = info["affinity"]
    if info["OMP_NUM_THREADS"]:
        threads = min(threads, int(info["OMP_NUM_THREADS"]))
    prm = sc["params"]
    mu0, la0 = snowmpm.lame(sc["E"], prm.nu)
    t_setup = time.perf_counter()
    st = orc.State(sc["x"], sc["v"], sc["m"], mu0=mu0, lambda0=la0)
    st.setup(prm.dx)
    t_setup = time.perf_counter() - t_setup
    cs = mpmref.orc_colliders(sc["colliders"])
    op = mpmref.orc_params(prm)
    st.run(cs, op, 1, nthreads=threads)  # warm-up step
    t0 = time.perf_counter()
    st.run(cs, op, steps, nthreads=threads)
    dt = time.perf_counter() - t0
    del st
    out = {"value": n * steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
           "sample": f"slab {n} particles{'' if scale == 1.0 else f' (scale {scale})'}: 1 warm-up + {steps} timed "
                     f"steps, {dt:.1f} s (setup {t_setup:.1f} s); FP64 oracle, std::unordered_map grid rebuilt "
                     f"each step, reference serial/OpenMP split; a 500-step run would take "
                     f"{500 * dt / steps / 3600:.2f} h (extrapolated)"}
    out.update(info)
    return out


def config_baselines(seed, no_cpu):
    """The other configurations' plain-method timings beside the FP64 oracle's,
    as the measurement plan lays them out: config 0 (the 64,000-particle cube,
    all 200 steps), config 1 (the two tori, 10 steps of the 2000), config 3
    (16 of the 256 batched scenes at 50 steps each, extrapolated to the
    1000-step scenes/s of all 256).  The GPU figures are the plain one-scene
    step (config 3: per scene, one at a time, the same 16 scenes)."""
    import numpy as np
    import snowmpm
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    info = _host_info()
    threads = min(info["affinity"], int(info["OMP_NUM_THREADS"] or info["affinity"]))
    out = {}

    def gpu_rate(sc, steps):
        n = sc["x"].shape[0]
        with snowmpm.Engine(capacity=n + 1024) as eng:
            eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"], E=sc.get("E"))
            eng.setup()
            eng.step(2)
            t0 = time.perf_counter()
            eng.step(steps)
            return n * steps / (time.perf_counter() - t0)

    def cpu_rate(sc, steps):
        import mpmref  # the oracle driver, used here only as the timed CPU baseline
        import oracle as orc
        prm = sc["params"]
        mu0 = la0 = None
        if sc.get("E") is not None:
            mu0, la0 = snowmpm.lame(sc["E"], prm.nu)
        st = orc.State(sc["x"], sc["v"], sc["m"], mu0=mu0, lambda0=la0)
        st.setup(prm.dx)
        t0 = time.perf_counter()
        st.run(mpmref.orc_colliders(sc["colliders"]), mpmref.orc_params(prm), steps, nthreads=threads)
        return sc["x"].shape[0] * steps / (time.perf_counter() - t0)

    for cfg, steps, label in ((0, 200, "cube 0.4 m, 64,000 particles, all 200 steps"),
                              (1, 10, "two tori, 1,365,824 particles, 10 of the 2000 steps")):
        sc = snowmpm.generate_scene(cfg, seed=seed)
        row = {"workload": label, "gpu_particle_steps_per_s": gpu_rate(sc, 200 if cfg == 0 else 100)}
        if not no_cpu:
            row["cpu_particle_steps_per_s"] = cpu_rate(sc, steps)
            row["gpu_over_cpu"] = row["gpu_particle_steps_per_s"] / row["cpu_particle_steps_per_s"]
        out[f"config{cfg}"] = row
    # config 3: 16 of the 256 scenes, 50 steps each, extrapolated to 1000 steps
    idx = np.linspace(0, 255, 16).astype(int)
    scenes = [snowmpm.generate_scene(3, seed=seed, index=int(i)) for i in idx]
    ns = np.array([s["x"].shape[0] for s in scenes])
    all_ns = np.array([snowmpm.generate_scene(3, seed=seed, index=i)["x"].shape[0] for i in range(256)])
    scale = all_ns.mean() / ns.mean()
    row = {"workload": "16 of the 256 batched scenes, 50 steps each, extrapolated to 1000-step scenes/s "
                       "(scaled by mean size)"}
    gpu_t = sum(s["x"].shape[0] * 50 / gpu_rate(s, 50) for s in scenes)
    row["gpu_plain_scenes_per_s"] = 1.0 / (gpu_t / len(scenes) * 20 * scale)
    if not no_cpu:
        cpu_t = sum(s["x"].shape[0] * 50 / cpu_rate(s, 50) for s in scenes)
        row["cpu_scenes_per_s"] = 1.0 / (cpu_t / len(scenes) * 20 * scale)
        row["gpu_plain_over_cpu"] = row["gpu_plain_scenes_per_s"] / row["cpu_scenes_per_s"]
    out["config3"] = row
    out["cpu"] = {"cores": threads, "kind": "port", "note": "FP64 oracle, the reference's serial/OpenMP split"}
    return out


def issue_probe():
    """SIMD cycles per iteration of 4 independent FP32 MFMAs, of 4 K FP32
    FMAs and of both interleaved, at 1, 2 and 4 waves per SIMD
    (mpm_probe_issue): if the matrix core and the VALU issued side by side
    the mixed loop would cost the larger of the two, not their sum."""
    import snowmpm
    rows = []
    for w in (1, 2, 4):
        for k in (1, 2, 4, 8, 16):
            r = snowmpm.probe_issue(w, k)
            rows.append({"waves_per_simd": w, "valu_per_mfma": k, "mfma_cycles": round(r["mfma"], 1),
                         "valu_cycles": round(r["valu"], 1), "mixed_cycles": round(r["mixed"], 1),
                         "mixed_over_sum": round(r["mixed"] / (r["mfma"] + r["valu"]), 3),
                         "mixed_over_max": round(r["mixed"] / max(r["mfma"], r["valu"]), 3),
                         "kernel_us": [round(u, 1) for u in r["us"]]})
    return rows


def stream_probe():
    """HBM rate with no arithmetic (mpm_probe_stream): a 1 GiB float4 copy,
    and G2P's particle streams on the 8M slab (19 FP32 arrays read, 22
    written) by 512-element workgroups as G2P takes a home block, and one
    element per thread."""
    import snowmpm
    names = {0: "copy 1 GiB, float4", 1: "G2P streams, 128 threads per 512 elements",
             2: "G2P streams, one element per thread"}
    rows = []
    for mode in (0, 1, 2):
        best = None
        for _ in range(3):
            r = snowmpm.probe_stream(mode)
            best = r if best is None or r["us"] < best["us"] else best
        rows.append({"mode": mode, "pattern": names[mode], "us": round(best["us"], 1), "bytes": int(best["bytes"]),
                     "tb_per_s": round(best["tb_per_s"], 3)})
    return rows


def copy_calibration(nbytes=1 << 30):
    """One device-to-device hipMemcpy of known size: the reference point that
    scales rocprofv3's FETCH_SIZE / WRITE_SIZE into HBM bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    t0 = time.perf_counter()
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    dt = time.perf_counter() - t0
    hip.hipFree(a)
    hip.hipFree(b)
    return {"bytes": nbytes, "GB_per_s_read_plus_write": 2 * nbytes / dt / 1e9}


def slab_contact(sc, steps, warmup=10):
    """The same 8M-particle slab driven into its box: the bottom layer starts
    1 mm above the box top, falling at 1.98 m/s (a 0.2 m drop), so after the
    warm-up steps the timed steps run with contact, friction and plastic
    yielding under way at the bottom (reported beside value, not in it)."""
    import numpy as np
    import snowmpm
    x = sc["x"] - np.array([0.0, sc["x"][:, 1].min() - 0.001, 0.0])
    v = np.zeros_like(x)
    v[:, 1] = -1.98
    n = x.shape[0]
    eng = snowmpm.Engine(capacity=n + 1024)
    try:
        sid = eng.add_scene(sc["params"], sc["colliders"], x, v, sc["m"], E=sc["E"])
        eng.setup()
        eng.step(warmup)
        t = snowmpm.StepTimings()
        t0 = time.perf_counter()
        eng.step(steps, t)
        wall = time.perf_counter() - t0
        kt = eng.kernel_times()
        st, fb = eng.status(sid)
        jp = eng.download(sid, fields=("Jp",))["Jp"]
    finally:
        eng.close()
    return {"workload": "slab driven into its box top (1 mm gap, 1.98 m/s), steps %d-%d after the drop starts"
                        % (warmup + 1, warmup + steps),
            "ms_per_step": wall * 1e3 / steps, "particle_steps_per_s": n * steps / wall,
            "kernel_ms": {k: round(val, 4) for k, val in kt.items() if k != "steps"},
            "yielded": int((jp < 1.0).sum()), "status": st, "fp64_fallbacks": fb}


def semi_implicit(seed, no_cpu, steps=20, warmup=2, beta=1.0, tol=1e-6, E=1e7, dt=5e-4):
    """The 64,000-particle cube (falling at 2 m/s onto the ground plane) made
    stiff (E = 1e7) and stepped at dt = 5e-4, about 4x the explicit stability
    limit dx / sqrt(E / rho), with the semi-implicit velocity solve, beta = 1,
    conjugate residuals to |r|_M <= 1e-6 |v*|_M: step time, CR iterations per
    step, and the FP64 oracle's step with the same solve and tolerance as the
    CPU baseline."""
    import numpy as np
    import snowmpm

    sc = snowmpm.generate_scene(0, seed=seed)
    prm = type(sc["params"]).from_buffer_copy(sc["params"])
    prm.beta, prm.cr_tolerance, prm.cr_max_iterations = beta, tol, 200
    prm.mu0, prm.lambda0 = snowmpm.lame(E, prm.nu)
    prm.dt = dt
    n = sc["x"].shape[0]
    iters = []
    # tile sizes and re-sort period from the autotuner on a copy of the scene
    # (results do not depend on them), then a fresh engine for the timed steps
    with snowmpm.Engine(capacity=n + 1024) as eng:
        eng.add_scene(prm, sc["colliders"], sc["x"], sc["v"], sc["m"])
        eng.setup()
        eng.step(warmup)
        tuned = eng.autotune(1)
    with snowmpm.Engine(capacity=n + 1024, p2g_chunk=tuned["p2g_chunk"], g2p_threads=tuned["g2p_threads"],
                        permute_period=tuned["permute_period"]) as eng:
        sid = eng.add_scene(prm, sc["colliders"], sc["x"], sc["v"], sc["m"])
        eng.setup()
        eng.step(warmup)
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.step(1)
            iters.append(eng.cr_stats(sid)["iterations"])
        wall = time.perf_counter() - t0
        st, _ = eng.status(sid)
    out = {"workload": f"cube {n} particles falling onto the ground plane, E {E:g}, dx 0.02, dt {dt:g}, "
                       f"beta {beta:g}, CR to {tol:g} |v*|_M, steps {warmup + 1}-{warmup + steps}",
           "ms_per_step": wall * 1e3 / steps, "particle_steps_per_s": n * steps / wall,
           "cr_iterations_mean": float(np.mean(iters)), "cr_iterations_max": int(max(iters)), "status": st,
           "autotune": tuned}
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import mpmref  # the oracle driver, used here only as the timed CPU baseline
        import oracle as orc
        info = _host_info()
        threads = min(info["affinity"], int(info["OMP_NUM_THREADS"] or info["affinity"]))
        st = orc.State(sc["x"], sc["v"], sc["m"])
        st.setup(prm.dx)
        cs = mpmref.orc_colliders(sc["colliders"])
        op = mpmref.orc_params(prm, cr_tolerance=tol)
        st.run(cs, op, 1, nthreads=threads)
        t0 = time.perf_counter()
        st.run(cs, op, 2, nthreads=threads)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n * 2 / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
                               "sample": f"same cube, 1 warm-up + 2 timed steps, {dt:.1f} s; FP64 oracle, same "
                                         "solve and tolerance"}
    return out


def batched_and_served(seed, n_scenes=256, steps=1000, plain_sample=16, rates=(10, 25, 50, 100, 200),
                       trace_s=2.5, min_requests=64, batched_only=False, multi=True):
    """Batched form (256 uneven scenes in one engine, graph-replayed step,
    autotuned launch shapes) and served form (Poisson arrivals, continuous
    batching) against the plain per-scene GPU step one scene at a time.
    All forms give bitwise-identical per-scene results (tests/test_gpu_batched.py)."""
    import numpy as np
    import snowmpm

    scenes = [snowmpm.generate_scene(3, seed=seed, index=i) for i in range(n_scenes)]
    ns = np.array([len(s["x"]) for s in scenes])
    total = int(ns.sum())
    out = {"scenes": n_scenes, "particles_total": total, "particles_min": int(ns.min()),
           "particles_max": int(ns.max()), "steps_per_scene": steps}
    # batched: all scenes resident, one launch per stage per step
    eng = snowmpm.Engine(capacity=total + 4096)
    for s in scenes:
        eng.add_scene(s["params"], s["colliders"], s["x"], s["v"], s["m"])
    eng.setup()
    out["autotune"] = eng.autotune(2)
    eng.step_graph(5)
    t0 = time.perf_counter()
    eng.step_graph(steps)
    wall = time.perf_counter() - t0
    eng.close()
    out["batched_scenes_per_s"] = n_scenes / wall
    out["batched_ms_per_step"] = wall * 1e3 / steps
    out["batched_particle_steps_per_s"] = total * steps / wall
    if not multi:
        return out
    # the same scenes over two engines (alternating by size), stepped
    # together on two streams (mpm_step_graph_multi)
    halves = [np.argsort(ns)[k::2] for k in range(2)]
    engs = []
    for h in halves:
        e = snowmpm.Engine(capacity=int(ns[h].sum()) + 4096)
        for i in h:
            s = scenes[i]
            e.add_scene(s["params"], s["colliders"], s["x"], s["v"], s["m"])
        e.setup()
        e.autotune(2)
        engs.append(e)
    snowmpm.step_graph_multi(engs, 5)
    t0 = time.perf_counter()
    snowmpm.step_graph_multi(engs, steps)
    wall2 = time.perf_counter() - t0
    for e in engs:
        e.close()
    out["two_engines_scenes_per_s"] = n_scenes / wall2
    out["two_engines_ms_per_step"] = wall2 * 1e3 / steps
    if batched_only:
        return out
    # plain: one scene at a time, plain launches (sampled)
    idx = np.linspace(0, n_scenes - 1, plain_sample).astype(int)
    eng = snowmpm.Engine(capacity=int(ns.max()) + 4096)
    t_plain = 0.0
    for i in idx:
        s = scenes[i]
        t0 = time.perf_counter()
        sid = eng.add_scene(s["params"], s["colliders"], s["x"], s["v"], s["m"])
        eng.step(steps)
        t_plain += time.perf_counter() - t0
        eng.clear()
    eng.close()
    plain_per_scene = t_plain / len(idx) * (ns.mean() / ns[idx].mean())
    out["plain_scenes_per_s"] = 1.0 / plain_per_scene
    out["plain_sample"] = f"{len(idx)} of the {n_scenes} scenes, {steps} steps each, scaled by mean size"
    out["batched_vs_plain"] = out["batched_scenes_per_s"] / out["plain_scenes_per_s"]

    # served: Poisson arrivals, steps ~ U{100..2000}
    rng = np.random.default_rng(seed)
    served = []
    cap = 16_000_000
    eng = snowmpm.Engine(capacity=cap)
    for lam in rates:
        # about trace_s seconds of arrivals (at least min_requests requests)
        nreq = max(min_requests, int(round(lam * trace_s)))
        gaps = rng.exponential(1.0 / lam, size=nreq)
        arr = np.cumsum(gaps) - gaps[0]
        reqs = []
        for k in range(nreq):
            s = scenes[int(rng.integers(0, n_scenes))]
            reqs.append(dict(s, arrival_s=float(arr[k]), steps=int(rng.integers(100, 2001))))
        r = eng.serve(reqs, mode=0, max_chunk=25)
        # steady state: completions between the arrival of the request a
        # quarter into the trace and the last arrival (no ramp-up, no drain)
        t1, t2 = arr[nreq // 4], arr[-1]
        done = np.array(r["done_s"])
        steady = float(((done >= t1) & (done <= t2)).sum() / (t2 - t1))
        served.append({"rate_per_s": lam, "requests": nreq, "scenes_per_s": r["scenes_per_s"],
                       "steady_scenes_per_s": steady, "p50_s": r["p50_s"], "p99_s": r["p99_s"],
                       "mean_s": r["mean_s"], "max_resident": r["max_resident"], "wall_s": r["wall_s"],
                       "chunks": r["chunks"], "steps_per_chunk": r["chunk_steps"] / max(r["chunks"], 1),
                       "time_step_s": r["time_step_s"], "time_admit_s": r["time_admit_s"],
                       "time_retire_s": r["time_retire_s"]})
    eng.close()
    # plain serving at saturation (all requests queued at t=0), sampled
    eng = snowmpm.Engine(capacity=int(ns.max()) + 4096)
    reqs = []
    for k in range(8):
        s = scenes[int(rng.integers(0, n_scenes))]
        reqs.append(dict(s, arrival_s=0.0, steps=int(rng.integers(100, 2001))))
    r = eng.serve(reqs, mode=1)
    eng.close()
    out["served"] = served
    out["served_plain_saturated"] = {"requests": 8, "scenes_per_s": r["scenes_per_s"], "p50_s": r["p50_s"],
                                     "p99_s": r["p99_s"]}
    return out


def _torus_mesh(R, r, nu, nv):
    """Closed torus around the z axis: nu x nv quads as 2 nu nv triangles."""
    import numpy as np
    u = 2 * np.pi * np.arange(nu) / nu
    w = 2 * np.pi * np.arange(nv) / nv
    U, W = np.meshgrid(u, w, indexing="ij")
    v = np.stack([(R + r * np.cos(W)) * np.cos(U), (R + r * np.cos(W)) * np.sin(U), r * np.sin(W)], -1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a, b = i * nv + j, ((i + 1) % nu) * nv + j
    c, d = ((i + 1) % nu) * nv + (j + 1) % nv, i * nv + (j + 1) % nv
    f = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)])
    return v, f.astype(np.int32)


def seeding(no_cpu):
    """Mesh seeding (the Blender add-on's seed_mesh, on the GPU): a torus of
    9,216 triangles on a 1.35M-point lattice; the CPU baseline is the oracle
    on the same mesh at a coarser spacing (the work per lattice point does
    not depend on the spacing)."""
    import numpy as np

    import snowmpm
    v, f = _torus_mesh(0.3, 0.12, 96, 48)
    spacing = 0.005
    snowmpm.seed_mesh(v, f, spacing)  # warm-up (module load, allocations)
    best = None
    for _ in range(3):
        t0 = time.perf_counter()
        pts = snowmpm.seed_mesh(v, f, spacing)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    lo, hi = v.astype(np.float32).min(0), v.astype(np.float32).max(0)
    s32, npts = float(np.float32(spacing)), 1
    for e in range(3):  # lattice size (seed_export.py:109-127)
        k, x = 0, float(lo[e]) + s32 * 0.5
        while x < float(hi[e]):
            k, x = k + 1, x + s32
        npts *= k
    out = {"mesh": "torus R=0.3 r=0.12, 9216 triangles", "spacing": spacing, "lattice_points": npts,
           "particles": int(pts.shape[0]), "call_ms": best * 1e3, "lattice_points_per_s": npts / best,
           "note": "whole call: host lattice, uploads, kernel, download, compaction"}
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc  # timed CPU baseline only
        cs = 0.03
        t0 = time.perf_counter()
        ref = orc.seed_mesh(v, f, cs)
        dt = time.perf_counter() - t0
        nc = 1
        for e in range(3):
            nc *= orc.seed_axis(float(lo[e]), float(hi[e]), cs).shape[0]
        out["cpu_baseline"] = {"value": nc / dt, "unit": "lattice points/s", "cores": 1, "kind": "port",
                               "sample": f"same mesh at spacing {cs}: {nc} lattice points, {ref.shape[0]} "
                                         f"particles, scalar oracle, {dt:.1f} s"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the slab workload is 500 steps long; timing all of them covers the
    # whole fall (velocities grow to 0.25 m/s, particles change cells and
    # blocks between re-sorts), not only the first steps from rest
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-scale", type=float, default=1.0,
                    help="CPU baseline on a slab of this scale (1 = the full 8M slab, ~32 GB of host memory)")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the batched / served measurements and the slab-contact run")
    ap.add_argument("--no-seeding", action="store_true", help="skip the mesh-seeding measurement")
    ap.add_argument("--no-implicit", action="store_true", help="skip the semi-implicit measurement")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched form (256 scenes), --steps graph-replayed steps each (for profiling)")
    ap.add_argument("--no-multi", action="store_true",
                    help="with --batched-only: one engine only, without the two-engine leg")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/traffic.json from two rocprofv3 counter CSVs (no GPU work)")
    ap.add_argument("--issue-probe", action="store_true",
                    help="only the MFMA/VALU issue-rate probe (one JSON line)")
    ap.add_argument("--config-baselines", action="store_true",
                    help="only configs 0, 1 and 3 on the plain GPU step beside the FP64 oracle (a few minutes "
                         "of CPU work; one JSON line)")
    ap.add_argument("--stream-probe", action="store_true",
                    help="only the HBM streaming probe (copy and G2P's particle streams; one JSON line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the run, time one 1 GiB device-to-device copy (traffic-counter calibration)")
    args = ap.parse_args()
    if args.traffic_from:
        write_traffic(args.traffic_from[0], args.traffic_from[1], args.steps)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine runs on a single MI355X and has no multi-GPU path; "
              "scenes are independent, so more GPUs would each run their own share of the scenes",
              file=sys.stderr)
        sys.exit(2)

    import snowmpm

    if args.issue_probe:
        print(json.dumps({"issue_probe": issue_probe()}))
        return
    if args.stream_probe:
        print(json.dumps({"stream_probe": stream_probe()}))
        return
    if args.config_baselines:
        print(json.dumps({"configs_plain": config_baselines(args.seed, args.no_cpu)}))
        return
    if args.batched_only:
        print(json.dumps({"batched": batched_and_served(args.seed, steps=args.steps, batched_only=True,
                                                        multi=not args.no_multi)}))
        return
    t0 = time.perf_counter()
    sc = snowmpm.generate_scene(2, seed=args.seed, scale=1.0)
    n = sc["x"].shape[0]
    eng = snowmpm.Engine(capacity=n + 1024)
    t_up = time.perf_counter()
    sid = eng.add_scene(sc["params"], sc["colliders"], sc["x"], sc["v"], sc["m"], E=sc["E"])
    t_up = time.perf_counter() - t_up
    eng.setup()
    print(f"# slab: {n} particles resident, setup {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    if args.warmup:
        eng.step(args.warmup)
    timings = snowmpm.StepTimings()
    t0 = time.perf_counter()
    eng.step(args.steps, timings)  # synchronizes the device before returning
    wall = time.perf_counter() - t0
    kt = eng.kernel_times()
    period = eng.permute_period()
    st, fb = eng.status(sid)
    t_down = time.perf_counter()
    eng.download(sid)
    t_down = time.perf_counter() - t_down
    t_down_x = time.perf_counter()
    eng.download(sid, fields=("x",))
    t_down_x = time.perf_counter() - t_down_x
    eng.close()

    ms_step = wall * 1e3 / args.steps
    value = n * args.steps / wall
    per_kernel = {k: kt[k] for k in ("p2g", "g2p", "grid")}
    dom = max(("p2g", "g2p"), key=lambda k: per_kernel[k])
    achieved = ALG_BYTES[dom] * n / (per_kernel[dom] * 1e-3) / 1e9
    traffic, traffic_note = None, None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            tj = json.load(f)
        if tj.get("kernels_sha256") == _kernels_sha():
            traffic = tj.get(dom)
            traffic_note = "rocprofv3 FETCH_SIZE/WRITE_SIZE of these kernels (profiles/traffic.json)"
            if dom == "g2p" and tj.get("g2p_permute"):
                # the timed launches mix ordinary and re-sort G2P as the
                # re-sort period does
                traffic = int((tj["g2p"] * (period - 1) + tj["g2p_permute"]) / period)
                traffic_note += (f"; G2P: ordinary and re-sort launches weighted {period - 1}:1 "
                                 f"(re-sort period {period})")
        else:
            traffic_note = ("not measured for this kernel source (profiles/traffic.json was measured for "
                            f"mpm_kernels.hip sha256 {str(tj.get('kernels_sha256'))[:12]})")
    # both particle kernels against HBM; P2G also against the FP32 matrix peak
    kernel_roofline = {}
    for k in ("p2g", "g2p"):
        gbs = ALG_BYTES[k] * n / (per_kernel[k] * 1e-3) / 1e9
        kernel_roofline[k] = {"ms": round(per_kernel[k], 4), "hbm_gbs": round(gbs, 1),
                              "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    tfl = 2.0 * P2G_MACS_PER_PARTICLE * n / (per_kernel["p2g"] * 1e-3) / 1e12
    kernel_roofline["p2g"].update({"mfma_tflops": round(tfl, 2), "mfma_frac": round(tfl / MFMA_FP32_PEAK_TFLOPS, 4)})
    out = {
        "metric": "particle-steps/s (8M-particle slab)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "slab 4x0.5x4 m, 8,000,000 particles (8 ppc, dx=0.02, jitter 0.1), "
                               "per-particle E~U[1e5,2e5], box collider, dt=5e-5, FP32 with FP64 SVD fallback",
                   "particles": n, "layout": "FP32 SoA, J_P (no full F_P)"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_note": traffic_note,
                     "alg_bytes_per_particle": ALG_BYTES[dom]},
        "kernel_ms": {k: round(v, 4) for k, v in kt.items() if k != "steps"},
        "kernel_roofline": kernel_roofline,
        "status": st,
        "fp64_fallbacks": fb,
        # Host buffers crossing PCIe (not part of value): scene upload from double
        # host arrays, download of the whole state and of the positions alone to
        # double host arrays, and the rate when the positions of every 80-step
        # frame are read back, as the reference driver does to export them.
        "host_io": {"upload_s": t_up, "download_s": t_down, "download_x_s": t_down_x,
                    "particle_steps_per_s_with_frame_download":
                        n * 80 / (80 * ms_step * 1e-3 + t_down_x)},
    }
    if not args.no_serving:
        out["slab_contact"] = slab_contact(sc, args.steps)
        out["batched"] = batched_and_served(args.seed)
    if not args.no_implicit:
        out["semi_implicit"] = semi_implicit(args.seed, args.no_cpu)
    if not args.no_seeding:
        out["seeding"] = seeding(args.no_cpu)
    if args.copy_calibration:
        out["copy_calibration"] = copy_calibration()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, args.seed, scale=args.cpu_scale)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
