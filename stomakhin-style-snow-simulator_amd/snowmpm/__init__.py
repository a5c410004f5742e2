"""snowmpm -- Python host binding of the MI355X MPM step engine (libsnowmpm.so).

A thin ctypes layer over the C ABI in include/snowmpm.h.  It mirrors the
reference's solver interface (include/snow/grid.hpp:96-97,
include/snow/collision.hpp:14, include/snow/debug.hpp:3-15): particles are
numpy arrays in the reference's per-particle layout (x, v: n x 3; F_E, F_P:
n x 3 x 3), ``setup`` computes V_p0, ``step`` advances every resident scene,
``make_plane`` builds the reference's collider.

The library is required: importing this module fails loudly when
libsnowmpm.so is missing, and every call runs on the GPU.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libsnowmpm.so")
if not os.path.exists(_LIB):
    raise ImportError(f"libsnowmpm.so not built ({_LIB}); run __graft_entry__.build()")
lib = C.CDLL(_LIB)

OK = 0
ERR_ARG, ERR_HIP, ERR_OOM, ERR_CAPACITY, ERR_STATE, ERR_DEVICE, ERR_IO = -1, -2, -3, -4, -5, -6, -7
STATUS_NAN, STATUS_OVERFLOW, STATUS_HASH_FULL, STATUS_RANGE = 0x1, 0x2, 0x4, 0x8
STATUS_REJECTED = 0x100  # serving: a request larger than the engine's particle pool
MAX_COLLIDERS = 8
PLANE, BOX, SPHERE = 0, 1, 2
PROJECT = 0x1  # Collider.flags: project particles that end a step inside the body onto its surface


class MpmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Params(C.Structure):
    """Scalar arguments of the reference step() (grid.hpp:97) + nu."""
    _fields_ = [("gravity", C.c_double * 3), ("dt", C.c_double), ("dx", C.c_double),
                ("mu0", C.c_double), ("lambda0", C.c_double), ("hardening", C.c_double),
                ("max_compression", C.c_double), ("max_stretch", C.c_double),
                ("alpha", C.c_double), ("nu", C.c_double), ("beta", C.c_double),
                ("cr_tolerance", C.c_double), ("cr_max_iterations", C.c_int)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["gravity"] = list(self.gravity)
        return d


class Collider(C.Structure):
    _fields_ = [("type", C.c_int), ("flags", C.c_int), ("point", C.c_double * 3),
                ("normal", C.c_double * 3), ("half_extent", C.c_double * 3),
                ("radius", C.c_double), ("velocity", C.c_double * 3), ("friction", C.c_double)]


class StepTimings(C.Structure):
    """StepTimings (debug.hpp:3-15), milliseconds accumulated with +=."""
    _fields_ = [(n, C.c_double) for n in (
        "stencil", "p2g", "forces_constitutive", "forces_accumulative", "grid_velocity",
        "grid_collisions", "stress_update", "g2p", "particle_collisions", "position_update")]


class Config(C.Structure):
    _fields_ = [("particle_capacity", C.c_size_t), ("hash_log2", C.c_int), ("track_fp", C.c_int),
                ("device", C.c_int), ("permute_period", C.c_int), ("kappa", C.c_double),
                ("p2g_chunk", C.c_int), ("g2p_threads", C.c_int), ("dense_blocks", C.c_int)]


class Request(C.Structure):
    """mpm_request: one scene request for the serving loop."""
    _fields_ = [("arrival_s", C.c_double), ("steps", C.c_int), ("ncolliders", C.c_int), ("n", C.c_size_t),
                ("x", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double)), ("m", C.POINTER(C.c_double)),
                ("E", C.POINTER(C.c_double)), ("prm", Params), ("colliders", Collider * MAX_COLLIDERS),
                ("out_x", C.POINTER(C.c_double)), ("out_v", C.POINTER(C.c_double)),
                ("t_admit_s", C.c_double), ("t_done_s", C.c_double), ("status", C.c_uint32),
                ("footprint_bytes", C.c_size_t)]


class ServeStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("scenes_per_s", C.c_double), ("p50_s", C.c_double),
                ("p99_s", C.c_double), ("mean_s", C.c_double), ("max_s", C.c_double),
                ("particle_steps", C.c_double), ("completed", C.c_int), ("max_resident", C.c_int),
                ("failed", C.c_int), ("max_resident_bytes", C.c_double), ("time_step_s", C.c_double),
                ("time_admit_s", C.c_double), ("time_retire_s", C.c_double), ("chunks", C.c_int), ("chunk_steps", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_pd = C.POINTER(C.c_double)
_ppd = C.POINTER(_pd)
_SIGS = {
    "mpm_create": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "mpm_destroy": (C.c_int, [C.c_void_p]),
    "mpm_last_error": (C.c_char_p, []),
    "mpm_scene_add": (C.c_int, [C.c_void_p, C.POINTER(Params), C.POINTER(Collider), C.c_int, C.c_size_t,
                                _pd, _pd, _pd, _pd, _pd, _pd, _pd, C.POINTER(C.c_int)]),
    "mpm_scene_remove": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_clear": (C.c_int, [C.c_void_p]),
    "mpm_scene_count": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_size_t)]),
    "mpm_setup": (C.c_int, [C.c_void_p]),
    "mpm_step": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(StepTimings)]),
    "mpm_scene_download": (C.c_int, [C.c_void_p, C.c_int, _pd, _pd, _pd, _pd, _pd, _pd]),
    "mpm_scene_download_noalloc": (C.c_int, [C.c_void_p, C.c_int, _pd, _pd, _pd, _pd, _pd, _pd,
                                             C.POINTER(C.c_size_t)]),
    "mpm_scene_status": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
    "mpm_scene_sums": (C.c_int, [C.c_void_p, C.c_int, _pd]),
    "mpm_kernel_times": (C.c_int, [C.c_void_p, _pd, C.POINTER(C.c_int)]),
    "mpm_slots": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "mpm_probe_issue": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _pd]),
    "mpm_probe_stream": (C.c_int, [C.c_int, C.c_int, _pd]),
    "mpm_scene_steps": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint64)]),
    "mpm_cr_stats": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), _pd, C.c_int]),
    "mpm_request_footprint": (C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_size_t),
                                        C.POINTER(C.c_size_t)]),
    "mpm_step_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "mpm_step_graph_multi": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int]),
    "mpm_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "mpm_permute_period": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "mpm_tile_sizes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mpm_serve": (C.c_int, [C.c_void_p, C.POINTER(Request), C.c_int, C.c_int, C.c_int, C.POINTER(ServeStats)]),
    "mpm_make_plane": (Collider, [_pd, _pd, C.c_double]),
    "mpm_make_box": (Collider, [_pd, _pd, C.c_double]),
    "mpm_make_sphere": (Collider, [_pd, C.c_double, C.c_double]),
    "mpm_import_particles": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_size_t), _ppd, _ppd, _ppd]),
    "mpm_export_frame": (C.c_int, [C.c_char_p, C.c_int, C.c_size_t, _pd]),
    "mpm_free": (None, [C.c_void_p]),
    "mpm_scene_generate": (C.c_int, [C.c_int, C.c_uint64, C.c_int, C.c_double, C.POINTER(C.c_size_t),
                                     _ppd, _ppd, _ppd, _ppd, C.POINTER(Params), C.POINTER(Collider),
                                     C.POINTER(C.c_int)]),
    "mpm_seed_mesh": (C.c_int, [C.c_int, _pd, C.c_size_t, C.POINTER(C.c_int32), C.c_size_t, C.c_double,
                                C.POINTER(C.c_size_t), _ppd]),
}
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

# Symbols declared in include/snowmpm.h (checked by the ABI test).
EXPORTED = tuple(_SIGS)


def last_error():
    return lib.mpm_last_error().decode(errors="replace")


def _check(rc):
    if rc != OK:
        raise MpmError(rc, last_error())
    return rc


def _dptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_pd)


def _as_f64(a, shape):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape)


def _v3(x):
    return (C.c_double * 3)(*map(float, x))


def _flagged(c, project, velocity):
    if project:
        c.flags |= PROJECT
    if velocity is not None:
        c.velocity[:] = list(map(float, velocity))
    return c


def make_plane(point, normal, friction, project=False, velocity=None):
    """make_plane (collision.hpp:14, collision.cpp:5-27).  velocity sets the
    body's constant velocity (CollisionBody::velocity, collision.hpp:10,
    which make_plane leaves at zero); project=True adds the position
    projection, which the reference does not have."""
    return _flagged(lib.mpm_make_plane(_v3(point), _v3(normal), float(friction)), project, velocity)


def make_box(centre, half_extent, friction, project=False, velocity=None):
    return _flagged(lib.mpm_make_box(_v3(centre), _v3(half_extent), float(friction)), project, velocity)


def make_sphere(centre, radius, friction, project=False, velocity=None):
    return _flagged(lib.mpm_make_sphere(_v3(centre), float(radius), float(friction)), project, velocity)


def lame(E, nu):
    """mu_0, lambda_0 from E and nu (main.cpp:28-29)."""
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


def params(dt, dx, mu0, lambda0, hardening, max_compression, max_stretch, alpha, gravity=(0, -9.81, 0), nu=0.2,
           beta=0.0, cr_tolerance=0.0, cr_max_iterations=0):
    """step() scalars (grid.hpp:97) + nu; beta > 0 adds the semi-implicit
    velocity solve (cr_tolerance 0 = 1e-6, cr_max_iterations 0 = 100)."""
    p = Params()
    p.beta, p.cr_tolerance, p.cr_max_iterations = float(beta), float(cr_tolerance), int(cr_max_iterations)
    p.gravity[:] = list(map(float, gravity))
    p.dt, p.dx, p.mu0, p.lambda0 = float(dt), float(dx), float(mu0), float(lambda0)
    p.hardening, p.max_compression, p.max_stretch = float(hardening), float(max_compression), float(max_stretch)
    p.alpha, p.nu = float(alpha), float(nu)
    return p


def _take(ptr, n, cols):
    if not ptr:
        return None
    a = np.ctypeslib.as_array(ptr, shape=(n * cols,)).copy()
    lib.mpm_free(C.cast(ptr, C.c_void_p))
    return a.reshape(n, cols) if cols > 1 else a


def generate_scene(kind, seed=20261203, index=0, scale=1.0):
    """Seeded workload: kind 0 cube, 1 two tori, 2 slab, 3 batched-distribution scene."""
    n = C.c_size_t()
    x, v, m, E = _pd(), _pd(), _pd(), _pd()
    prm = Params()
    coll = (Collider * 8)()
    nc = C.c_int()
    _check(lib.mpm_scene_generate(int(kind), C.c_uint64(seed), int(index), float(scale), C.byref(n),
                                  C.byref(x), C.byref(v), C.byref(m), C.byref(E), C.byref(prm), coll,
                                  C.byref(nc)))
    k = n.value
    out = dict(x=_take(x, k, 3), v=_take(v, k, 3), m=_take(m, k, 1), E=_take(E, k, 1) if E else None,
               params=prm, colliders=[coll[i] for i in range(nc.value)])
    return out


def import_particles(groups_csv, particles_csv):
    """import_particles (import.hpp:25, import.cpp:131-137) -> (x, v, m)."""
    n = C.c_size_t()
    x, v, m = _pd(), _pd(), _pd()
    _check(lib.mpm_import_particles(os.fsencode(groups_csv), os.fsencode(particles_csv), C.byref(n),
                                    C.byref(x), C.byref(v), C.byref(m)))
    k = n.value
    return _take(x, k, 3), _take(v, k, 3), _take(m, k, 1)


def export_frame(directory, frame_no, x):
    """export_particles (export.cpp:10-22) into directory/frame_%06d.csv."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    _check(lib.mpm_export_frame(os.fsencode(directory), int(frame_no), x.shape[0], _dptr(x)))


def step_graph_multi(engines, nsteps=1):
    """step_graph() of several engines together, each launching on its own
    stream so their kernels share the GPU; the scenes' results are those of
    step_graph() on each engine alone."""
    hs = (C.c_void_p * len(engines))(*[e.h for e in engines])
    return _check(lib.mpm_step_graph_multi(hs, len(engines), int(nsteps)))


def probe_issue(waves_per_simd, valu_per_mfma, iterations=4096, device=0):
    """SIMD cycles per iteration of 4 independent FP32 MFMAs (mfma), of
    4 * valu_per_mfma FP32 FMAs (valu) and of both interleaved (mixed), with
    the kernel times in microseconds (see mpm_probe_issue)."""
    out = (C.c_double * 6)()
    _check(lib.mpm_probe_issue(int(device), int(waves_per_simd), int(valu_per_mfma), int(iterations), out))
    return dict(mfma=out[0], valu=out[1], mixed=out[2], us=(out[3], out[4], out[5]))


def probe_stream(mode, device=0):
    """HBM rate with no arithmetic: mode 0 a 1 GiB float4 copy, mode 1 G2P's
    particle streams on the 8M slab (19 FP32 arrays read, 22 written) by
    512-element workgroups, mode 2 the same one element per thread
    (see mpm_probe_stream)."""
    out = (C.c_double * 3)()
    _check(lib.mpm_probe_stream(int(device), int(mode), out))
    return dict(us=out[0], bytes=out[1], tb_per_s=out[2])


def seed_mesh(verts, tris, spacing, device=0):
    """Lattice particles inside a closed triangle mesh, on the GPU.

    seed_mesh / point_inside_mesh / ray_inside of the reference's Blender
    add-on (blender/seed_export.py:43-131): verts (n x 3) and tris (k x 3
    vertex indices) in the mesh's frame; returns the points (m x 3, float32
    values held as float64), x outer and z inner, in that frame.
    """
    v = np.ascontiguousarray(np.asarray(verts, dtype=np.float64).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.int32).reshape(-1, 3))
    n = C.c_size_t()
    x = _pd()
    _check(lib.mpm_seed_mesh(int(device), _dptr(v), v.shape[0], t.ctypes.data_as(C.POINTER(C.c_int32)),
                             t.shape[0], float(spacing), C.byref(n), C.byref(x)))
    if n.value == 0:
        return np.zeros((0, 3))
    return _take(x, n.value, 3)


class Engine:
    """Device-resident MPM engine holding any number of scenes."""

    def __init__(self, capacity, track_fp=False, device=0, hash_log2=0, kappa=0.0, permute_period=0, p2g_chunk=0,
                 g2p_threads=0, dense_blocks=False):
        cfg = Config(particle_capacity=int(capacity), hash_log2=int(hash_log2), track_fp=int(bool(track_fp)),
                     device=int(device), permute_period=int(permute_period), kappa=float(kappa),
                     p2g_chunk=int(p2g_chunk), g2p_threads=int(g2p_threads), dense_blocks=int(bool(dense_blocks)))
        h = C.c_void_p()
        _check(lib.mpm_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.track_fp = bool(track_fp)

    def close(self):
        if getattr(self, "h", None):
            lib.mpm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_scene(self, prm, colliders, x, v, m, V0=None, Fe=None, Fp=None, E=None):
        x = _as_f64(x, (-1, 3))
        n = x.shape[0]
        v = _as_f64(v, (n, 3))
        m = _as_f64(np.broadcast_to(np.asarray(m, dtype=np.float64), (n,)), (n,))
        V0 = None if V0 is None else _as_f64(np.broadcast_to(np.asarray(V0, dtype=np.float64), (n,)), (n,))
        Fe = _as_f64(Fe, (n, 9))
        Fp = _as_f64(Fp, (n, 9))
        E = None if E is None else _as_f64(np.broadcast_to(np.asarray(E, dtype=np.float64), (n,)), (n,))
        arr = (Collider * max(1, len(colliders)))(*colliders)
        sid = C.c_int()
        _check(lib.mpm_scene_add(self.h, C.byref(prm), arr, len(colliders), n, _dptr(x), _dptr(v), _dptr(m),
                                 _dptr(V0), _dptr(Fe), _dptr(Fp), _dptr(E), C.byref(sid)))
        return sid.value

    def remove_scene(self, sid):
        _check(lib.mpm_scene_remove(self.h, int(sid)))

    def clear(self):
        _check(lib.mpm_clear(self.h))

    def count(self, sid):
        n = C.c_size_t()
        _check(lib.mpm_scene_count(self.h, int(sid), C.byref(n)))
        return n.value

    def setup(self):
        """setup() (grid.cpp:340-357) for every scene added without V0."""
        _check(lib.mpm_setup(self.h))

    def step(self, nsteps=1, timings=None):
        """nsteps x step() (grid.cpp:359-450) on every active scene."""
        return _check(lib.mpm_step(self.h, int(nsteps), C.byref(timings) if timings is not None else None))

    def step_graph(self, nsteps=1):
        """Same steps as step(), replayed from a captured hipGraph."""
        return _check(lib.mpm_step_graph(self.h, int(nsteps)))

    def autotune(self, steps_per_config=2):
        chosen = (C.c_int * 4)()
        _check(lib.mpm_autotune(self.h, int(steps_per_config), chosen))
        return dict(grid_p=chosen[0], grid_b=chosen[1], grid_w=chosen[2], grid_k=chosen[3],
                    permute_period=self.permute_period(), **self.tile_sizes())

    def tile_sizes(self):
        """P2G staging chunk and G2P workgroup size in use."""
        c, t = C.c_int(), C.c_int()
        _check(lib.mpm_tile_sizes(self.h, C.byref(c), C.byref(t)))
        return dict(p2g_chunk=c.value, g2p_threads=t.value)

    def permute_period(self):
        """Steps between full re-sorts of the particle buffers."""
        p = C.c_int()
        _check(lib.mpm_permute_period(self.h, C.byref(p)))
        return p.value

    @staticmethod
    def _requests(requests, keep_outputs=False):
        arr = (Request * max(1, len(requests)))()
        keep = []
        outs = []
        for k, q in enumerate(requests):
            r = arr[k]
            r.arrival_s = float(q.get("arrival_s", 0.0))
            r.steps = int(q.get("steps", 0))
            x = np.ascontiguousarray(q["x"], dtype=np.float64)
            v = np.ascontiguousarray(q["v"], dtype=np.float64)
            m = np.ascontiguousarray(np.broadcast_to(np.asarray(q["m"], dtype=np.float64), (x.shape[0],)))
            E = None if q.get("E") is None else np.ascontiguousarray(q["E"], dtype=np.float64)
            keep += [x, v, m, E]
            r.n = x.shape[0]
            r.x, r.v, r.m, r.E = _dptr(x), _dptr(v), _dptr(m), _dptr(E)
            r.prm = q["params"]
            cs = q["colliders"]
            if len(cs) > MAX_COLLIDERS:
                raise MpmError(ERR_ARG, f"a request holds at most {MAX_COLLIDERS} colliders")
            r.ncolliders = len(cs)
            for j, c in enumerate(cs):
                r.colliders[j] = c
            if keep_outputs:
                ox = np.zeros_like(x)
                ov = np.zeros_like(v)
                outs.append((ox, ov))
                keep += [ox, ov]
                r.out_x, r.out_v = _dptr(ox), _dptr(ov)
            else:
                outs.append(None)
        return arr, keep, outs

    def footprint(self, request):
        """Device bytes a request would reserve, and the engine's pools
        (particles, blocks, partial sums) in bytes."""
        arr, keep, _ = self._requests([request])
        b = C.c_size_t()
        pools = (C.c_size_t * 3)()
        _check(lib.mpm_request_footprint(self.h, arr, C.byref(b), pools))
        return b.value, tuple(pools)

    def serve(self, requests, mode=0, max_chunk=25, keep_outputs=False):
        """Runs the request loop.  requests: dicts with arrival_s, steps and a
        generate_scene() payload (x, v, m, E, params, colliders).  mode 0 is
        continuous batching, mode 1 the plain per-scene step one at a time.
        Per request the result lists the latency, the status (0, or the
        device status bits that retired it) and the reserved footprint."""
        arr, keep, outs = self._requests(requests, keep_outputs)
        st = ServeStats()
        _check(lib.mpm_serve(self.h, arr, len(requests), int(mode), int(max_chunk), C.byref(st)))
        res = st.as_dict()
        res["latency_s"] = [arr[k].t_done_s - arr[k].arrival_s for k in range(len(requests))]
        res["status"] = [arr[k].status for k in range(len(requests))]
        res["footprint_bytes"] = [arr[k].footprint_bytes for k in range(len(requests))]
        res["admit_s"] = [arr[k].t_admit_s for k in range(len(requests))]
        res["done_s"] = [arr[k].t_done_s for k in range(len(requests))]
        if keep_outputs:
            res["outputs"] = outs
        return res

    def cr_stats(self, sid):
        """Iterations and |r|_M history of the scene's last semi-implicit solve."""
        it = C.c_int()
        res = (C.c_double * 257)()
        _check(lib.mpm_cr_stats(self.h, int(sid), C.byref(it), res, 257))
        return dict(iterations=it.value, residual=np.array(res[:it.value + 1]))

    def steps(self, sid):
        """Steps the scene has taken since it was added."""
        n = C.c_uint64()
        _check(lib.mpm_scene_steps(self.h, int(sid), C.byref(n)))
        return n.value

    def step_rc(self, nsteps=1):
        """Like step() but returns the status code instead of raising."""
        return lib.mpm_step(self.h, int(nsteps), None)

    def _download(self, sid, fields, noalloc):
        n = self.count(sid)
        want = set(fields) if fields is not None else {"x", "v", "Fe", "Jp", "V0", "Fp"}
        x = np.zeros((n, 3)) if "x" in want else None
        v = np.zeros((n, 3)) if "v" in want else None
        Fe = np.zeros((n, 9)) if "Fe" in want else None
        Jp = np.zeros(n) if "Jp" in want else None
        V0 = np.zeros(n) if "V0" in want else None
        Fp = np.zeros((n, 9)) if self.track_fp and "Fp" in want else None
        ptrs = (_dptr(x), _dptr(v), _dptr(Fe), _dptr(Fp), _dptr(Jp), _dptr(V0))
        if noalloc:
            _check(lib.mpm_scene_download_noalloc(self.h, int(sid), *ptrs, None))
        else:
            _check(lib.mpm_scene_download(self.h, int(sid), *ptrs))
        out = {k: a for k, a in (("x", x), ("v", v), ("Jp", Jp), ("V0", V0)) if a is not None}
        if Fe is not None:
            out["Fe"] = Fe.reshape(n, 3, 3)
        if Fp is not None:
            out["Fp"] = Fp.reshape(n, 3, 3)
        return out

    def download(self, sid, fields=None):
        """The scene's state in its original particle order: x, v (n x 3), Fe
        (and Fp with track_fp) (n x 3 x 3), Jp, V0 (n); fields limits it to
        the names given (only those are gathered and copied)."""
        return self._download(sid, fields, False)

    def download_noalloc(self, sid, fields=None):
        """download() staged through the library's static device buffer
        (mpm_scene_download_noalloc: no device allocation, 32 MiB limit)."""
        return self._download(sid, fields, True)

    def status(self, sid):
        st = C.c_uint32()
        fb = C.c_uint64()
        _check(lib.mpm_scene_status(self.h, int(sid), C.byref(st), C.byref(fb)))
        return st.value, fb.value

    def sums(self, sid):
        out = (C.c_double * 7)()
        _check(lib.mpm_scene_sums(self.h, int(sid), out))
        return np.array(out[:])

    def slots(self):
        """(slots in use, slots per buffer): in use counts up to the last
        particle, holes of the block-aligned layout and a removed scene's
        particles (until the next step) included (see mpm_slots)."""
        used, total = C.c_size_t(), C.c_size_t()
        _check(lib.mpm_slots(self.h, C.byref(used), C.byref(total)))
        return used.value, total.value

    def kernel_times(self):
        ms = (C.c_double * 5)()
        n = C.c_int()
        _check(lib.mpm_kernel_times(self.h, ms, C.byref(n)))
        return dict(bin_sort=ms[0], p2g=ms[1], grid=ms[2], g2p=ms[3], end=ms[4], steps=n.value)
