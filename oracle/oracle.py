"""ctypes binding of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product (libsnowmpm.so and the snowmpm package) never imports this module.

The C++ source (snow_oracle.cpp) restates the reference solver
src/grid.cpp:15-450 and src/collision.cpp:5-27 in FP64.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class Particle(C.Structure):
    _fields_ = [("x", C.c_double * 3), ("v", C.c_double * 3), ("m", C.c_double),
                ("V0", C.c_double), ("Fp", C.c_double * 9), ("Fe", C.c_double * 9),
                ("mu0", C.c_double), ("lambda0", C.c_double)]


class Collider(C.Structure):
    _fields_ = [("type", C.c_int), ("flags", C.c_int), ("point", C.c_double * 3),
                ("normal", C.c_double * 3), ("half_extent", C.c_double * 3),
                ("radius", C.c_double), ("velocity", C.c_double * 3), ("friction", C.c_double)]


class Params(C.Structure):
    _fields_ = [("gravity", C.c_double * 3), ("dt", C.c_double), ("dx", C.c_double),
                ("mu0", C.c_double), ("lambda0", C.c_double), ("hardening", C.c_double),
                ("max_compression", C.c_double), ("max_stretch", C.c_double),
                ("alpha", C.c_double), ("beta", C.c_double), ("cr_tolerance", C.c_double),
                ("cr_max_iterations", C.c_int)]


class CRStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("nodes", C.c_int), ("residual", C.c_double * 257)]


class Timings(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "stencil", "p2g", "forces_constitutive", "forces_accumulative", "grid_velocity",
        "grid_collisions", "stress_update", "g2p", "particle_collisions", "position_update")]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = C.CDLL(_LIB_PATH)
    d, pd, i = C.c_double, C.POINTER(C.c_double), C.c_int
    P, G = C.POINTER(Particle), C.c_void_p
    sig = {
        "orc_N": (d, [d, d]), "orc_N_dash": (d, [d, d]),
        "orc_det3": (d, [pd]), "orc_inv3": (i, [pd, pd]),
        "orc_svd3": (i, [pd, pd, pd, pd]),
        "orc_make_plane": (Collider, [pd, pd, d]),
        "orc_make_box": (Collider, [pd, pd, d]),
        "orc_make_sphere": (Collider, [pd, d, d]),
        "orc_collider_phi": (d, [C.POINTER(Collider), pd]),
        "orc_collision_helper": (None, [pd, pd, C.POINTER(Collider), pd]),
        "orc_grid_new": (G, []), "orc_grid_free": (None, [G]),
        "orc_grid_size": (C.c_size_t, [G]),
        "orc_grid_set": (None, [G, i, i, i, d, pd, pd, pd]),
        "orc_grid_get": (i, [G, i, i, i, pd, pd, pd, pd]),
        "orc_grid_dump": (C.c_size_t, [G, C.POINTER(C.c_int), pd, pd, pd, pd]),
        "orc_build_stencil": (None, [P, d, C.POINTER(C.c_int), pd, pd]),
        "orc_p2g": (None, [G, P, d]),
        "orc_calculate_volume": (None, [G, P, d]),
        "orc_force_constitutive": (None, [P, d, d, d, pd]),
        "orc_force_calculation": (None, [G, P, d, d, d, d]),
        "orc_grid_velocity": (None, [G, pd, d]),
        "orc_grid_collisions": (None, [G, C.POINTER(Collider), i, d]),
        "orc_stress_update": (None, [G, P, d, d, d, d]),
        "orc_g2p": (None, [G, P, d, d]),
        "orc_particle_collisions": (None, [P, C.POINTER(Collider), i]),
        "orc_position_update": (None, [P, d]),
        "orc_position_projection": (None, [P, C.POINTER(Collider), i]),
        "orc_constitutive_differential": (None, [P, pd, d, d, d, pd]),
        "orc_krylov_nodes": (C.c_size_t, [G, C.POINTER(C.c_int)]),
        "orc_apply_A": (None, [G, P, C.c_size_t, C.POINTER(Params), pd, pd]),
        "orc_last_cr_stats": (None, [C.POINTER(CRStats)]),
        "orc_setup": (None, [P, C.c_size_t, d]),
        "orc_step": (None, [P, C.c_size_t, C.POINTER(Collider), i, C.POINTER(Params),
                            C.POINTER(Timings), i, G]),
        "orc_run": (None, [P, C.c_size_t, C.POINTER(Collider), i, C.POINTER(Params), i,
                           C.POINTER(Timings), i]),
        "orc_seed_axis": (C.c_size_t, [C.c_float, C.c_float, d, C.POINTER(C.c_float), C.c_size_t]),
        "orc_seed_direction": (None, [i, C.POINTER(C.c_float)]),
        "orc_seed_ray_hits": (i, [pd, C.c_size_t, C.POINTER(C.c_int), C.c_size_t, pd, i]),
        "orc_seed_mesh": (C.c_long, [pd, C.c_size_t, C.POINTER(C.c_int), C.c_size_t, d, pd, C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _dp(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def vec(*xs):
    return (C.c_double * len(xs))(*xs)


def N(a, b):
    return lib.orc_N(a, b)


def N_dash(g, p):
    return lib.orc_N_dash(g, p)


def det3(A):
    A, p = _dp(A)
    return lib.orc_det3(p)


def svd3(A):
    A, pa = _dp(np.asarray(A, dtype=np.float64).reshape(9))
    U = np.zeros(9)
    S = np.zeros(3)
    V = np.zeros(9)
    lib.orc_svd3(pa, U.ctypes.data_as(C.POINTER(C.c_double)), S.ctypes.data_as(C.POINTER(C.c_double)),
                 V.ctypes.data_as(C.POINTER(C.c_double)))
    return U.reshape(3, 3), S, V.reshape(3, 3)


def make_particle(x, v, m, V0, Fe=None, Fp=None, mu0=-1.0, lambda0=-1.0):
    """Particle{x, v, m, V0} with F_P = F_E = I (particle.hpp:14-25)."""
    p = Particle()
    p.x[:] = list(map(float, x))
    p.v[:] = list(map(float, v))
    p.m = float(m)
    p.V0 = float(V0)
    eye = [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0]
    p.Fe[:] = list(np.asarray(Fe, dtype=float).reshape(9)) if Fe is not None else eye
    p.Fp[:] = list(np.asarray(Fp, dtype=float).reshape(9)) if Fp is not None else eye
    p.mu0 = mu0
    p.lambda0 = lambda0
    return p


def _moving(c, velocity):
    if velocity is not None:
        c.velocity[:] = list(map(float, velocity))
    return c


def make_plane(point, normal, friction, velocity=None):
    return _moving(lib.orc_make_plane(vec(*point), vec(*normal), float(friction)), velocity)


def make_box(centre, half_extent, friction, velocity=None):
    return _moving(lib.orc_make_box(vec(*centre), vec(*half_extent), float(friction)), velocity)


def make_sphere(centre, radius, friction, velocity=None):
    return _moving(lib.orc_make_sphere(vec(*centre), float(radius), float(friction)), velocity)


def position_projection(p, cs):
    """Optional projection of particle p out of the flagged colliders cs."""
    arr, n = colliders_array(cs)
    lib.orc_position_projection(C.byref(p), arr, n)


def colliders_array(cs):
    arr = (Collider * max(1, len(cs)))()
    for k, c in enumerate(cs):
        arr[k] = c
    return arr, len(cs)


class Grid:
    """The std::unordered_map sparse grid (grid.hpp:72)."""

    def __init__(self):
        self.h = lib.orc_grid_new()

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_grid_free(self.h)
            self.h = None

    def __len__(self):
        return lib.orc_grid_size(self.h)

    def set(self, ijk, mass=0.0, momentum=(0, 0, 0), velocity=(0, 0, 0), force=(0, 0, 0)):
        lib.orc_grid_set(self.h, *map(int, ijk), float(mass), vec(*momentum), vec(*velocity), vec(*force))

    def get(self, ijk):
        m = C.c_double()
        mom, vel, f = vec(0, 0, 0), vec(0, 0, 0), vec(0, 0, 0)
        ok = lib.orc_grid_get(self.h, *map(int, ijk), C.byref(m), mom, vel, f)
        if not ok:
            return None
        return dict(mass=m.value, momentum=np.array(mom[:]), velocity=np.array(vel[:]), force=np.array(f[:]))

    def dump(self):
        n = len(self)
        idx = np.zeros((n, 3), dtype=np.int32)
        mass = np.zeros(n)
        mom = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        f = np.zeros((n, 3))
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        lib.orc_grid_dump(self.h, idx.ctypes.data_as(C.POINTER(C.c_int)), dp(mass), dp(mom), dp(vel), dp(f))
        return dict(idx=idx, mass=mass, momentum=mom, velocity=vel, force=f)


def params(dt, dx, mu0, lambda0, hardening, max_compression, max_stretch, alpha, gravity=(0, -9.81, 0),
           beta=0.0, cr_tolerance=1e-10, cr_max_iterations=200):
    """step() scalars (grid.hpp:97); beta > 0 adds the semi-implicit solve."""
    p = Params()
    p.gravity[:] = list(map(float, gravity))
    p.dt, p.dx, p.mu0, p.lambda0 = float(dt), float(dx), float(mu0), float(lambda0)
    p.hardening, p.max_compression, p.max_stretch, p.alpha = (
        float(hardening), float(max_compression), float(max_stretch), float(alpha))
    p.beta, p.cr_tolerance, p.cr_max_iterations = float(beta), float(cr_tolerance), int(cr_max_iterations)
    return p


def constitutive_differential(p, dF, mu0, lambda0, xi):
    out = (C.c_double * 9)()
    _, dptr = _dp(np.asarray(dF, dtype=float).reshape(9))
    lib.orc_constitutive_differential(C.byref(p), dptr, float(mu0), float(lambda0), float(xi), out)
    return np.array(out[:]).reshape(3, 3)


def last_cr_stats():
    st = CRStats()
    lib.orc_last_cr_stats(C.byref(st))
    return dict(iterations=st.iterations, nodes=st.nodes, residual=np.array(st.residual[:st.iterations + 1]))


def lame(E, nu):
    """mu_0 and lambda_0 from Young's modulus and Poisson ratio (main.cpp:28-29)."""
    mu = E / (2 * (1 + nu))
    lam = (E * nu) / ((1 + nu) * (1 - 2 * nu))
    return mu, lam


class State:
    """A particle vector held as an array of oracle Particle structs."""

    def __init__(self, x, v, m, V0=None, Fe=None, Fp=None, mu0=None, lambda0=None):
        n = len(x)
        self.n = n
        self.arr = (Particle * n)()
        x = np.asarray(x, dtype=float)
        v = np.asarray(v, dtype=float)
        m = np.broadcast_to(np.asarray(m, dtype=float), (n,))
        V0 = np.zeros(n) if V0 is None else np.broadcast_to(np.asarray(V0, dtype=float), (n,))
        eye = np.eye(3).reshape(9)
        Fe = np.broadcast_to(eye, (n, 9)) if Fe is None else np.asarray(Fe, dtype=float).reshape(n, 9)
        Fp = np.broadcast_to(eye, (n, 9)) if Fp is None else np.asarray(Fp, dtype=float).reshape(n, 9)
        mu0 = np.full(n, -1.0) if mu0 is None else np.broadcast_to(np.asarray(mu0, dtype=float), (n,))
        lambda0 = np.full(n, -1.0) if lambda0 is None else np.broadcast_to(np.asarray(lambda0, dtype=float), (n,))
        buf = np.frombuffer(self.arr, dtype=np.float64).reshape(n, 28)
        buf[:, 0:3] = x
        buf[:, 3:6] = v
        buf[:, 6] = m
        buf[:, 7] = V0
        buf[:, 8:17] = Fp
        buf[:, 17:26] = Fe
        buf[:, 26] = mu0
        buf[:, 27] = lambda0

    @property
    def raw(self):
        return np.frombuffer(self.arr, dtype=np.float64).reshape(self.n, 28)

    def field(self, name):
        r = self.raw
        sl = {"x": slice(0, 3), "v": slice(3, 6), "m": slice(6, 7), "V0": slice(7, 8),
              "Fp": slice(8, 17), "Fe": slice(17, 26)}[name]
        out = r[:, sl].copy()
        return out[:, 0] if out.shape[1] == 1 else out

    def setup(self, dx):
        lib.orc_setup(self.arr, self.n, float(dx))

    def step(self, colliders, prm, nthreads=0, grid=None, timings=None):
        arr, nc = colliders_array(colliders)
        lib.orc_step(self.arr, self.n, arr, nc, C.byref(prm), C.byref(timings) if timings else None,
                     int(nthreads), grid.h if grid is not None else None)

    def krylov_nodes(self, grid):
        n = lib.orc_krylov_nodes(grid.h, None)
        idx = np.zeros((n, 3), dtype=np.int32)
        lib.orc_krylov_nodes(grid.h, idx.ctypes.data_as(C.POINTER(C.c_int)))
        return idx

    def apply_A(self, grid, prm, q):
        q = np.ascontiguousarray(q, dtype=float)
        y = np.zeros_like(q)
        lib.orc_apply_A(grid.h, self.arr, self.n, C.byref(prm), q.ctypes.data_as(C.POINTER(C.c_double)),
                        y.ctypes.data_as(C.POINTER(C.c_double)))
        return y

    def copy(self):
        c = State.__new__(State)
        c.n = self.n
        c.arr = (Particle * self.n)()
        C.memmove(c.arr, self.arr, C.sizeof(self.arr))
        return c

    def run(self, colliders, prm, nsteps, nthreads=0, timings=None):
        arr, nc = colliders_array(colliders)
        lib.orc_run(self.arr, self.n, arr, nc, C.byref(prm), int(nsteps),
                    C.byref(timings) if timings else None, int(nthreads))


# ---- mesh seeding (seed_oracle.cpp, blender/seed_export.py:43-131) ----

def _mesh(verts, tris):
    v = np.ascontiguousarray(np.asarray(verts, dtype=np.float64).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.int32).reshape(-1, 3))
    return v, t, v.ctypes.data_as(C.POINTER(C.c_double)), t.ctypes.data_as(C.POINTER(C.c_int))


def seed_axis(lo, hi, spacing):
    """Lattice coordinates of one axis (float32 values)."""
    n = lib.orc_seed_axis(float(np.float32(lo)), float(np.float32(hi)), float(spacing), None, 0)
    out = np.zeros(n, dtype=np.float32)
    lib.orc_seed_axis(float(np.float32(lo)), float(np.float32(hi)), float(spacing),
                      out.ctypes.data_as(C.POINTER(C.c_float)), n)
    return out


def seed_direction(k):
    d = (C.c_float * 3)()
    lib.orc_seed_direction(int(k), d)
    return np.array(list(d), dtype=np.float32)


def seed_ray_hits(verts, tris, p, k):
    v, t, pv, pt = _mesh(verts, tris)
    return lib.orc_seed_ray_hits(pv, v.shape[0], pt, t.shape[0], vec(*map(float, p)), int(k))


def seed_mesh(verts, tris, spacing):
    """Seeded points (n x 3, float64 holding float32 values), x outer, z inner."""
    v, t, pv, pt = _mesh(verts, tris)
    n = lib.orc_seed_mesh(pv, v.shape[0], pt, t.shape[0], float(spacing), None, 0)
    if n < 0:
        raise ValueError("bad mesh or spacing")
    out = np.zeros((n, 3))
    lib.orc_seed_mesh(pv, v.shape[0], pt, t.shape[0], float(spacing), out.ctypes.data_as(C.POINTER(C.c_double)), n)
    return out
