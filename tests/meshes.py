"""Closed triangle meshes for the seeding tests, and an independent inside test.

Every mesh is outward-oriented (counter-clockwise seen from outside).
``winding`` is the generalised winding number of a closed mesh at points
(the solid-angle sum of van Oosterom and Strackee, 1983, in float64): 1
inside, 0 outside, so it classifies points without any ray.  ``surface_dist``
is the distance from points to the mesh (closest point on each triangle,
Ericson, Real-Time Collision Detection 5.1.5), used to set aside points that
lie within rounding of the surface.
"""
import numpy as np


def box(lo, hi, flip=False):
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    v = np.array([[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]], [lo[0], hi[1], lo[2]],
                  [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]], [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]]])
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [2, 3, 7], [2, 7, 6], [1, 2, 6], [1, 6, 5], [0, 4, 7], [0, 7, 3]], dtype=np.int32)
    if flip:
        f = f[:, ::-1].copy()
    return v, f


def merge(*meshes):
    vs, fs, off = [], [], 0
    for v, f in meshes:
        vs.append(v)
        fs.append(f + off)
        off += v.shape[0]
    return np.concatenate(vs), np.concatenate(fs).astype(np.int32)


def cavity_box(lo, hi, ilo, ihi):
    """A solid box with a closed box-shaped cavity (inner faces turned inward)."""
    return merge(box(lo, hi), box(ilo, ihi, flip=True))


def octahedron(c, r):
    c = np.asarray(c, float)
    v = c + r * np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], float)
    f = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]],
                 dtype=np.int32)
    return v, f


def icosphere(c, r, subdiv):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    v = [list(np.asarray(p) / np.linalg.norm(p)) for p in v]
    for _ in range(subdiv):
        mid, nf = {}, []

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = (np.asarray(v[a]) + np.asarray(v[b])) / 2.0
                v.append(list(p / np.linalg.norm(p)))
                mid[key] = len(v) - 1
            return mid[key]

        for a, b, cc in f:
            ab, bc, ca = m(a, b), m(b, cc), m(cc, a)
            nf += [[a, ab, ca], [b, bc, ab], [cc, ca, bc], [ab, bc, ca]]
        f = nf
    return np.asarray(c, float) + r * np.asarray(v), np.asarray(f, dtype=np.int32)


def torus(c, R, r, nu, nv):
    """Torus around the z axis (quads split into two triangles)."""
    u = 2 * np.pi * np.arange(nu) / nu
    w = 2 * np.pi * np.arange(nv) / nv
    U, W = np.meshgrid(u, w, indexing="ij")
    x = (R + r * np.cos(W)) * np.cos(U)
    y = (R + r * np.cos(W)) * np.sin(U)
    z = r * np.sin(W)
    v = np.stack([x, y, z], -1).reshape(-1, 3) + np.asarray(c, float)
    f = []
    for i in range(nu):
        for j in range(nv):
            a, b = i * nv + j, ((i + 1) % nu) * nv + j
            cc, d = ((i + 1) % nu) * nv + (j + 1) % nv, i * nv + (j + 1) % nv
            f += [[a, b, cc], [a, cc, d]]
    return v, np.asarray(f, dtype=np.int32)


def winding(verts, tris, pts):
    """Generalised winding number of the mesh at each point (float64)."""
    pts = np.asarray(pts, float).reshape(-1, 3)
    out = np.zeros(pts.shape[0])
    T = np.asarray(verts, float)[np.asarray(tris)]
    for a, b, c in T:
        A, B, C = a - pts, b - pts, c - pts
        la, lb, lc = (np.linalg.norm(X, axis=1) for X in (A, B, C))
        num = np.einsum("ij,ij->i", A, np.cross(B, C))
        den = (la * lb * lc + np.einsum("ij,ij->i", A, B) * lc + np.einsum("ij,ij->i", B, C) * la
               + np.einsum("ij,ij->i", C, A) * lb)
        out += 2.0 * np.arctan2(num, den)
    return out / (4.0 * np.pi)


def _closest_on_triangle(p, a, b, c):
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b)
    den = 1.0 / (va + vb + vc)
    return a + ab * (vb * den) + ac * (vc * den)


def surface_dist(verts, tris, p):
    """Distance from one point to the mesh."""
    p = np.asarray(p, float)
    T = np.asarray(verts, float)[np.asarray(tris)]
    return min(np.linalg.norm(p - _closest_on_triangle(p, a, b, c)) for a, b, c in T)


def lattice(verts, spacing):
    """The add-on's lattice (seed_export.py:106-131) restated in plain Python:
    float32 bounds and spacing, double accumulation, float32 points."""
    v32 = np.asarray(verts, float).astype(np.float32)
    lo, hi = v32.min(axis=0), v32.max(axis=0)
    s = float(np.float32(spacing))
    axes = []
    for e in range(3):
        a, x = [], float(lo[e]) + s * 0.5
        while x < float(hi[e]):
            a.append(float(np.float32(x)))
            x += s
        axes.append(np.asarray(a))
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    return axes, np.stack([X, Y, Z], -1).reshape(-1, 3)
