// mpm_math.hpp -- small device math for the MPM kernels: B-spline weights,
// colliders, the 3x3 symmetric Jacobi eigen-solver used as the FP32 SVD, and
// the FP64 one-sided Jacobi SVD used by the fallback path.
#pragma once

#include <hip/hip_runtime.h>

#include "mpm_device.hpp"

namespace snowmpm {

// Hardware reciprocal / square root / reciprocal square root (about 1 ulp,
// no IEEE division sequence) for FP32; exact operations for FP64.
__device__ __forceinline__ float rcp_(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float rsq_(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rcp_(double x) { return 1.0 / x; }
__device__ __forceinline__ double sqrt_(double x) { return sqrt(x); }
__device__ __forceinline__ double rsq_(double x) { return 1.0 / sqrt(x); }

// Element i of a particle array as a 32-bit byte offset from the array's
// wave-uniform base: the access takes the base from scalar registers and needs
// no 64-bit address arithmetic per lane.  Particle indices are below 2^30
// (the engine's capacity limit), so the offset does not wrap.
template <typename T>
__device__ __forceinline__ T& at(T* p, uint32_t i) {
    static_assert(sizeof(T) == 4, "4-byte elements");
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + (i << 2));
}
// Component e of a vector field (x, v, F, Fp, S) of particle i: components
// are p.stride elements apart, so a kernel needs only the first one's base.
__device__ __forceinline__ float& comp(float* const* f, const PBuf& p, int e, uint32_t i) {
    return at(f[0] + (size_t)e * p.stride, i);
}

// Base node of the particle's stencil, floor(x/dx) - 1 (grid.cpp:64-67).
__device__ __forceinline__ void base_node(float inv_dx, float x0, float x1, float x2, int b[3], float f[3]) {
    const float g0 = x0 * inv_dx, g1 = x1 * inv_dx, g2 = x2 * inv_dx;
    const float c0 = floorf(g0), c1 = floorf(g1), c2 = floorf(g2);
    f[0] = g0 - c0;
    f[1] = g1 - c1;
    f[2] = g2 - c2;
    b[0] = (int)c0 - 1;
    b[1] = (int)c1 - 1;
    b[2] = (int)c2 - 1;
}
__device__ __forceinline__ void base_node(const SceneDev& sc, float x0, float x1, float x2, int b[3], float f[3]) {
    base_node(sc.inv_dx, x0, x1, x2, b, f);
}

// Lame parameters of particle i: per-particle Young's modulus or the scene's.
__device__ __forceinline__ void lame(const SceneDev& sc, const PBuf& p, uint32_t i, float& mu0, float& la0) {
    if (sc.use_E) {
        const float E = at(p.E, i);
        mu0 = E * sc.mu_per_E;
        la0 = E * sc.la_per_E;
    } else {
        mu0 = sc.mu0;
        la0 = sc.lambda0;
    }
}

// Cubic B-spline weights of the 4 stencil nodes c-1, c, c+1, c+2 for a
// particle at fractional offset f in [0,1) inside cell c, and their
// derivatives with respect to the particle coordinate (grid units).  These
// are the pieces of N and N_dash (reference src/grid.cpp:24-58) that apply on
// each node's distance interval (1+f, f, 1-f, 2-f), written as polynomials
// in f.  sign(p - node) is +1, sign(f), -1, -1, and N_dash vanishes at f = 0
// for node c exactly as signum(0) = 0 makes it in the reference.
__device__ __forceinline__ void bspline(float f, float w[4], float dw[4]) {
    const float t = 1.0f - f;
    const float f2 = f * f, t2 = t * t;
    w[0] = (1.0f / 6.0f) * t2 * t;
    w[1] = 0.5f * f2 * f - f2 + (2.0f / 3.0f);
    w[2] = 0.5f * t2 * t - t2 + (2.0f / 3.0f);
    w[3] = (1.0f / 6.0f) * f2 * f;
    dw[0] = -0.5f * t2;
    dw[1] = 1.5f * f2 - 2.0f * f;
    dw[2] = -(1.5f * t2 - 2.0f * t);
    dw[3] = 0.5f * f2;
}

// Collider signed distance and normal (plane: reference src/collision.cpp:5-27;
// box and sphere: solid SDF bodies).
template <typename T>
__device__ __forceinline__ T collider_phi(const ColliderDev& c, T x0, T x1, T x2, T n[3]) {
    if (c.type == 0) {
        n[0] = c.normal[0];
        n[1] = c.normal[1];
        n[2] = c.normal[2];
        return n[0] * (x0 - (T)c.point[0]) + n[1] * (x1 - (T)c.point[1]) + n[2] * (x2 - (T)c.point[2]);
    }
    if (c.type == 1) {
        T d[3] = {x0 - (T)c.point[0], x1 - (T)c.point[1], x2 - (T)c.point[2]};
        T q[3];
        T qmax = -1e30f, outside = 0;
        int amax = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            q[a] = fabs(d[a]) - (T)c.half_extent[a];
            if (q[a] > qmax) {
                qmax = q[a];
                amax = a;
            }
            if (q[a] > 0) outside += q[a] * q[a];
        }
        if (qmax > 0) {
            T l = 0;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                n[a] = q[a] > 0 ? q[a] * (d[a] >= 0 ? (T)1 : (T)-1) : (T)0;
                l += n[a] * n[a];
            }
            l = rsq_(l);
#pragma unroll
            for (int a = 0; a < 3; ++a) n[a] *= l;
        } else {
#pragma unroll
            for (int a = 0; a < 3; ++a) n[a] = (a == amax) ? (d[a] >= 0 ? (T)1 : (T)-1) : (T)0;
        }
        return sqrt_(outside) + (qmax < 0 ? qmax : (T)0);
    }
    T d0 = x0 - (T)c.point[0], d1 = x1 - (T)c.point[1], d2 = x2 - (T)c.point[2];
    T l = sqrt_(d0 * d0 + d1 * d1 + d2 * d2);
    if (l > 0) {
        T il = rcp_(l);
        n[0] = d0 * il;
        n[1] = d1 * il;
        n[2] = d2 * il;
    } else {
        n[0] = 0;
        n[1] = 1;
        n[2] = 0;
    }
    return l - (T)c.radius;
}

// collision_helper (reference src/grid.cpp:206-231), bodies applied in list
// order (grid.cpp:238-241, 329-332).
template <typename T>
__device__ __forceinline__ void collide(const SceneDev& sc, T x0, T x1, T x2, T& v0, T& v1, T& v2) {
    for (int k = 0; k < sc.ncoll; ++k) {
        const ColliderDev& c = sc.coll[k];
        T n[3];
        T phi = collider_phi<T>(c, x0, x1, x2, n);
        if (phi <= 0) {
            T r0 = v0 - (T)c.velocity[0], r1 = v1 - (T)c.velocity[1], r2 = v2 - (T)c.velocity[2];
            T vn = r0 * n[0] + r1 * n[1] + r2 * n[2];
            if (vn < 0) {
                T t0 = r0 - vn * n[0], t1 = r1 - vn * n[1], t2 = r2 - vn * n[2];
                T tn = sqrt_(t0 * t0 + t1 * t1 + t2 * t2);
                T mu = (T)c.friction;
                if (tn <= -1 * mu * vn) {
                    r0 = r1 = r2 = 0;
                } else {
                    T s = mu * vn * rcp_(tn);
                    r0 = t0 + s * t0;
                    r1 = t1 + s * t1;
                    r2 = t2 + s * t2;
                }
                v0 = r0 + (T)c.velocity[0];
                v1 = r1 + (T)c.velocity[1];
                v2 = r2 + (T)c.velocity[2];
            }
        }
    }
}

// Optional position projection after advection (colliders flagged
// kProjectFlag, in list order): a particle inside the body (phi < 0) moves to
// x - phi n, its closest surface point.  Not in the reference, which leaves
// positions to advection (grid.cpp:335-338); velocities are not touched.
template <typename T>
__device__ __forceinline__ void project(const SceneDev& sc, T& x0, T& x1, T& x2) {
    for (int k = 0; k < sc.ncoll; ++k) {
        const ColliderDev& c = sc.coll[k];
        if (!(c.flags & kProjectFlag)) continue;
        T n[3];
        const T phi = collider_phi<T>(c, x0, x1, x2, n);
        if (phi < 0) {
            x0 -= phi * n[0];
            x1 -= phi * n[1];
            x2 -= phi * n[2];
        }
    }
}

template <typename T>
__device__ __forceinline__ T det3(const T A[9]) {
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}

// One cyclic-Jacobi rotation zeroing a[P][Q] of a symmetric 3x3 matrix and
// accumulating the rotation into the eigenvector matrix v (columns).
template <typename T, int P, int Q>
__device__ __forceinline__ void jacobi_rot(T (&a)[3][3], T (&v)[3][3]) {
    constexpr int R = 3 - P - Q;
    const T apq = a[P][Q];
    if (apq == (T)0) return;
    const T theta = (a[Q][Q] - a[P][P]) * (T)0.5 * rcp_(apq);
    const T at = fabs(theta);
    // t = 1/(|theta| + sqrt(theta^2+1)); for huge |theta| the square
    // overflows to inf and t becomes 0, i.e. no rotation, as it should.
    T t = rcp_(at + sqrt_(at * at + 1));
    if (theta < 0) t = -t;
    const T c = rsq_(t * t + 1);
    const T s = t * c;
    a[P][P] -= t * apq;
    a[Q][Q] += t * apq;
    a[P][Q] = a[Q][P] = 0;
    const T arp = a[R][P], arq = a[R][Q];
    a[R][P] = a[P][R] = c * arp - s * arq;
    a[R][Q] = a[Q][R] = s * arp + c * arq;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const T vp = v[r][P], vq = v[r][Q];
        v[r][P] = c * vp - s * vq;
        v[r][Q] = s * vp + c * vq;
    }
}

// Eigen-decomposition of the symmetric C = F^T F: C = V diag(lam) V^T.
template <typename T, int SWEEPS>
__device__ __forceinline__ void sym_eig3(T (&a)[3][3], T (&v)[3][3], T lam[3]) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[r][c] = (r == c) ? (T)1 : (T)0;
#pragma unroll
    for (int s = 0; s < SWEEPS; ++s) {
        jacobi_rot<T, 0, 1>(a, v);
        jacobi_rot<T, 0, 2>(a, v);
        jacobi_rot<T, 1, 2>(a, v);
    }
    lam[0] = a[0][0];
    lam[1] = a[1][1];
    lam[2] = a[2][2];
}

// FP64 one-sided Jacobi SVD with U completion for (near) singular input:
// A = U diag(S) V^T, S >= 0, U and V orthogonal (Eigen's unsigned form).
__device__ inline void svd3_fp64(const double A[9], double U[9], double S[3], double V[9]) {
    double W[3][3], Vc[3][3];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            W[c][r] = A[r * 3 + c];
            Vc[c][r] = (r == c) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 30; ++sweep) {
        bool rotated = false;
        for (int pq = 0; pq < 3; ++pq) {
            const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
            double al = 0, be = 0, ga = 0;
            for (int r = 0; r < 3; ++r) {
                al += W[p][r] * W[p][r];
                be += W[q][r] * W[q][r];
                ga += W[p][r] * W[q][r];
            }
            if (ga == 0.0 || fabs(ga) <= 1e-15 * sqrt(al * be)) continue;
            rotated = true;
            const double zeta = (be - al) / (2.0 * ga);
            const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
            for (int r = 0; r < 3; ++r) {
                const double wp = W[p][r], wq = W[q][r];
                W[p][r] = c * wp - s * wq;
                W[q][r] = s * wp + c * wq;
                const double vp = Vc[p][r], vq = Vc[q][r];
                Vc[p][r] = c * vp - s * vq;
                Vc[q][r] = s * vp + c * vq;
            }
        }
        if (!rotated) break;
    }
    double sig[3];
    for (int c = 0; c < 3; ++c) sig[c] = sqrt(W[c][0] * W[c][0] + W[c][1] * W[c][1] + W[c][2] * W[c][2]);
    // order descending (3-element sort network)
    int o[3] = {0, 1, 2};
    if (sig[o[0]] < sig[o[1]]) { int t = o[0]; o[0] = o[1]; o[1] = t; }
    if (sig[o[1]] < sig[o[2]]) { int t = o[1]; o[1] = o[2]; o[2] = t; }
    if (sig[o[0]] < sig[o[1]]) { int t = o[0]; o[0] = o[1]; o[1] = t; }
    double u[3][3];
    const double smax = sig[o[0]];
    for (int k = 0; k < 3; ++k) {
        const int c = o[k];
        S[k] = sig[c];
        for (int r = 0; r < 3; ++r) V[r * 3 + k] = Vc[c][r];
        if (sig[c] > 1e-12 * smax && sig[c] > 0.0) {
            for (int r = 0; r < 3; ++r) u[k][r] = W[c][r] / sig[c];
        } else if (k == 0) {
            u[0][0] = 1; u[0][1] = 0; u[0][2] = 0;
        } else if (k == 1) {
            double e[3] = {fabs(u[0][0]) < 0.9 ? 1.0 : 0.0, fabs(u[0][0]) < 0.9 ? 0.0 : 1.0, 0.0};
            double d = e[0] * u[0][0] + e[1] * u[0][1] + e[2] * u[0][2];
            double t[3] = {e[0] - d * u[0][0], e[1] - d * u[0][1], e[2] - d * u[0][2]};
            double l = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
            for (int r = 0; r < 3; ++r) u[1][r] = t[r] / l;
        } else {
            u[2][0] = u[0][1] * u[1][2] - u[0][2] * u[1][1];
            u[2][1] = u[0][2] * u[1][0] - u[0][0] * u[1][2];
            u[2][2] = u[0][0] * u[1][1] - u[0][1] * u[1][0];
        }
    }
    for (int k = 0; k < 3; ++k)
        for (int r = 0; r < 3; ++r) U[r * 3 + k] = u[k][r];
}

}  // namespace snowmpm
