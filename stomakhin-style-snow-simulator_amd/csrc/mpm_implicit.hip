// mpm_implicit.hip -- the semi-implicit velocity solve on the GPU.
//
// The reference declares this solve and never defines it (grid.hpp:77
// constitutive_differential, 81 index_nodes, 89 deformation_differential, 90
// force_differential, 91 apply_A, 92 Conjugate_residual).  It follows
// Stomakhin et al. 2013 (section 4): after the grid update and collisions,
//   (I + beta dt^2 M^-1 H) v = v*
// on the nodes with mass of scenes with beta > 0 that no collider changed
// (the unknowns), v* the explicit grid velocity, H q = -df(q)/dt the elastic
// Hessian applied to a grid velocity field q:
//   dF_p = dt (sum_i q_i grad w_ip^T) F_Ep          k_imp_dF
//   dP_p = dP/dF (F_Ep) : dF_p  (corotated, hardened) k_imp_dF
//   df_i = - sum_p V_p0 dP_p F_Ep^T grad w_ip        P2G's force channels (k_p2g<.., true>)
//   A q = q - beta dt M^-1 df(q)                      k_imp_apply
// M A is symmetric, so the conjugate-residual iteration runs in the mass inner
// product <a, b> = sum m_i a_i . b_i, per scene (every scene has its own
// scalars, so a scene's solve does not depend on the others).  Dot products
// are summed per block, then per scene over the blocks in key order, each in
// a fixed order: the solve is as deterministic as the explicit step.  Vectors
// are FP32 with the grid velocity's layout; dot products and the CR scalars
// are FP64.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "mpm_device.hpp"
#include "mpm_engine.hpp"
#include "mpm_kernels.hpp"
#include "mpm_math.hpp"

namespace snowmpm {

// ---------------------------------------------------------------------------
// Block order: all blocks of the pass sorted by key, and each scene's range.

__global__ __launch_bounds__(256) void k_imp_keys(StepArgs a, uint32_t n, uint64_t* kin, uint32_t* vin) {
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < n; b += gridDim.x * blockDim.x) {
        const uint32_t slot = a.g.all_list[b];
        kin[b] = a.g.keys[slot];
        vin[b] = slot;
    }
}

__global__ __launch_bounds__(256) void k_imp_clear_seg(StepArgs a) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < 2u * kMaxScenes; k += gridDim.x * blockDim.x)
        a.imp.seg[k] = 0u;
}

__global__ __launch_bounds__(256) void k_imp_segments(StepArgs a, uint32_t n) {
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < n; b += gridDim.x * blockDim.x) {
        const uint32_t sid = (uint32_t)(a.imp.skey[b] >> 51);
        if (b == 0 || (uint32_t)(a.imp.skey[b - 1] >> 51) != sid) a.imp.seg[2 * sid] = b;
        if (b + 1 == n || (uint32_t)(a.imp.skey[b + 1] >> 51) != sid) a.imp.seg[2 * sid + 1] = b + 1;
    }
}

// ---------------------------------------------------------------------------
// Per-block vector kernels: one wave per block (64 nodes, one per lane), four
// blocks per workgroup, blocks in key order.  dot() is the block's share of a
// mass-weighted dot product, summed over the wave in a fixed order.

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ float* vec_at(float* v, uint32_t slot, int c, int lane) {
    return v + ((size_t)slot * 3 + c) * 64 + lane;
}

// X = v* on the unknowns (0 elsewhere); R, P, AR, AP = 0.
__global__ __launch_bounds__(256) void k_imp_init(StepArgs a, uint32_t n) {
    const int lane = (int)(threadIdx.x & 63);
    for (uint32_t b = blockIdx.x * 4 + (threadIdx.x >> 6); b < n; b += gridDim.x * 4) {
        const uint32_t slot = a.imp.ord[b];
        const bool unk = a.imp.gm[(size_t)slot * 64 + lane] > 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float vs = a.g.vu[((size_t)slot * kVelCh + c) * 64 + lane];
            *vec_at(a.imp.vec[0], slot, c, lane) = unk ? vs : 0.0f;
#pragma unroll
            for (int k = 1; k < kCRVecs; ++k) *vec_at(a.imp.vec[k], slot, c, lane) = 0.0f;
        }
    }
}

// The force differential summed from the home blocks' partial sums (channels
// 4..6, layout at k_grid), then Y = Q - beta dt df / m on the unknowns.
//   mode 0: Q = X:  R = v* - Y;            dots <R,R>, <v*,v*>        (slots 0, 1)
//   mode 1: Q = R:  AR = AP = Y, P = R;    dots <R,AR>, <AR,AR>       (slots 0, 1)
//   mode 2: Q = R:  AR = Y;                dots <R,AR>, <AR,AR>, <AR,AP> (slots 1, 2, 3)
// (per block 4 partial slots in dp; slot 0 of mode 2 is k_imp_vec's <R,R>)
__global__ __launch_bounds__(256) void k_imp_apply(StepArgs a, uint32_t n, int mode) {
    const int lane = (int)(threadIdx.x & 63);
    const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
    float* X = a.imp.vec[0];
    float* R = a.imp.vec[1];
    float* P = a.imp.vec[2];
    float* AR = a.imp.vec[3];
    float* AP = a.imp.vec[4];
    for (uint32_t b = blockIdx.x * 4 + (threadIdx.x >> 6); b < n; b += gridDim.x * 4) {
        const uint32_t slot = a.imp.ord[b];
        const uint32_t sid = (uint32_t)(a.imp.skey[b] >> 51);
        const SceneDev& sc = a.scenes[sid];
        const uint32_t revl = lane < 8 ? a.g.rev[(size_t)slot * 8 + lane] : kInvalid;
        float df[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int nx = ((t & 1) ? 3 : 4), ny = (((t >> 1) & 1) ? 3 : 4), nz = ((t >> 2) ? 3 : 4), nxy = nx * ny;
            // xy piece t & 3 of a plane starts at the padded offset of the
            // pieces before it (7 nx ny floats each, rounded up to 32)
            int xyb = 0;
            for (int u = 0; u < (t & 3); ++u) xyb += (kGridCh * ((u & 1) ? 3 : 4) * (((u >> 1) & 1) ? 3 : 4) + 31) / 32 * 32;
            const uint32_t hb = (uint32_t)__builtin_amdgcn_readlane((int)revl, t);
            if (hb != kInvalid && hb < a.g.max_home && lx < nx && ly < ny && lz < nz) {
                const float* src = a.g.part + (size_t)hb * kPartFloats + (4 * (t >> 2) + lz) * kPlaneFloats + xyb +
                                   ly * nx + lx;
#pragma unroll
                for (int r = 0; r < 3; ++r) df[r] += src[(4 + r) * nxy];
            }
        }
        const float m = a.imp.gm[(size_t)slot * 64 + lane];
        const float* Q = mode == 0 ? X : R;
        const float s = m > 0.0f ? sc.beta * sc.dt * rcp_(m) : 0.0f;
        double d0 = 0.0, d1 = 0.0, d2 = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float q = *vec_at(const_cast<float*>(Q), slot, c, lane);
            const float y = m > 0.0f ? q - s * df[c] : 0.0f;
            if (mode == 0) {
                const float vs = m > 0.0f ? a.g.vu[((size_t)slot * kVelCh + c) * 64 + lane] : 0.0f;
                const float r = vs - y;
                *vec_at(R, slot, c, lane) = r;
                d0 += (double)m * (double)r * (double)r;
                d1 += (double)m * (double)vs * (double)vs;
            } else {
                *vec_at(AR, slot, c, lane) = y;
                d0 += (double)m * (double)q * (double)y;
                d1 += (double)m * (double)y * (double)y;
                if (mode == 1) {
                    *vec_at(AP, slot, c, lane) = y;
                    *vec_at(P, slot, c, lane) = q;
                } else {
                    d2 += (double)m * (double)y * (double)*vec_at(AP, slot, c, lane);
                }
            }
        }
        d0 = wave_sum(d0);
        d1 = wave_sum(d1);
        if (mode == 2) d2 = wave_sum(d2);
        if (lane == 0) {
            double* o = a.imp.dp + 4 * (size_t)b;
            if (mode == 2) {
                o[1] = d0;
                o[2] = d1;
                o[3] = d2;
            } else {
                o[0] = d0;
                o[1] = d1;
            }
        }
    }
}

// CR vector updates with the scene's scalars, mode 0: the direction update
// of the previous iteration, P = R + beta P and AP = AR + beta AP, then
// X += alpha P and R -= alpha AP; dot <R,R> (slot 0).  Mode 2: the solution
// into the grid: v = X and u = u* + (X - v*) on the unknowns.
__global__ __launch_bounds__(256) void k_imp_vec(StepArgs a, uint32_t n, int mode) {
    const int lane = (int)(threadIdx.x & 63);
    float* X = a.imp.vec[0];
    float* R = a.imp.vec[1];
    float* P = a.imp.vec[2];
    float* AR = a.imp.vec[3];
    float* AP = a.imp.vec[4];
    for (uint32_t b = blockIdx.x * 4 + (threadIdx.x >> 6); b < n; b += gridDim.x * 4) {
        const uint32_t slot = a.imp.ord[b];
        const uint32_t sid = (uint32_t)(a.imp.skey[b] >> 51);
        const float m = a.imp.gm[(size_t)slot * 64 + lane];
        if (mode == 2) {
            if (m > 0.0f)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    float* vu = a.g.vu + ((size_t)slot * kVelCh + c) * 64 + lane;
                    const float x = *vec_at(X, slot, c, lane), vs = vu[0];
                    vu[0] = x;
                    vu[3 * 64] = vu[3 * 64] + (x - vs);
                }
            continue;
        }
        const float alpha = (float)a.imp.cr[sid].alpha, beta = (float)a.imp.cr[sid].beta;
        double d0 = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float* p = vec_at(P, slot, c, lane);
            float* ap = vec_at(AP, slot, c, lane);
            float* x = vec_at(X, slot, c, lane);
            float* r = vec_at(R, slot, c, lane);
            const float rc = *r;
            const float pn = rc + beta * *p;
            const float apn = *vec_at(AR, slot, c, lane) + beta * *ap;
            *p = pn;
            *ap = apn;
            *x = *x + alpha * pn;
            const float rn = rc - alpha * apn;
            *r = rn;
            d0 += (double)m * (double)rn * (double)rn;
        }
        d0 = wave_sum(d0);
        if (lane == 0) a.imp.dp[4 * (size_t)b] = d0;
    }
}

// Per scene (one wave each): the block partials summed over the scene's
// blocks in key order (lane-strided, then a fixed butterfly), and the CR
// scalars updated:
//   kInit: rr, bb from mode-0 apply; active if rr > tol^2 bb
//   kRho0: rho = <R,AR>, den = <AR,AR>, alpha = rho / den, beta = 0
//   kIter: rr = <R,R> after k_imp_vec's update (stop at the tolerance or the
//          iteration cap); then rho' = <R,AR>, beta = rho' / rho and, for the
//          next direction AP' = AR + beta AP,
//          den = <AP',AP'> = <AR,AR> + 2 beta <AR,AP> + beta^2 <AP,AP>,
//          alpha = rho' / den (both 0 once the scene stops)
// Scenes still iterating are counted into nactive[slot].
enum { kInit = 0, kRho0 = 1, kIter = 2 };
__global__ __launch_bounds__(256) void k_imp_reduce(StepArgs a, uint32_t nscenes, int mode, int slot) {
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t sid = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (sid >= nscenes) return;
    const uint32_t b0 = a.imp.seg[2 * sid], b1 = a.imp.seg[2 * sid + 1];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (uint32_t b = b0 + lane; b < b1; b += 64) {
        const double* d = a.imp.dp + 4 * (size_t)b;
        s0 += d[0];
        s1 += d[1];
        if (mode == kIter) {
            s2 += d[2];
            s3 += d[3];
        }
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    if (mode == kIter) {
        s2 = wave_sum(s2);
        s3 = wave_sum(s3);
    }
    if (lane != 0) return;
    CRScene& c = a.imp.cr[sid];
    const SceneDev& sc = a.scenes[sid];
    double* hist = a.imp.hist + (size_t)sid * kCRHist;
    if (mode == kInit) {
        c.rr = s0;
        c.bb = s1;
        c.iters = 0;
        c.alpha = c.beta = 0.0;
        hist[0] = sqrt(s0);
        c.active = (b1 > b0 && sc.beta > 0.0f && s1 > 0.0 && s0 > (double)sc.cr_tol2 * s1 && sc.cr_maxit > 0) ? 1 : 0;
    } else if (mode == kRho0) {
        c.rho = s0;
        c.den = s1;
        c.beta = 0.0;
        c.alpha = (c.active && c.den > 0.0 && c.rho != 0.0) ? c.rho / c.den : 0.0;
        if (c.alpha == 0.0) c.active = 0;
    } else {  // kIter
        if (c.active) {
            c.rr = s0;
            c.iters += 1;
            hist[c.iters] = sqrt(s0);
            if (s0 <= (double)sc.cr_tol2 * c.bb || c.iters >= sc.cr_maxit) c.active = 0;
        }
        if (c.active && c.rho != 0.0) {
            const double beta = s1 / c.rho;
            const double den = s2 + 2.0 * beta * s3 + beta * beta * c.den;
            c.beta = beta;
            c.rho = s1;
            c.den = den;
            c.alpha = (den > 0.0 && s1 != 0.0) ? s1 / den : 0.0;
            if (c.alpha == 0.0) c.active = 0;
        }
        if (!c.active) c.alpha = c.beta = 0.0;
    }
    if (c.active) atomicAdd(a.imp.nactive + slot, 1u);
}

// ---------------------------------------------------------------------------
// The Hessian product's particle half: per particle the deformation
// differential from the staged velocity field Q of its 8 blocks, the
// constitutive differential of the corotated model at F_E, and
// G = -V0 dP F_E^T for P2G's force channels.  A workgroup of 128 threads per
// home block (the grid covers the home blocks).

// dR = U W V^T, W_ij = (M_ij - M_ji) / (s_i + s_j), M = U^T dF V: the
// polar factor's differential (R^T dR skew, R^T dR S + S R^T dR = R^T dF - dF^T R).
template <typename T>
__device__ __forceinline__ void corotated_dP(const T F[9], const T U[9], const T s[3], const T V[9], const T dF[9],
                                             T mu, T la, T dP[9]) {
    T Mh[9], tmp[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            tmp[i * 3 + j] = dF[i * 3 + 0] * V[0 * 3 + j] + dF[i * 3 + 1] * V[1 * 3 + j] + dF[i * 3 + 2] * V[2 * 3 + j];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Mh[i * 3 + j] = U[0 * 3 + i] * tmp[0 * 3 + j] + U[1 * 3 + i] * tmp[1 * 3 + j] + U[2 * 3 + i] * tmp[2 * 3 + j];
    T W[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) W[i * 3 + j] = i == j ? (T)0 : (Mh[i * 3 + j] - Mh[j * 3 + i]) / (s[i] + s[j]);
    // dR = U (W V^T)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            tmp[i * 3 + j] = W[i * 3 + 0] * V[j * 3 + 0] + W[i * 3 + 1] * V[j * 3 + 1] + W[i * 3 + 2] * V[j * 3 + 2];
    T C[9], dC[9], dJ = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            C[i * 3 + j] = F[i1 * 3 + j1] * F[i2 * 3 + j2] - F[i1 * 3 + j2] * F[i2 * 3 + j1];
            dC[i * 3 + j] = dF[i1 * 3 + j1] * F[i2 * 3 + j2] + F[i1 * 3 + j1] * dF[i2 * 3 + j2] -
                            dF[i1 * 3 + j2] * F[i2 * 3 + j1] - F[i1 * 3 + j2] * dF[i2 * 3 + j1];
            dJ += C[i * 3 + j] * dF[i * 3 + j];
        }
    const T J = det3<T>(F);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const T dR = U[i * 3 + 0] * tmp[0 * 3 + j] + U[i * 3 + 1] * tmp[1 * 3 + j] + U[i * 3 + 2] * tmp[2 * 3 + j];
            dP[i * 3 + j] = (T)2 * mu * (dF[i * 3 + j] - dR) + la * (dJ * C[i * 3 + j] + (J - (T)1) * dC[i * 3 + j]);
        }
}

constexpr int kImpThreads = 128;

// With few home blocks the particles of a block are split over `splits`
// workgroups (each stages the block's nodes and takes every splits-th group
// of kImpThreads particles), so a small scene is not bound by one
// workgroup's passes over its block.
__global__ __launch_bounds__(kImpThreads) void k_imp_dF(StepArgs a, const float* Q, uint32_t splits) {
    __shared__ float qs[3][512];  // [channel][Z][Y][X] of the 8x8x8 nodes of the block's 8 blocks
    const uint32_t nh = a.cnt->n_home;
    for (uint32_t wg = blockIdx.x; wg < nh * splits; wg += gridDim.x) {
        const uint32_t blk = wg / splits, part = wg % splits;
        const uint32_t slot = a.g.home_list[blk];
        uint32_t sid;
        int bx, by, bz;
        unpack_key(a.g.keys[slot], sid, bx, by, bz);
        const SceneDev& sc = a.scenes[sid];
        __syncthreads();  // the previous block's reads are done
        for (int e = threadIdx.x; e < 3 * 512; e += kImpThreads) {
            const int tg = e / 192, rem = e % 192, c = rem >> 6, nn = rem & 63;
            const uint32_t ts = a.g.nb[(size_t)slot * 8 + tg];
            const int X = (tg & 1) * 4 + (nn & 3), Y = ((tg >> 1) & 1) * 4 + ((nn >> 2) & 3), Z = (tg >> 2) * 4 + (nn >> 4);
            qs[c][Z * 64 + Y * 8 + X] = Q[((size_t)ts * 3 + c) * 64 + nn];
        }
        __syncthreads();
        const uint32_t off = a.g.blkoff[slot], cnt = a.g.blkcnt[slot];
        const bool implicit = sc.beta > 0.0f;
        for (uint32_t j = off + part * kImpThreads + threadIdx.x; j < off + cnt; j += splits * kImpThreads) {
            const uint32_t i = at(a.bin.sidx, j);
            float G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            if (implicit) {
                const float x0 = comp(a.cur.x, a.cur, 0, i), x1 = comp(a.cur.x, a.cur, 1, i),
                            x2 = comp(a.cur.x, a.cur, 2, i);
                int b[3];
                float f[3];
                base_node(sc, x0, x1, x2, b, f);
                const int lx = min(max(b[0] - 4 * bx, 0), 3), ly = min(max(b[1] - 4 * by, 0), 3),
                          lz = min(max(b[2] - 4 * bz, 0), 3);
                float wx[4], dwx[4], wy[4], dwy[4], wz[4], dwz[4];
                bspline(f[0], wx, dwx);
                bspline(f[1], wy, dwy);
                bspline(f[2], wz, dwz);
                // L = sum_i q_i grad w_ip^T, separable over x, then y and z
                float L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int dd = 0; dd < 4; ++dd)
#pragma unroll
                    for (int bb = 0; bb < 4; ++bb) {
                        const int rowb = (lz + dd) * 64 + (ly + bb) * 8 + lx;
                        float S[3] = {0, 0, 0}, D[3] = {0, 0, 0};
#pragma unroll
                        for (int aa = 0; aa < 4; ++aa)
#pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                const float q = qs[c][rowb + aa];
                                S[c] += wx[aa] * q;
                                D[c] += dwx[aa] * q;
                            }
                        const float yz = wy[bb] * wz[dd], dyz = dwy[bb] * wz[dd], ydz = wy[bb] * dwz[dd];
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            L[c * 3 + 0] += yz * D[c];
                            L[c * 3 + 1] += dyz * S[c];
                            L[c * 3 + 2] += ydz * S[c];
                        }
                    }
                float F[9], dF[9];
#pragma unroll
                for (int e = 0; e < 9; ++e) F[e] = comp(a.cur.F, a.cur, e, i);
                const float sdt = sc.dt * sc.inv_dx;  // grad w in grid units: / dx
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        dF[r * 3 + c] = sdt * (L[r * 3 + 0] * F[0 * 3 + c] + L[r * 3 + 1] * F[1 * 3 + c] +
                                               L[r * 3 + 2] * F[2 * 3 + c]);
                float mu0, la0;
                lame(sc, a.cur, i, mu0, la0);
                const float h = expf(sc.hardening * (1.0f - at(a.cur.Jp, i)));
                const float mu = mu0 * h, la = la0 * h;
                // SVD of F_E through the eigenproblem of F^T F (F_E is well
                // conditioned after the plasticity clamp); FP64 otherwise
                float A3[3][3], V3[3][3], lam[3];
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = r; c < 3; ++c) {
                        const float s = F[0 * 3 + r] * F[0 * 3 + c] + F[1 * 3 + r] * F[1 * 3 + c] + F[2 * 3 + r] * F[2 * 3 + c];
                        A3[r][c] = s;
                        A3[c][r] = s;
                    }
                sym_eig3<float, 4>(A3, V3, lam);
                float s[3], U[9], V[9];
#pragma unroll
                for (int k = 0; k < 3; ++k) s[k] = sqrt_(fmaxf(lam[k], 0.0f));
                const float smin = fminf(s[0], fminf(s[1], s[2])), smax = fmaxf(s[0], fmaxf(s[1], s[2]));
                float dP[9];
                if (det3<float>(F) > 0.0f && smin > 0.0f && smax <= a.kappa * smin) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float is = rcp_(s[k]);
#pragma unroll
                        for (int r = 0; r < 3; ++r) {
                            V[r * 3 + k] = V3[r][k];
                            U[r * 3 + k] = (F[r * 3 + 0] * V3[0][k] + F[r * 3 + 1] * V3[1][k] + F[r * 3 + 2] * V3[2][k]) * is;
                        }
                    }
                    corotated_dP<float>(F, U, s, V, dF, mu, la, dP);
                } else {
                    double Fd[9], Ud[9], Sd[3], Vd[9], dFd[9], dPd[9];
                    for (int e = 0; e < 9; ++e) {
                        Fd[e] = F[e];
                        dFd[e] = dF[e];
                    }
                    svd3_fp64(Fd, Ud, Sd, Vd);
                    corotated_dP<double>(Fd, Ud, Sd, Vd, dFd, (double)mu, (double)la, dPd);
                    for (int e = 0; e < 9; ++e) dP[e] = (float)dPd[e];
                }
                const float nv = -at(a.cur.V0, i);
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        G[r * 3 + c] = nv * (dP[r * 3 + 0] * F[c * 3 + 0] + dP[r * 3 + 1] * F[c * 3 + 1] +
                                             dP[r * 3 + 2] * F[c * 3 + 2]);
            }
#pragma unroll
            for (int e = 0; e < 9; ++e) at(a.imp.G + (size_t)e * a.cur.stride, i) = G[e];
        }
    }
}

// ---------------------------------------------------------------------------
// Host side.

namespace {
inline unsigned blocks_for(uint32_t n, unsigned per, unsigned cap) {
    const unsigned g = (unsigned)((n + per - 1) / per);
    return g == 0 ? 1u : (g < cap ? g : cap);
}
}  // namespace

int Engine::implicit_alloc() {
    if (imp_.gm) return MPM_OK;
    const size_t N = slots_, H = H_;  // G is indexed like the particle buffers
    int rc;
    if ((rc = dalloc((void**)&imp_.gm, H * 64 * 4))) return rc;
    for (int k = 0; k < kCRVecs; ++k)
        if ((rc = dalloc((void**)&imp_.vec[k], H * 3 * 64 * 4))) return rc;
    if ((rc = dalloc((void**)&imp_.G, N * 9 * 4))) return rc;
    if ((rc = dalloc((void**)&imp_.ord, H * 4))) return rc;
    if ((rc = dalloc((void**)&imp_.skey, H * 8))) return rc;
    if ((rc = dalloc((void**)&imp_.dp, H * 4 * 8))) return rc;
    if ((rc = dalloc((void**)&imp_.seg, 2 * kMaxScenes * 4))) return rc;
    if ((rc = dalloc((void**)&imp_.cr, kMaxScenes * sizeof(CRScene)))) return rc;
    if ((rc = dalloc((void**)&imp_.hist, (size_t)kMaxScenes * kCRHist * 8))) return rc;
    if ((rc = dalloc((void**)&imp_.nactive, (kCRHist + 1) * 4))) return rc;
    if ((rc = dalloc((void**)&imp_kin_, H * 8))) return rc;
    if ((rc = dalloc((void**)&imp_vin_, H * 4))) return rc;
    imp_sort_bytes_ = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, imp_sort_bytes_, imp_kin_, imp_.skey, imp_vin_, imp_.ord,
                                               (int)H, 0, 64, stream_));
    if ((rc = dalloc(&imp_sort_tmp_, imp_sort_bytes_))) return rc;
    HIP_TRY(hipMemsetAsync(imp_.hist, 0, (size_t)kMaxScenes * kCRHist * 8, stream_));
    HIP_TRY(hipMemsetAsync(imp_.cr, 0, kMaxScenes * sizeof(CRScene), stream_));
    return MPM_OK;
}

// One CR solve for the step whose grid launch_step_head just built.
int Engine::implicit_solve(const StepArgs& a) {
    Counters c;
    HIP_TRY(hipMemcpyAsync(&c, cnt_, sizeof c, hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    const uint32_t n = c.n_home + c.n_extra;
    if (n == 0) return MPM_OK;
    if (n > H_) {
        set_error("implicit solve: more blocks than hash slots");
        return MPM_ERR_STATE;
    }
    const hipStream_t s = stream_;
    k_imp_keys<<<blocks_for(n, 256, 4096), 256, 0, s>>>(a, n, imp_kin_, imp_vin_);
    size_t bytes = imp_sort_bytes_;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(imp_sort_tmp_, bytes, imp_kin_, imp_.skey, imp_vin_, imp_.ord, (int)n,
                                               0, 64, s));
    k_imp_clear_seg<<<64, 256, 0, s>>>(a);
    k_imp_segments<<<blocks_for(n, 256, 4096), 256, 0, s>>>(a, n);
    const unsigned gb = blocks_for(n, 4, (unsigned)cus_ * 16);
    const uint32_t nsc = (uint32_t)scenes_.size();
    const unsigned gs = blocks_for(nsc, 4, 65535);
    const unsigned gw = (unsigned)std::max(launch_.grid_w, launch_.grid_h);
    int maxit = 0;
    for (size_t k = 0; k < scenes_.size(); ++k)
        if (scenes_[k].used && h_scenes_[k].beta > 0.0f) maxit = std::max(maxit, h_scenes_[k].cr_maxit);
    // nactive[k]: scenes still iterating after the k-th reduction of this
    // solve (cleared once here, so the loop has no memory-set nodes)
    HIP_TRY(hipMemsetAsync(imp_.nactive, 0, (kCRHist + 1) * 4, s));
    auto reduce = [&](int mode, int slot) { k_imp_reduce<<<gs, 256, 0, s>>>(a, nsc, mode, slot); };
    // k_imp_dF: split blocks over workgroups while there are fewer than
    // about 4 workgroups per CU (at most 4 ways)
    const uint32_t splits = std::max(1u, std::min(4u, (uint32_t)(cus_ * 4) / std::max(1u, c.n_home)));
    auto apply = [&](const float* q, int mode) {
        k_imp_dF<<<gw, kImpThreads, 0, s>>>(a, q, splits);
        launch_p2g_force(a, launch_, s);
        k_imp_apply<<<gb, 256, 0, s>>>(a, n, mode);
    };
    auto any_active = [&](int slot, bool* any) -> int {
        uint32_t act = 0;
        HIP_TRY(hipMemcpyAsync(&act, imp_.nactive + slot, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        *any = act != 0;
        return MPM_OK;
    };
    int rc;
    k_imp_init<<<gb, 256, 0, s>>>(a, n);
    apply(imp_.vec[0], 0);
    reduce(kInit, 0);
    bool any = false;
    if ((rc = any_active(0, &any))) return rc;
    if (any) {
        apply(imp_.vec[1], 1);
        reduce(kRho0, 1);
        // Per iteration: the direction and solution updates with their <R,R>,
        // the operator on R with its three dots, and one per-scene reduction;
        // the host looks at the active count every 4 iterations.
        for (int it = 0; it < maxit; ++it) {
            k_imp_vec<<<gb, 256, 0, s>>>(a, n, 0);
            apply(imp_.vec[1], 2);
            reduce(kIter, it + 2);
            if ((it & 3) == 3 || it + 1 == maxit) {
                if ((rc = any_active(it + 2, &any))) return rc;
                if (!any) break;
            }
        }
    }
    k_imp_vec<<<gb, 256, 0, s>>>(a, n, 2);  // v = X, u = u* + (X - v*)
    HIP_TRY(hipGetLastError());
    return MPM_OK;
}

int Engine::cr_stats(int sid, int* iters, double* residual, int cap) {
    if (sid < 0 || sid >= (int)scenes_.size() || !scenes_[sid].used) {
        set_error("bad scene id");
        return MPM_ERR_ARG;
    }
    *iters = 0;
    if (!imp_.gm) return MPM_OK;
    CRScene c;
    HIP_TRY(hipMemcpyAsync(&c, imp_.cr + sid, sizeof c, hipMemcpyDeviceToHost, stream_));
    std::vector<double> h(kCRHist);
    HIP_TRY(hipMemcpyAsync(h.data(), imp_.hist + (size_t)sid * kCRHist, kCRHist * 8, hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    *iters = c.iters;
    if (residual)
        for (int k = 0; k <= c.iters && k < cap && k < kCRHist; ++k) residual[k] = h[k];
    return MPM_OK;
}

}  // namespace snowmpm

extern "C" int mpm_cr_stats(mpm_engine* e, int scene, int* iterations, double* residual, int cap) {
    if (!e || !iterations || cap < 0) return MPM_ERR_ARG;
    return e->eng.cr_stats(scene, iterations, residual, cap);
}
