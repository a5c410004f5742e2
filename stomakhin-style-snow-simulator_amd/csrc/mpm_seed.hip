// mpm_seed.hip -- lattice seeding of a closed triangle mesh on the GPU.
//
// Replaces the reference's Blender add-on seeding (blender/seed_export.py):
//   world_bounds       (43-62)   axis-aligned bounds of the mesh, float32
//   ray_inside         (64-86)   hits counted along a ray by re-casting from
//                                just past each hit (at most 256 casts,
//                                epsilon 1e-5); inside when the count is odd
//   point_inside_mesh  (88-104)  majority vote (>= 2) of three fixed rays
//                                (RAY_DIRECTIONS, 7-11)
//   seed_mesh          (106-131) lattice min + spacing/2 + k spacing per axis,
//                                accumulated in double as the add-on's Python
//                                loop does, points stored as float32
// Blender's parts are taken from their definitions: mathutils vectors and
// float properties are float32, Vector.normalized() scales by 1/sqrt(dot),
// Object.ray_cast returns the nearest double-sided hit at distance >= 0 below
// its default distance 1.70141e38, at origin + t dir.  The ray/triangle test
// is the watertight test of Woop, Benthin and Wald (JCGT 2(1), 2013): shear
// to the ray's dominant axis, edge functions U, V, W (redone in double when
// one is zero), no back-face culling, t = T/det correctly rounded.
//
// One thread per lattice point traces the three rays; the triangle loop is
// the same for every lane, so triangle corners come through scalar loads.
// No floating-point contraction: every product and sum rounds on its own, so
// the result does not depend on the compiler's fusing.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/snowmpm.h"

namespace snowmpm {
void set_error(const std::string& s);
}

namespace {

constexpr float kSeedDistMax = 1.70141183e38f;  // Object.ray_cast default distance
constexpr float kSeedEps = 1e-5f;               // seed_export.py:68
constexpr int kSeedCasts = 256;                 // seed_export.py:70
constexpr int kSeedBlock = 256;

// Three rays: unit direction, dominant axis kz (kx, ky the sheared axes,
// swapped when the direction along kz is negative) and shear constants.
struct SeedRays {
    float d[3][3];
    int kx[3], ky[3], kz[3];
    float sx[3], sy[3], sz[3];
};

SeedRays seed_rays() {
#pragma clang fp contract(off)
    static const float raw[3][3] = {{1.0f, 0.371f, 0.529f}, {-0.427f, 1.0f, 0.233f}, {0.191f, -0.463f, 1.0f}};
    SeedRays r{};
    for (int k = 0; k < 3; ++k) {
        const float* v = raw[k];
        const float len2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        const float s = 1.0f / std::sqrt(len2);
        for (int i = 0; i < 3; ++i) r.d[k][i] = v[i] * s;
        int kz = 0;
        if (std::fabs(r.d[k][1]) > std::fabs(r.d[k][kz])) kz = 1;
        if (std::fabs(r.d[k][2]) > std::fabs(r.d[k][kz])) kz = 2;
        int kx = kz == 2 ? 0 : kz + 1;
        int ky = kx == 2 ? 0 : kx + 1;
        if (r.d[k][kz] < 0.0f) std::swap(kx, ky);
        r.kx[k] = kx;
        r.ky[k] = ky;
        r.kz[k] = kz;
        r.sx[k] = r.d[k][kx] / r.d[k][kz];
        r.sy[k] = r.d[k][ky] / r.d[k][kz];
        r.sz[k] = 1.0f / r.d[k][kz];
    }
    return r;
}

// The add-on's three directions are dominant along x, y and z in turn, each
// with a positive dominant component, so ray k shears onto axis KZ = k with
// KX = k + 1, KY = k + 2 (mod 3); seed_rays() derives the same axes and the
// launch checks that it did, so the kernel can fix them at compile time.
template <int KZ>
__device__ __forceinline__ float seed_hit(const SeedRays& r, const float o[3], const float A[3], const float B[3],
                                          const float C[3]) {
#pragma clang fp contract(off)
    constexpr int kx = (KZ + 1) % 3, ky = (KZ + 2) % 3, kz = KZ;
    float a[3], b[3], c[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        a[i] = A[i] - o[i];
        b[i] = B[i] - o[i];
        c[i] = C[i] - o[i];
    }
    const float sx = r.sx[KZ], sy = r.sy[KZ], sz = r.sz[KZ];
    const float akz = a[kz], bkz = b[kz], ckz = c[kz];
    const float ax = a[kx] - sx * akz, ay = a[ky] - sy * akz;
    const float bx = b[kx] - sx * bkz, by = b[ky] - sy * bkz;
    const float cx = c[kx] - sx * ckz, cy = c[ky] - sy * ckz;
    float U = cx * by - cy * bx;
    float V = ax * cy - ay * cx;
    float W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)cx * (double)by - (double)cy * (double)bx);
        V = (float)((double)ax * (double)cy - (double)ay * (double)cx);
        W = (float)((double)bx * (double)ay - (double)by * (double)ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return kSeedDistMax;
    const float det = U + V + W;
    if (det == 0.0f) return kSeedDistMax;
    const float T = U * (sz * akz) + V * (sz * bkz) + W * (sz * ckz);
    if ((det > 0.0f && T < 0.0f) || (det < 0.0f && T > 0.0f)) return kSeedDistMax;
    return (float)((double)T / (double)det);  // double quotient rounded once: the correctly rounded float
}

// inside[p] = 1 when lattice point p = (i, j, k) (k fastest) is inside the
// mesh by the three-ray vote.
__global__ __launch_bounds__(kSeedBlock) void k_seed(const float* __restrict__ tri, uint32_t ntris,
                                                     const float* __restrict__ ax, const float* __restrict__ ay,
                                                     const float* __restrict__ az, uint32_t ny, uint32_t nz,
                                                     uint64_t npts, SeedRays r, uint8_t* __restrict__ inside) {
#pragma clang fp contract(off)
    const uint64_t p = (uint64_t)blockIdx.x * kSeedBlock + threadIdx.x;
    if (p >= npts) return;
    const uint32_t k = (uint32_t)(p % nz), j = (uint32_t)((p / nz) % ny), i = (uint32_t)(p / ((uint64_t)nz * ny));
    const float P[3] = {ax[i], ay[j], az[k]};
    float o[3][3];
    int hits[3] = {0, 0, 0};
    bool live[3] = {true, true, true};
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int e = 0; e < 3; ++e) o[q][e] = P[e];
    for (int cast = 0; cast < kSeedCasts && (live[0] || live[1] || live[2]); ++cast) {
        float best[3] = {kSeedDistMax, kSeedDistMax, kSeedDistMax};
        for (uint32_t t = 0; t < ntris; ++t) {
            const float* T = tri + 9 * (size_t)t;
            const float A[3] = {T[0], T[1], T[2]}, B[3] = {T[3], T[4], T[5]}, C[3] = {T[6], T[7], T[8]};
            if (live[0]) best[0] = fminf(best[0], seed_hit<0>(r, o[0], A, B, C));
            if (live[1]) best[1] = fminf(best[1], seed_hit<1>(r, o[1], A, B, C));
            if (live[2]) best[2] = fminf(best[2], seed_hit<2>(r, o[2], A, B, C));
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (!live[q]) continue;
            if (best[q] < kSeedDistMax) {
                ++hits[q];
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    const float loc = o[q][e] + r.d[q][e] * best[q];
                    o[q][e] = loc + r.d[q][e] * kSeedEps;
                }
            } else {
                live[q] = false;
            }
        }
    }
    inside[p] = (uint8_t)(((hits[0] & 1) + (hits[1] & 1) + (hits[2] & 1)) >= 2);
}

// Lattice coordinates along one axis (seed_mesh, seed_export.py:109-127):
// the spacing is the float32 property value, the loop variable a double.
std::vector<float> lattice_axis(float lo, float hi, double spacing) {
    const double s = (double)(float)spacing;
    std::vector<float> out;
    for (double x = (double)lo + s * 0.5; x < (double)hi; x += s) out.push_back((float)x);
    return out;
}

int hip_fail(hipError_t e, const char* what) {
    snowmpm::set_error(std::string("mpm_seed_mesh: ") + what + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? MPM_ERR_OOM : MPM_ERR_HIP;
}

}  // namespace

extern "C" int mpm_seed_mesh(int device, const double* verts, size_t nverts, const int32_t* tris, size_t ntris,
                             double spacing, size_t* n, double** x) {
    if (!n || !x || (nverts && !verts) || (ntris && !tris) || !(spacing > 0.0) || !std::isfinite(spacing) ||
        ntris > 0xFFFFFFFFull) {
        snowmpm::set_error("mpm_seed_mesh: bad arguments");
        return MPM_ERR_ARG;
    }
    *n = 0;
    *x = nullptr;
    // Triangle corners as float32 (Blender stores mesh coordinates so).
    std::vector<float> tri(9 * ntris);
    for (size_t t = 0; t < ntris; ++t)
        for (int c = 0; c < 3; ++c) {
            const int32_t id = tris[3 * t + c];
            if (id < 0 || (size_t)id >= nverts) {
                snowmpm::set_error("mpm_seed_mesh: triangle " + std::to_string(t) + " names vertex " +
                                   std::to_string(id) + " of " + std::to_string(nverts));
                return MPM_ERR_ARG;
            }
            for (int e = 0; e < 3; ++e) tri[9 * t + 3 * c + e] = (float)verts[3 * (size_t)id + e];
        }
    // world_bounds: an empty mesh has an all-zero box, which seeds nothing.
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
    for (int e = 0; e < 3 && nverts; ++e) {
        lo[e] = hi[e] = (float)verts[e];
        for (size_t v = 1; v < nverts; ++v) {
            const float c = (float)verts[3 * v + e];
            lo[e] = std::fmin(lo[e], c);
            hi[e] = std::fmax(hi[e], c);
        }
    }
    const std::vector<float> axis[3] = {lattice_axis(lo[0], hi[0], spacing), lattice_axis(lo[1], hi[1], spacing),
                                        lattice_axis(lo[2], hi[2], spacing)};
    const uint64_t npts = (uint64_t)axis[0].size() * axis[1].size() * axis[2].size();
    if (npts > (1ull << 32)) {
        snowmpm::set_error("mpm_seed_mesh: lattice of " + std::to_string(npts) + " points exceeds 2^32");
        return MPM_ERR_ARG;
    }
    if (npts == 0 || ntris == 0) return MPM_OK;
    const SeedRays rays = seed_rays();
    for (int k = 0; k < 3; ++k)
        if (rays.kz[k] != k || rays.kx[k] != (k + 1) % 3 || rays.ky[k] != (k + 2) % 3) {
            snowmpm::set_error("mpm_seed_mesh: ray axes differ from the kernel's");
            return MPM_ERR_STATE;
        }

    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    float *d_tri = nullptr, *d_ax[3] = {nullptr, nullptr, nullptr};
    uint8_t* d_in = nullptr;
    hipStream_t s = nullptr;
    std::vector<uint8_t> inside(npts);
    int rc = MPM_OK;
    auto check = [&](hipError_t err, const char* what) {
        if (err != hipSuccess && rc == MPM_OK) rc = hip_fail(err, what);
        return rc == MPM_OK;
    };
    if (check(hipStreamCreate(&s), "hipStreamCreate") && check(hipMalloc(&d_tri, tri.size() * 4), "hipMalloc") &&
        check(hipMalloc(&d_ax[0], axis[0].size() * 4), "hipMalloc") &&
        check(hipMalloc(&d_ax[1], axis[1].size() * 4), "hipMalloc") &&
        check(hipMalloc(&d_ax[2], axis[2].size() * 4), "hipMalloc") && check(hipMalloc(&d_in, npts), "hipMalloc") &&
        check(hipMemcpyAsync(d_tri, tri.data(), tri.size() * 4, hipMemcpyHostToDevice, s), "hipMemcpy") &&
        check(hipMemcpyAsync(d_ax[0], axis[0].data(), axis[0].size() * 4, hipMemcpyHostToDevice, s), "hipMemcpy") &&
        check(hipMemcpyAsync(d_ax[1], axis[1].data(), axis[1].size() * 4, hipMemcpyHostToDevice, s), "hipMemcpy") &&
        check(hipMemcpyAsync(d_ax[2], axis[2].data(), axis[2].size() * 4, hipMemcpyHostToDevice, s), "hipMemcpy")) {
        const uint64_t blocks = (npts + kSeedBlock - 1) / kSeedBlock;
        k_seed<<<dim3((uint32_t)blocks), dim3(kSeedBlock), 0, s>>>(d_tri, (uint32_t)ntris, d_ax[0], d_ax[1], d_ax[2],
                                                                   (uint32_t)axis[1].size(), (uint32_t)axis[2].size(),
                                                                   npts, rays, d_in);
        if (check(hipGetLastError(), "k_seed launch") &&
            check(hipMemcpyAsync(inside.data(), d_in, npts, hipMemcpyDeviceToHost, s), "hipMemcpy"))
            check(hipStreamSynchronize(s), "k_seed");
    }
    if (d_in) (void)hipFree(d_in);
    for (float* p : d_ax)
        if (p) (void)hipFree(p);
    if (d_tri) (void)hipFree(d_tri);
    if (s) (void)hipStreamDestroy(s);
    if (rc != MPM_OK) return rc;

    size_t cnt = 0;
    for (uint8_t f : inside) cnt += f;
    if (cnt == 0) return MPM_OK;
    double* out = static_cast<double*>(std::malloc(cnt * 3 * sizeof(double)));
    if (!out) {
        snowmpm::set_error("out of host memory");
        return MPM_ERR_OOM;
    }
    const size_t ny = axis[1].size(), nz = axis[2].size();
    size_t w = 0;
    for (uint64_t p = 0; p < npts; ++p)
        if (inside[p]) {
            out[3 * w] = axis[0][p / (nz * ny)];
            out[3 * w + 1] = axis[1][(p / nz) % ny];
            out[3 * w + 2] = axis[2][p % nz];
            ++w;
        }
    *n = cnt;
    *x = out;
    return MPM_OK;
}
