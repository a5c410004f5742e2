// mpm_probe.hip -- measurements behind the kernels' bounds.  The issue-rate
// probe: how many SIMD cycles a v_mfma_f32_16x16x4f32 and a wave64 FP32 FMA
// take, alone and interleaved (if the matrix core and the VALU issued side
// by side, the mixed loop would cost the larger of the two alone; it costs
// their sum).  The streaming probe: the HBM rate of a plain copy and of
// G2P's particle streams with no arithmetic.
//
// Every wave runs `iters` iterations of 4 independent MFMA accumulations
// (mode 0), of 4 K FMAs over 8 independent chains (mode 1), or of both
// interleaved (mode 2), and records its shader-clock cycles.  The grid is
// 4 waves per CU per requested wave per SIMD; no memory is read and each
// thread writes one float (keeping the results alive) and each wave one
// cycle count.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/snowmpm.h"

namespace snowmpm {

typedef float probe_v4f __attribute__((ext_vector_type(4)));

template <int MODE, int K>
__global__ __launch_bounds__(256) void k_probe_issue(int iters, unsigned long long* cycles, float* sink) {
    const float a = 1.0f + 1e-7f * (float)threadIdx.x, b = 1.0f - 1e-7f * (float)threadIdx.x;
    probe_v4f c0 = {0.f, 0.f, 0.f, 0.f}, c1 = c0, c2 = c0, c3 = c0;
    float x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = (float)(threadIdx.x + j);
    const float y = 0.999999f, z = 1e-6f;
    const unsigned long long t0 = clock64();
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (MODE != 1) {
                if (q == 0) c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c0, 0, 0, 0);
                if (q == 1) c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c1, 0, 0, 0);
                if (q == 2) c2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c2, 0, 0, 0);
                if (q == 3) c3 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c3, 0, 0, 0);
            }
            if (MODE != 0) {
#pragma unroll
                for (int k = 0; k < K; ++k) x[(q * K + k) & 7] = __builtin_fmaf(x[(q * K + k) & 7], y, z);
            }
        }
    }
    const unsigned long long t1 = clock64();
    float s = c0[0] + c1[1] + c2[2] + c3[3];
#pragma unroll
    for (int j = 0; j < 8; ++j) s += x[j];
    const unsigned gtid = blockIdx.x * blockDim.x + threadIdx.x;
    sink[gtid] = s;
    if ((threadIdx.x & 63) == 0) cycles[gtid >> 6] = t1 - t0;
}

template <int K>
static int probe_run(int mode, int blocks, int iters, unsigned long long* dcyc, float* dsink, hipEvent_t e0,
                     hipEvent_t e1, double& cyc, double& us) {
    auto launch = [&]() {
        if (mode == 0) k_probe_issue<0, K><<<blocks, 256>>>(iters, dcyc, dsink);
        if (mode == 1) k_probe_issue<1, K><<<blocks, 256>>>(iters, dcyc, dsink);
        if (mode == 2) k_probe_issue<2, K><<<blocks, 256>>>(iters, dcyc, dsink);
    };
    launch();  // warm-up (clocks ramp, code loads)
    if (hipEventRecord(e0, nullptr) != hipSuccess) return MPM_ERR_HIP;
    launch();
    if (hipEventRecord(e1, nullptr) != hipSuccess) return MPM_ERR_HIP;
    if (hipEventSynchronize(e1) != hipSuccess) return MPM_ERR_HIP;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return MPM_ERR_HIP;
    const size_t nw = (size_t)blocks * 4;
    std::vector<unsigned long long> h(nw);
    if (hipMemcpy(h.data(), dcyc, nw * sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess)
        return MPM_ERR_HIP;
    cyc = (double)*std::max_element(h.begin(), h.end()) / iters;
    us = ms * 1e3;
    return MPM_OK;
}

// ---------------------------------------------------------------------------
// HBM streaming probe: the rate a kernel reaches when it only moves bytes,
// for a plain copy and for the particle streams of G2P on the 8M-particle
// slab (19 FP32 arrays read: x v F_E J_P V0 E and the sorted index; 22
// written: x v F_E J_P S), so a kernel's traffic can be set against what
// its access pattern allows on this part.
constexpr uint32_t kStreamN = 8000000u;
constexpr int kStreamR = 19, kStreamW = 22;

__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) b[i] = a[i];
}
// ONE = false: a 128-thread workgroup per 512 consecutive elements (4
// passes), as G2P takes a home block; ONE = true: one element per thread.
template <bool ONE>
__global__ __launch_bounds__(256) void k_stream_fields(const float* __restrict__ in, float* __restrict__ out,
                                                       uint32_t n) {
    const uint32_t first = ONE ? blockIdx.x * blockDim.x + threadIdx.x : blockIdx.x * 512u + threadIdx.x;
    const uint32_t last = ONE ? first + 1u : blockIdx.x * 512u + 512u;
    for (uint32_t j = first; j < last && j < n; j += blockDim.x) {
        float s = 0.f;
#pragma unroll
        for (int e = 0; e < kStreamR; ++e) s += in[(size_t)e * n + j];
#pragma unroll
        for (int e = 0; e < kStreamW; ++e) out[(size_t)e * n + j] = s + (float)e;
    }
}

}  // namespace snowmpm

extern "C" int mpm_probe_stream(int device, int mode, double out[3]) {
    using namespace snowmpm;
    if (!out || mode < 0 || mode > 2) return MPM_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return MPM_ERR_HIP;
    const size_t copy_bytes = (size_t)1 << 30;
    const size_t nin = mode == 0 ? copy_bytes : (size_t)kStreamR * kStreamN * sizeof(float);
    const size_t nout = mode == 0 ? copy_bytes : (size_t)kStreamW * kStreamN * sizeof(float);
    const double moved = (double)(nin + nout);
    void *din = nullptr, *dout = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MPM_OK;
    if (hipMalloc(&din, nin) != hipSuccess || hipMalloc(&dout, nout) != hipSuccess ||
        hipMemset(din, 0, nin) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        rc = MPM_ERR_HIP;
    auto launch = [&]() {
        if (mode == 0) {
            const size_t n4 = copy_bytes / sizeof(float4);
            k_stream_copy<<<(unsigned)((n4 + 255) / 256), 256>>>((const float4*)din, (float4*)dout, n4);
        } else if (mode == 1) {
            k_stream_fields<false><<<(kStreamN + 511) / 512, 128>>>((const float*)din, (float*)dout, kStreamN);
        } else {
            k_stream_fields<true><<<(kStreamN + 255) / 256, 256>>>((const float*)din, (float*)dout, kStreamN);
        }
    };
    constexpr int kTimed = 10;
    float ms = 0.f;
    if (rc == MPM_OK) {
        launch();  // warm-up
        launch();
        if (hipEventRecord(e0, nullptr) != hipSuccess) rc = MPM_ERR_HIP;
        for (int k = 0; k < kTimed && rc == MPM_OK; ++k) launch();
        if (rc == MPM_OK && (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                             hipEventElapsedTime(&ms, e0, e1) != hipSuccess || hipGetLastError() != hipSuccess))
            rc = MPM_ERR_HIP;
    }
    if (rc == MPM_OK) {
        out[0] = ms * 1e3 / kTimed;
        out[1] = moved;
        out[2] = moved / (out[0] * 1e-6) / 1e12;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    return rc;
}

extern "C" int mpm_probe_issue(int device, int waves_per_simd, int valu_per_mfma, int iterations, double out[6]) {
    using namespace snowmpm;
    if (!out || waves_per_simd < 1 || waves_per_simd > 8 || iterations < 1 || iterations > (1 << 20))
        return MPM_ERR_ARG;
    const int K = valu_per_mfma;
    if (K != 1 && K != 2 && K != 4 && K != 8 && K != 16) return MPM_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return MPM_ERR_HIP;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return MPM_ERR_HIP;
    const int blocks = cus * waves_per_simd;  // 256-thread workgroups: one wave per SIMD each
    unsigned long long* dcyc = nullptr;
    float* dsink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MPM_OK;
    if (hipMalloc((void**)&dcyc, (size_t)blocks * 4 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc((void**)&dsink, (size_t)blocks * 256 * sizeof(float)) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = MPM_ERR_HIP;
    for (int mode = 0; mode < 3 && rc == MPM_OK; ++mode) {
        double cyc = 0.0, us = 0.0;
        switch (K) {
            case 1: rc = probe_run<1>(mode, blocks, iterations, dcyc, dsink, e0, e1, cyc, us); break;
            case 2: rc = probe_run<2>(mode, blocks, iterations, dcyc, dsink, e0, e1, cyc, us); break;
            case 4: rc = probe_run<4>(mode, blocks, iterations, dcyc, dsink, e0, e1, cyc, us); break;
            case 8: rc = probe_run<8>(mode, blocks, iterations, dcyc, dsink, e0, e1, cyc, us); break;
            default: rc = probe_run<16>(mode, blocks, iterations, dcyc, dsink, e0, e1, cyc, us); break;
        }
        out[mode] = cyc;
        out[3 + mode] = us;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (dcyc) (void)hipFree(dcyc);
    if (dsink) (void)hipFree(dsink);
    return rc;
}
