// mpm_noalloc.hip -- the engine's one entry point that allocates no device
// memory: a scene download staged through a static device buffer.
//
// mpm_scene_download stages the caller's arrays on the GPU (a kernel writes
// them in the caller's layout, then one copy per array) in a buffer it
// allocates per call.  Callers that cannot allocate (tight serving loops,
// harnesses that forbid it) use mpm_scene_download_noalloc instead, which
// stages through `scratch`, a file-scope device array of kScratchBytes that
// every GPU holds for the life of the process.  Its device address is looked
// up once per device when an engine is created there (mpm_create), with
// that device current, and only that address is passed to kernels and
// copies.  Every call on a GPU shares that GPU's array: calls that overlap on
// one GPU (two host threads, or two engines on the same device) race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/snowmpm.h"
#include "mpm_engine.hpp"

namespace snowmpm {

constexpr size_t kScratchBytes = (size_t)32 << 20;  // 32 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Called by Engine::init with `device` current.
int scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device ordinal out of range for the static staging buffer");
        return MPM_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return MPM_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return MPM_ERR_HIP;
    }
    g_scratch_ptr[device] = p;
    return MPM_OK;
}

static double* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return static_cast<double*>(g_scratch_ptr[device]);
}

}  // namespace snowmpm

extern "C" int mpm_scene_download_noalloc(mpm_engine* e, int scene, double* x, double* v, double* Fe, double* Fp,
                                          double* Jp, double* V0, size_t* max_bytes) {
    if (max_bytes) *max_bytes = snowmpm::kScratchBytes;
    if (!e) return MPM_ERR_ARG;
    double* p = snowmpm::scratch_ptr(e->eng.device());
    if (!p) {
        snowmpm::set_error("static staging buffer not set up on this device");
        return MPM_ERR_STATE;
    }
    return e->eng.scene_download_static(scene, x, v, Fe, Fp, Jp, V0, p, snowmpm::kScratchBytes);
}
