// mpm_engine.hpp -- host engine class behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/snowmpm.h"
#include "mpm_device.hpp"
#include "mpm_kernels.hpp"

namespace snowmpm {

extern thread_local std::string g_last_error;
void set_error(const std::string& s);
// Looks up, once per device, the address of the static staging buffer of
// mpm_scene_download_noalloc (mpm_noalloc.hip); call with `device` current.
int scratch_setup(int device);

#define HIP_TRY(expr)                                                          \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));      \
            return MPM_ERR_HIP;                                                \
        }                                                                      \
    } while (0)

struct SceneHost {
    bool used = false;
    bool pending_drop = false;  // removed; particles leave at the next (permuting) step
    size_t n = 0;
    uint64_t steps = 0;  // steps taken since the scene was added
    mpm_params prm{};
};

// Device bytes a scene holds in the engine's three pools.
struct Footprint {
    size_t particles, blocks, partials;
    size_t total() const { return particles + blocks + partials; }
};

class Engine {
  public:
    Engine() = default;
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    int init(const mpm_config& cfg);
    int scene_add(const mpm_params& prm, const mpm_collider* coll, int ncoll, size_t n, const double* x,
                  const double* v, const double* m, const double* V0, const double* Fe, const double* Fp,
                  const double* E, int* sid);
    int scene_remove(int sid);
    int clear();
    int scene_count(int sid, size_t* n) const {
        if (sid < 0 || sid >= (int)scenes_.size() || !scenes_[sid].used) return MPM_ERR_ARG;
        *n = scenes_[sid].n;
        return MPM_OK;
    }
    int setup();
    int step(int nsteps, mpm_timings* t);
    int step_graph(int nsteps);
    // step_graph in three parts, so several engines can interleave their
    // launches on their own streams: graph_begin sets up new scenes (or
    // reports that a semi-implicit scene makes the step host-driven),
    // graph_enqueue launches one step without waiting, graph_finish waits and
    // checks the scenes' status.
    int graph_begin(bool* host_driven);
    int graph_enqueue();
    int graph_finish(int nsteps);
    int autotune(int steps_per_config, int chosen[4]);
    int permute_period() const { return cfg_.permute_period; }
    void tile_sizes(int* chunk, int* threads) const {
        *chunk = launch_.p2g_chunk;
        *threads = launch_.g2p_threads;
    }
    int serve(mpm_request* reqs, int nreq, int mode, int max_chunk, mpm_serve_stats* st);
    int scene_download(int sid, double* x, double* v, double* Fe, double* Fp, double* Jp, double* V0);
    // The same download staged through a caller-given device buffer of
    // scratch_bytes (no device allocation); larger requests are refused.
    int scene_download_static(int sid, double* x, double* v, double* Fe, double* Fp, double* Jp, double* V0,
                              double* scratch, size_t scratch_bytes);
    int device() const { return cfg_.device; }
    int scene_status(int sid, uint32_t* status, uint64_t* fb);
    int scene_sums(int sid, double out[7]);
    int scene_steps(int sid, uint64_t* steps) const {
        if (sid < 0 || sid >= (int)scenes_.size() || !scenes_[sid].used) {
            set_error("bad scene id");
            return MPM_ERR_ARG;
        }
        *steps = scenes_[sid].steps;
        return MPM_OK;
    }
    int cr_stats(int sid, int* iters, double* residual, int cap);
    Footprint footprint(size_t n, const double* x, double dx) const;
    Footprint pools() const;
    void kernel_times(double ms[5], int* nsteps) const {
        for (int j = 0; j < 5; ++j) ms[j] = kernel_ms_[j];
        if (nsteps) *nsteps = kernel_steps_;
    }
    // Particle slots in use (holes and removed scenes' particles included
    // until the next re-sort) and the slots of each buffer.
    int slots(size_t* used, size_t* total) {
        uint32_t n = 0;
        const int rc = device_count(&n);
        if (rc) return rc;
        *used = n;
        *total = slots_;
        return 0;
    }

  private:
    int dalloc(void** p, size_t bytes);
    void release();
    StepArgs args(uint32_t pass_flag, int permute = 1) const;
    PBuf view(int mov, int con) const { return compose_pbuf(buf_[mov], buf_[con]); }
    void advance(int permute) {
        mov_ ^= 1;
        if (permute) con_ ^= 1;
        for (auto& sh : scenes_)
            if (sh.used) ++sh.steps;
    }
    int device_count(uint32_t* n);
    int download_bytes(int sid, const double* const host[6], bool want_fp, size_t* bytes);
    // semi-implicit solve (mpm_implicit.hip)
    bool implicit_active() const;
    int implicit_alloc();
    int implicit_solve(const StepArgs& a);
    // one step of every active scene (explicit launches, or the solve between
    // the two halves when a scene has beta > 0); ev may be null
    int one_step(int permute, hipEvent_t* ev, mpm_timings* t);
    int download_via(int sid, double* const host[6], double* stage, bool fill_kernel);
    int upload_scene_table(int sid);
    int check_status();
    int ensure_graph(int permute);
    bool take_permute();
    void drop_graphs();

    mpm_config cfg_{};
    hipStream_t stream_ = nullptr;
    LaunchCfg launch_{};
    uint32_t H_ = 0;
    // Particle slots per buffer (and sorted positions): the capacity plus room
    // for the holes that start every home block's range on a 32-particle
    // (128-byte) boundary.
    size_t slots_ = 0;
    PBuf buf_[2]{};
    // Current particle state: moving fields (x v F_E S) in buf_[mov_], the
    // others in buf_[con_].  Every step writes the moving fields to the other
    // buffer and updates J_P and F_P in place where the plastic update changed
    // them; a permuting step moves all fields to the other buffers.
    int mov_ = 0;
    int con_ = 0;
    GridDev grid_{};
    BinDev bin_{};
    FallbackDev fb_{};
    StatusDev st_{};
    Counters* cnt_ = nullptr;
    SceneDev* d_scenes_ = nullptr;
    double* d_sums_ = nullptr;
    void* host_stage_ = nullptr;
    bool any_E_ = false;
    bool any_project_ = false;  // some scene was added with projecting colliders
    std::vector<SceneDev> h_scenes_;
    std::vector<SceneHost> scenes_;
    std::vector<void*> allocs_;
    std::vector<hipEvent_t> events_;
    // Steps update the particles at their index and re-sort the buffers (a
    // permuting step: particles move to their sorted positions, dropped scenes
    // are compacted away) every cfg_.permute_period steps, and on the first
    // step after scenes were added or removed.
    bool need_permute_ = true;
    bool tiles_tuned_ = false;  // mpm_autotune chose the tile sizes
    uint64_t step_count_ = 0;
    hipGraphExec_t gexec_[8] = {};  // [(mov_ * 2 + con_) * 2 + permute]
    StepArgs gargs_[8]{};
    LaunchCfg glaunch_[8]{};
    ImpDev imp_{};
    uint64_t* imp_kin_ = nullptr;
    uint32_t* imp_vin_ = nullptr;
    void* imp_sort_tmp_ = nullptr;
    size_t imp_sort_bytes_ = 0;
    int cus_ = 256;
    double kernel_ms_[5] = {0, 0, 0, 0, 0};
    int kernel_steps_ = 0;
};

}  // namespace snowmpm

// The opaque handle of the C ABI.
struct mpm_engine {
    snowmpm::Engine eng;
};
