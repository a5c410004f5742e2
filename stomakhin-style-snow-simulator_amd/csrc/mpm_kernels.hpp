// mpm_kernels.hpp -- kernel argument block and host launch wrappers.
#pragma once

#include <hip/hip_runtime.h>

#include "mpm_device.hpp"

namespace snowmpm {

struct StepArgs {
    PBuf cur;             // particles read this pass
    PBuf nxt;             // particles written by G2P: all fields at sorted positions
                          // (permute), else the moving fields at the same index
    GridDev g;
    BinDev bin;
    StatusDev st;
    FallbackDev fb;
    Counters* cnt;
    const SceneDev* scenes;
    uint32_t pass_flag;   // kSceneActive for steps, kSceneSetup for setup
    int mass_only;
    int copy_E;           // some scene carries per-particle E
    float kappa;          // FP64 fallback bound on cond(trial F_E)
    int permute;          // 1: G2P writes every particle to nxt in sorted order (and
                          //    compacts); 0: only x v F_E J_P [F_P] S, at index i
    int project;          // some scene has colliders with position projection
    uint32_t hole_room;   // particle slots per buffer beyond the capacity (room for holes)
    int align_blocks;     // home blocks' sorted ranges start on 32-particle boundaries when the holes fit
    ImpDev imp;           // semi-implicit solve (imp.gm null: explicit steps only)
};

struct LaunchCfg {
    int grid_p;  // per-particle grid-stride kernels
    int grid_b;  // one wave per block kernels (4 blocks per workgroup)
    int grid_w;  // one workgroup per block kernels (P2G, G2P)
    int grid_l;  // hash-slot compaction (4096 slots per workgroup)
    int grid_k;  // k_blocks (16 blocks per workgroup)
    int grid_h;  // lower bound of the P2G grid (one wave per home block)
    int p2g_chunk;    // particles P2G stages per pass: 16, 32, 48 or 64
    int g2p_threads;  // G2P workgroup size: 64, 128 or 256
};

// Kernel slots timed by launch_step's events:
// [0,1) binning+sort, [1,2) P2G, [2,3) grid, [3,4) G2P, [4,5) end.
constexpr int kStepEvents = 6;

void launch_step(const StepArgs& a, const LaunchCfg& c, hipStream_t s, hipEvent_t* ev);
// The same step in two parts around the semi-implicit solve: binning, P2G and
// the grid update (with imp.gm set: the unknowns' masses), then G2P and the end.
void launch_step_head(const StepArgs& a, const LaunchCfg& c, hipStream_t s);
void launch_step_tail(const StepArgs& a, const LaunchCfg& c, hipStream_t s);
// P2G of the force differential only: reads imp.G instead of the force matrix.
void launch_p2g_force(const StepArgs& a, const LaunchCfg& c, hipStream_t s);
void launch_setup(const StepArgs& a, const LaunchCfg& c, hipStream_t s);
void launch_init_stress(const StepArgs& a, uint32_t first, uint32_t n, hipStream_t s);
void launch_set_count(Counters* c, uint32_t n, hipStream_t s);
// Device outputs of a scene download, in the reference's host layout
// (doubles); a null member is not gathered.
struct SceneOut {
    double *x, *v, *Fe, *Fp, *Jp, *V0;
};
// Device copies of a scene's host arrays (reference layout, doubles); x, v
// and m are required, a null V0, Fe, Fp or E takes its default (0, I, I, 0).
struct SceneIn {
    const double *x, *v, *m, *V0, *Fe, *Fp, *E;
};
void launch_put_scene(const StepArgs& a, const LaunchCfg& c, uint32_t first, uint32_t n, uint32_t sid,
                      const SceneIn& in, hipStream_t s);
void launch_gather_scene(const StepArgs& a, const LaunchCfg& c, uint32_t sid, uint32_t nscene, const SceneOut& o,
                         hipStream_t s);
// Sets n doubles at p to zero (a kernel, so it can stand in captured work).
void launch_fill_zero(double* p, size_t n, hipStream_t s);
void launch_scene_sums(const StepArgs& a, uint32_t sid, double* out, int nblocks, hipStream_t s);

}  // namespace snowmpm
