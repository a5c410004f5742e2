// mpm_engine.cpp -- C++20 host engine behind the C ABI (include/snowmpm.h).
//
// Owns the device state of many resident scenes (particle double buffer,
// block hash grid, binning arrays, status words) and drives the kernels of
// mpm_kernels.hip on one HIP stream.  All scenes step together: one launch
// per stage covers every active scene, because block keys carry the scene id.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/snowmpm.h"
#include "mpm_device.hpp"
#include "mpm_engine.hpp"
#include "mpm_kernels.hpp"

namespace snowmpm {

thread_local std::string g_last_error;

void set_error(const std::string& s) { g_last_error = s; }

Engine::~Engine() { release(); }

void Engine::drop_graphs() {
    for (auto& g : gexec_)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

void Engine::release() {
    drop_graphs();
    for (void* p : allocs_) (void)hipFree(p);
    allocs_.clear();
    for (auto& ev : events_) (void)hipEventDestroy(ev);
    events_.clear();
    if (stream_) (void)hipStreamDestroy(stream_);
    stream_ = nullptr;
    if (host_stage_) (void)hipHostFree(host_stage_);
    host_stage_ = nullptr;
}

int Engine::dalloc(void** p, size_t bytes) {
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        set_error("hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
        return MPM_ERR_OOM;
    }
    allocs_.push_back(*p);
    return MPM_OK;
}

int Engine::init(const mpm_config& cfg) {
    cfg_ = cfg;
    // Kernels address particle arrays with 32-bit byte offsets (index * 4).
    if (cfg_.particle_capacity == 0 || cfg_.particle_capacity >= (1ull << 30)) {
        set_error("particle_capacity must be in [1, 2^30)");
        return MPM_ERR_ARG;
    }
    if (cfg_.kappa <= 0) cfg_.kappa = 1.5;
    if (cfg_.permute_period < 0) {
        set_error("permute_period < 0");
        return MPM_ERR_ARG;
    }
    if (cfg_.permute_period == 0) cfg_.permute_period = 16;
    if ((cfg_.p2g_chunk != 0 && cfg_.p2g_chunk != 16 && cfg_.p2g_chunk != 32 && cfg_.p2g_chunk != 48 &&
         cfg_.p2g_chunk != 64) ||
        (cfg_.g2p_threads != 0 && cfg_.g2p_threads != 64 && cfg_.g2p_threads != 128 && cfg_.g2p_threads != 256)) {
        set_error("p2g_chunk must be 0, 16, 32, 48 or 64 and g2p_threads 0, 64, 128 or 256");
        return MPM_ERR_ARG;
    }
    HIP_TRY(hipSetDevice(cfg_.device));
    int rc0 = scratch_setup(cfg_.device);  // the static staging buffer's address on this device
    if (rc0) return rc0;
    HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    int cus = 256;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, cfg_.device));
    cus_ = cus;
    const size_t N = cfg_.particle_capacity;
    // Home blocks take sorted ranges that start on 32-particle boundaries
    // when their holes fit in the slots beyond the capacity (k_blocks), so a
    // re-sort leaves holes in the buffers: a quarter of the capacity is kept
    // for them.
    slots_ = cfg_.dense_blocks ? N : std::min<size_t>((N + N / 4 + 64 + 63) & ~(size_t)63, ((size_t)1 << 30) - 64);
    if (slots_ < N) slots_ = N;
    launch_.grid_p = (int)std::min<size_t>((N + 255) / 256, (size_t)cus * 16);
    launch_.grid_b = cus * 8;
    launch_.grid_w = cus * 8;

    int hl = cfg_.hash_log2;
    if (hl <= 0) {
        size_t want = std::max<size_t>(65536, N / 16);
        hl = 16;
        while (((size_t)1 << hl) < want) ++hl;
    }
    // A cell key is (slot << 6) | cell in 32 bits and 0xFFFFFFFF marks an
    // unbinned particle, so slots must stay below 2^25 for no key to equal it.
    if (hl > 25) {
        set_error("hash_log2 must be <= 25");
        return MPM_ERR_ARG;
    }
    H_ = 1u << hl;
    launch_.grid_l = (int)std::max<uint32_t>(1u, H_ / 4096u);
    launch_.grid_k = cus * 2;
    // P2G: enough single-wave workgroups for every home block of a scene
    // with >= 256 particles per occupied block (a dense 8-per-cell block has
    // 512); more home blocks than that fall back to the persistent loop.
    launch_.grid_h = (int)std::min<size_t>((N + 255) / 256, (size_t)1 << 20);
    launch_.p2g_chunk = cfg_.p2g_chunk ? cfg_.p2g_chunk : 48;
    launch_.g2p_threads = cfg_.g2p_threads ? cfg_.g2p_threads : 128;
    const bool track = cfg_.track_fp != 0;

    int rc;
    for (int b = 0; b < 2; ++b) {
        PBuf& p = buf_[b];
        std::memset(&p, 0, sizeof p);
        const int nf = 3 + 3 + 9 + 1 + (track ? 9 : 0) + 1 + 1 + 1 + 6;
        const size_t S = slots_;
        float* base;
        if ((rc = dalloc((void**)&base, S * sizeof(float) * nf))) return rc;
        float* q = base;
        auto take = [&]() { float* r = q; q += S; return r; };
        for (int e = 0; e < 3; ++e) p.x[e] = take();
        for (int e = 0; e < 3; ++e) p.v[e] = take();
        for (int e = 0; e < 9; ++e) p.F[e] = take();
        p.Jp = take();
        if (track)
            for (int e = 0; e < 9; ++e) p.Fp[e] = take();
        p.m = take();
        p.V0 = take();
        p.E = take();
        for (int e = 0; e < 6; ++e) p.S[e] = take();
        p.stride = (uint32_t)S;  // components of x, v, F, Fp and S lie slots_ elements apart
        if ((rc = dalloc((void**)&p.pid, S * 4))) return rc;
        if ((rc = dalloc((void**)&p.sid, S * 4))) return rc;
    }
    if ((rc = dalloc((void**)&bin_.pkey, slots_ * 4))) return rc;
    if ((rc = dalloc((void**)&bin_.prank, slots_ * 4))) return rc;
    if ((rc = dalloc((void**)&bin_.sidx, slots_ * 4))) return rc;
    if ((rc = dalloc((void**)&bin_.spid, slots_ * 4))) return rc;
    if ((rc = dalloc((void**)&fb_.pos, N * 4))) return rc;
    if ((rc = dalloc((void**)&fb_.L, N * 36))) return rc;

    grid_.mask = H_ - 1;
    grid_.max_probe = std::min<uint32_t>(4096u, H_);  // a full table is scanned once at most
    if ((rc = dalloc((void**)&grid_.keys, (size_t)H_ * 8))) return rc;
    if ((rc = dalloc((void**)&grid_.cellcnt, (size_t)H_ * 64 * 4))) return rc;
    if ((rc = dalloc((void**)&grid_.celloff, (size_t)H_ * 64 * 4))) return rc;
    if ((rc = dalloc((void**)&grid_.blkoff, (size_t)H_ * 4))) return rc;
    if ((rc = dalloc((void**)&grid_.blkcnt, (size_t)H_ * 4))) return rc;
    if ((rc = dalloc((void**)&grid_.nb, (size_t)H_ * 8 * 4))) return rc;
    // P2G partials: the 7 node planes of a 7x7x7-node region, 7 channels, per
    // home block (kPartFloats); a home block holds at least one particle, so
    // N bounds the count as well as H / 2.
    grid_.max_home = (uint32_t)std::min<size_t>(N + 64, H_ / 2);
    if ((rc = dalloc((void**)&grid_.part, (size_t)grid_.max_home * kPartFloats * 4))) return rc;
    if ((rc = dalloc((void**)&grid_.rev, (size_t)H_ * 8 * 4))) return rc;
    if ((rc = dalloc((void**)&grid_.vu, (size_t)H_ * kVelCh * 64 * 4))) return rc;
    if ((rc = dalloc((void**)&grid_.home_list, (size_t)H_ * 4))) return rc;
    if ((rc = dalloc((void**)&grid_.all_list, (size_t)H_ * 4))) return rc;
    HIP_TRY(hipMemsetAsync(grid_.keys, 0xFF, (size_t)H_ * 8, stream_));
    HIP_TRY(hipMemsetAsync(grid_.cellcnt, 0, (size_t)H_ * 64 * 4, stream_));
    HIP_TRY(hipMemsetAsync(grid_.rev, 0xFF, (size_t)H_ * 8 * 4, stream_));

    if ((rc = dalloc((void**)&cnt_, sizeof(Counters)))) return rc;
    HIP_TRY(hipMemsetAsync(cnt_, 0, sizeof(Counters), stream_));
    if ((rc = dalloc((void**)&d_scenes_, sizeof(SceneDev) * kMaxScenes))) return rc;
    HIP_TRY(hipMemsetAsync(d_scenes_, 0, sizeof(SceneDev) * kMaxScenes, stream_));
    if ((rc = dalloc((void**)&st_.scene_status, 4 * kMaxScenes))) return rc;
    if ((rc = dalloc((void**)&st_.fp64_count, 8 * kMaxScenes))) return rc;
    HIP_TRY(hipMemsetAsync(st_.scene_status, 0, 4 * kMaxScenes, stream_));
    HIP_TRY(hipMemsetAsync(st_.fp64_count, 0, 8 * kMaxScenes, stream_));
    if ((rc = dalloc((void**)&d_sums_, sizeof(double) * 7 * 1024))) return rc;
    HIP_TRY(hipStreamSynchronize(stream_));
    h_scenes_.assign(kMaxScenes, SceneDev{});
    scenes_.clear();
    return MPM_OK;
}

StepArgs Engine::args(uint32_t pass_flag, int permute) const {
    StepArgs a;
    std::memset(&a, 0, sizeof a);
    a.cur = view(mov_, con_);
    a.nxt = view(mov_ ^ 1, permute ? con_ ^ 1 : con_);
    a.g = grid_;
    a.bin = bin_;
    a.st = st_;
    a.fb = fb_;
    a.cnt = cnt_;
    a.scenes = d_scenes_;
    a.pass_flag = pass_flag;
    a.mass_only = 0;
    a.copy_E = any_E_ ? 1 : 0;
    a.kappa = (float)cfg_.kappa;
    a.permute = permute;
    a.project = any_project_ ? 1 : 0;
    a.hole_room = (uint32_t)(slots_ - cfg_.particle_capacity);
    a.align_blocks = cfg_.dense_blocks ? 0 : 1;
    if (implicit_active()) a.imp = imp_;
    return a;
}

bool Engine::implicit_active() const {
    for (size_t s = 0; s < scenes_.size(); ++s)
        if (scenes_[s].used && h_scenes_[s].beta > 0.0f) return true;
    return false;
}

int Engine::one_step(int permute, hipEvent_t* ev, mpm_timings* t) {
    if (!implicit_active()) {
        const StepArgs a = args(kSceneActive, permute);
        launch_step(a, launch_, stream_, ev);
        advance(a.permute);
        return MPM_OK;
    }
    int rc = implicit_alloc();
    if (rc) return rc;
    const StepArgs a = args(kSceneActive, permute);
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    struct Events {  // destroyed on every return path
        hipEvent_t* e;
        ~Events() {
            for (int k = 0; k < 3; ++k)
                if (e[k]) (void)hipEventDestroy(e[k]);
        }
    } guard{e};
    if (t)
        for (auto& x : e) HIP_TRY(hipEventCreate(&x));
    launch_step_head(a, launch_, stream_);
    if (t) HIP_TRY(hipEventRecord(e[0], stream_));
    if ((rc = implicit_solve(a))) return rc;
    if (t) HIP_TRY(hipEventRecord(e[1], stream_));
    launch_step_tail(a, launch_, stream_);
    if (t) HIP_TRY(hipEventRecord(e[2], stream_));
    HIP_TRY(hipGetLastError());
    advance(a.permute);
    if (t) {
        HIP_TRY(hipEventSynchronize(e[2]));
        float solve = 0, tail = 0;
        HIP_TRY(hipEventElapsedTime(&solve, e[0], e[1]));
        HIP_TRY(hipEventElapsedTime(&tail, e[1], e[2]));
        t->grid_collisions += solve;  // the solve follows grid_collisions
        t->g2p += tail;
    }
    return MPM_OK;
}

bool Engine::take_permute() {
    const bool p = need_permute_ || (step_count_ % (uint64_t)cfg_.permute_period) == 0;
    need_permute_ = false;
    ++step_count_;
    return p;
}

int Engine::device_count(uint32_t* n) {
    Counters c;
    HIP_TRY(hipMemcpyAsync(&c, cnt_, sizeof c, hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    *n = c.n_particles;
    return MPM_OK;
}

int Engine::upload_scene_table(int sid) {
    HIP_TRY(hipMemcpyAsync(d_scenes_ + sid, &h_scenes_[sid], sizeof(SceneDev), hipMemcpyHostToDevice, stream_));
    return MPM_OK;
}

namespace {
// Stream-ordered staging memory, freed on every return path.
struct StagingBuffer {
    double* p = nullptr;
    hipStream_t s;
    explicit StagingBuffer(hipStream_t st) : s(st) {}
    hipError_t release() {
        const hipError_t e = p ? hipFreeAsync(p, s) : hipSuccess;
        p = nullptr;
        return e;
    }
    ~StagingBuffer() { (void)release(); }
};
}  // namespace

int Engine::scene_add(const mpm_params& prm, const mpm_collider* coll, int ncoll, size_t n, const double* x,
                      const double* v, const double* m, const double* V0, const double* Fe, const double* Fp,
                      const double* E, int* out_sid) {
    if (!x || !v || !m || n == 0 || ncoll < 0 || ncoll > kMaxColliders || (ncoll > 0 && !coll)) {
        set_error("mpm_scene_add: bad arguments");
        return MPM_ERR_ARG;
    }
    if (!(prm.dx > 0) || !(prm.dt > 0)) {
        set_error("mpm_scene_add: dx and dt must be positive");
        return MPM_ERR_ARG;
    }
    if (!(prm.beta >= 0.0) || prm.cr_tolerance < 0.0 || prm.cr_max_iterations < 0 || prm.cr_max_iterations > 256) {
        set_error("mpm_scene_add: beta must be >= 0, cr_tolerance >= 0 and cr_max_iterations in [0, 256]");
        return MPM_ERR_ARG;
    }
    int sid = -1;
    for (int s = 0; s < (int)scenes_.size(); ++s)
        if (!scenes_[s].used && !scenes_[s].pending_drop) {
            sid = s;
            break;
        }
    if (sid < 0) {
        if ((int)scenes_.size() >= kMaxScenes - 1) {
            set_error("scene table full");
            return MPM_ERR_CAPACITY;
        }
        sid = (int)scenes_.size();
        scenes_.push_back(SceneHost{});
    }
    uint32_t n0;
    int rc = device_count(&n0);
    if (rc) return rc;
    size_t live = 0;
    for (const auto& sh : scenes_)
        if (sh.used) live += sh.n;
    if (live + n > cfg_.particle_capacity) {
        set_error("particle capacity exceeded: " + std::to_string(live) + " + " + std::to_string(n) + " > " +
                  std::to_string(cfg_.particle_capacity));
        return MPM_ERR_CAPACITY;
    }
    // The new particles go after the buffer's last slot in use (holes and
    // removed scenes included until the next re-sort compacts them).
    if ((size_t)n0 + n > slots_) {
        set_error("particle slots exceeded: " + std::to_string(n0) + " in use (removed scenes included until the "
                  "next step compacts them) + " + std::to_string(n) + " > " + std::to_string(slots_));
        return MPM_ERR_CAPACITY;
    }
    // The caller's double arrays go to the GPU as they are; a kernel converts
    // them into the FP32 buffers at particles n0 .. n0 + n - 1.
    const double* const host[7] = {x, v, m, V0, Fe, Fp, E};
    const size_t per[7] = {3, 3, 1, 1, 9, 9, 1};
    size_t total = 0;
    for (int k = 0; k < 7; ++k)
        if (host[k]) total += per[k] * n;
    StagingBuffer stage(stream_);
    HIP_TRY(hipMallocAsync((void**)&stage.p, total * sizeof(double), stream_));
    double* const d_in = stage.p;
    const double* dev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t off = 0;
    for (int k = 0; k < 7; ++k)
        if (host[k]) {
            HIP_TRY(hipMemcpyAsync(d_in + off, host[k], per[k] * n * sizeof(double), hipMemcpyHostToDevice, stream_));
            dev[k] = d_in + off;
            off += per[k] * n;
        }
    {
        const StepArgs a = args(kSceneActive);
        launch_put_scene(a, launch_, n0, (uint32_t)n, (uint32_t)sid, SceneIn{dev[0], dev[1], dev[2], dev[3], dev[4], dev[5], dev[6]},
                         stream_);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(stage.release());
    SceneDev& sd = h_scenes_[sid];
    std::memset(&sd, 0, sizeof sd);
    sd.dx = (float)prm.dx;
    sd.inv_dx = (float)(1.0 / prm.dx);
    sd.dt = (float)prm.dt;
    for (int e = 0; e < 3; ++e) sd.gravity[e] = (float)prm.gravity[e];
    sd.mu0 = (float)prm.mu0;
    sd.lambda0 = (float)prm.lambda0;
    sd.nu = (float)prm.nu;
    sd.mu_per_E = (float)(1.0 / (2.0 * (1.0 + prm.nu)));
    sd.la_per_E = (float)(prm.nu / ((1.0 + prm.nu) * (1.0 - 2.0 * prm.nu)));
    sd.hardening = (float)prm.hardening;
    sd.theta_c = (float)prm.max_compression;
    sd.theta_s = (float)prm.max_stretch;
    sd.alpha = (float)prm.alpha;
    sd.beta = (float)prm.beta;
    {
        const double tol = prm.cr_tolerance > 0.0 ? prm.cr_tolerance : 1e-6;
        sd.cr_tol2 = (float)(tol * tol);
    }
    sd.cr_maxit = prm.cr_max_iterations > 0 ? prm.cr_max_iterations : 100;
    sd.use_E = E ? 1 : 0;
    sd.ncoll = ncoll;
    sd.nproject = 0;
    for (int k = 0; k < ncoll; ++k) {
        ColliderDev& c = sd.coll[k];
        c.type = coll[k].type;
        c.flags = coll[k].flags & kProjectFlag;
        sd.nproject += c.flags ? 1 : 0;
        for (int e = 0; e < 3; ++e) {
            c.point[e] = (float)coll[k].point[e];
            c.normal[e] = (float)coll[k].normal[e];
            c.half_extent[e] = (float)coll[k].half_extent[e];
            c.velocity[e] = (float)coll[k].velocity[e];
        }
        c.radius = (float)coll[k].radius;
        c.friction = (float)coll[k].friction;
    }
    sd.flags = V0 ? kSceneActive : kSceneSetup;
    if (E) any_E_ = true;
    if (sd.nproject) any_project_ = true;
    if ((rc = upload_scene_table(sid))) return rc;
    HIP_TRY(hipMemsetAsync(st_.scene_status + sid, 0, 4, stream_));
    HIP_TRY(hipMemsetAsync(st_.fp64_count + sid, 0, 8, stream_));
    launch_set_count(cnt_, n0 + (uint32_t)n, stream_);
    StepArgs a = args(kSceneActive);
    launch_init_stress(a, n0, (uint32_t)n, stream_);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream_));
    SceneHost& sh = scenes_[sid];
    sh = SceneHost{};
    sh.used = true;
    sh.n = n;
    sh.prm = prm;
    *out_sid = sid;
    need_permute_ = true;  // bring the appended particles into sorted order
    return MPM_OK;
}

int Engine::scene_remove(int sid) {
    if (sid < 0 || sid >= (int)scenes_.size() || !scenes_[sid].used) {
        set_error("bad scene id");
        return MPM_ERR_ARG;
    }
    // Particles of a scene with no flags are not binned; the next step is a
    // permuting one, whose G2P writes only binned particles and so compacts
    // the buffer.
    h_scenes_[sid].flags = 0;
    int rc = upload_scene_table(sid);
    if (rc) return rc;
    scenes_[sid].used = false;
    scenes_[sid].pending_drop = true;
    need_permute_ = true;
    return MPM_OK;
}

int Engine::clear() {
    for (size_t s = 0; s < scenes_.size(); ++s) {
        h_scenes_[s].flags = 0;
        int rc = upload_scene_table((int)s);
        if (rc) return rc;
    }
    scenes_.clear();
    any_E_ = false;
    any_project_ = false;
    need_permute_ = true;
    launch_set_count(cnt_, 0, stream_);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream_));
    return MPM_OK;
}

int Engine::setup() {
    bool any = false;
    for (size_t s = 0; s < scenes_.size(); ++s)
        if (scenes_[s].used && (h_scenes_[s].flags & kSceneSetup)) any = true;
    if (!any) return MPM_OK;
    StepArgs a = args(kSceneSetup);
    launch_setup(a, launch_, stream_);
    HIP_TRY(hipGetLastError());
    for (size_t s = 0; s < scenes_.size(); ++s) {
        if (!scenes_[s].used || !(h_scenes_[s].flags & kSceneSetup)) continue;
        SceneDev& sd = h_scenes_[s];
        sd.flags = kSceneActive;
        int rc = upload_scene_table((int)s);
        if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(stream_));
    return check_status();
}

int Engine::check_status() {
    if (scenes_.empty()) return MPM_OK;
    std::vector<uint32_t> st(scenes_.size());
    HIP_TRY(hipMemcpyAsync(st.data(), st_.scene_status, 4 * scenes_.size(), hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    for (size_t s = 0; s < scenes_.size(); ++s)
        if (scenes_[s].used && st[s]) {
            char buf[128];
            std::snprintf(buf, sizeof buf, "scene %zu raised device status 0x%x", s, st[s]);
            set_error(buf);
            return MPM_ERR_DEVICE;
        }
    return MPM_OK;
}

int Engine::step(int nsteps, mpm_timings* t) {
    if (nsteps < 0) {
        set_error("nsteps < 0");
        return MPM_ERR_ARG;
    }
    int rc = setup();
    if (rc) return rc;
    const bool timed = t != nullptr;
    if (timed) {
        // Timing-only events: no system-scope fence (L2 write-back) at each
        // record, which would add its own time between the stages measured.
        // The stream is synchronized before the times are read.
        const size_t need = (size_t)nsteps * kStepEvents;
        while (events_.size() < need) {
            hipEvent_t ev;
            HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableSystemFence));
            events_.push_back(ev);
        }
    }
    if (implicit_active()) {
        // host-driven solve between the halves of each step: no per-kernel
        // event slots; the solve's time is added to timings->grid_collisions
        for (int s = 0; s < nsteps; ++s)
            if ((rc = one_step(take_permute() ? 1 : 0, nullptr, t))) return rc;
        HIP_TRY(hipStreamSynchronize(stream_));
        if (nsteps > 0)
            for (auto& sh : scenes_) sh.pending_drop = false;
        kernel_steps_ = 0;
        return check_status();
    }
    for (int s = 0; s < nsteps; ++s) {
        const StepArgs a = args(kSceneActive, take_permute() ? 1 : 0);
        launch_step(a, launch_, stream_, timed ? &events_[(size_t)s * kStepEvents] : nullptr);
        advance(a.permute);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream_));
    if (nsteps > 0)
        for (auto& sh : scenes_) sh.pending_drop = false;
    if (timed) {
        double k[5] = {0, 0, 0, 0, 0};
        for (int s = 0; s < nsteps; ++s)
            for (int j = 0; j < 5; ++j) {
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, events_[(size_t)s * kStepEvents + j],
                                            events_[(size_t)s * kStepEvents + j + 1]));
                k[j] += ms;
            }
        t->stencil += k[0] + k[4];
        t->p2g += k[1];
        t->grid_velocity += k[2];
        t->g2p += k[3];
        for (int j = 0; j < 5; ++j) kernel_ms_[j] = nsteps ? k[j] / nsteps : 0.0;
        kernel_steps_ = nsteps;
    }
    return check_status();
}

// Bytes of the staging a download of the given outputs needs (0 for none).
int Engine::download_bytes(int sid, const double* const host[6], bool want_fp, size_t* bytes) {
    if (sid < 0 || sid >= (int)scenes_.size() || !scenes_[sid].used) {
        set_error("bad scene id");
        return MPM_ERR_ARG;
    }
    if (want_fp && cfg_.track_fp == 0) {
        set_error("F_P download needs track_fp");
        return MPM_ERR_STATE;
    }
    const size_t per[6] = {3, 3, 9, 9, 1, 1};
    size_t total = 0;
    for (int k = 0; k < 6; ++k)
        if (host[k]) total += per[k] * scenes_[sid].n;
    *bytes = total * sizeof(double);
    return MPM_OK;
}

// The GPU writes the requested arrays into `stage` in the caller's layout
// (doubles, original particle order; zeros for particles a failed pass
// dropped); the host only copies them.  With fill_kernel the zeros come from
// a fill kernel instead of hipMemsetAsync.
int Engine::download_via(int sid, double* const host[6], double* stage, bool fill_kernel) {
    const size_t n = scenes_[sid].n;
    const size_t per[6] = {3, 3, 9, 9, 1, 1};
    size_t total = 0;
    for (int k = 0; k < 6; ++k)
        if (host[k]) total += per[k] * n;
    if (fill_kernel) {
        launch_fill_zero(stage, total, stream_);
        HIP_TRY(hipGetLastError());
    } else {
        HIP_TRY(hipMemsetAsync(stage, 0, total * sizeof(double), stream_));
    }
    double* dev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t off = 0;
    for (int k = 0; k < 6; ++k)
        if (host[k]) {
            dev[k] = stage + off;
            off += per[k] * n;
        }
    const SceneOut o{dev[0], dev[1], dev[2], dev[3], dev[4], dev[5]};
    StepArgs a = args(kSceneActive);
    launch_gather_scene(a, launch_, (uint32_t)sid, (uint32_t)n, o, stream_);
    HIP_TRY(hipGetLastError());
    for (int k = 0; k < 6; ++k)
        if (host[k]) HIP_TRY(hipMemcpyAsync(host[k], dev[k], per[k] * n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    return MPM_OK;
}

int Engine::scene_download(int sid, double* x, double* v, double* Fe, double* Fp, double* Jp, double* V0) {
    double* const host[6] = {x, v, Fe, Fp, Jp, V0};
    size_t bytes = 0;
    int rc = download_bytes(sid, host, Fp != nullptr, &bytes);
    if (rc) return rc;
    if (bytes == 0) return MPM_OK;
    StagingBuffer stage(stream_);
    HIP_TRY(hipMallocAsync((void**)&stage.p, bytes, stream_));
    if ((rc = download_via(sid, host, stage.p, false))) return rc;
    HIP_TRY(stage.release());
    HIP_TRY(hipStreamSynchronize(stream_));
    return MPM_OK;
}

int Engine::scene_download_static(int sid, double* x, double* v, double* Fe, double* Fp, double* Jp, double* V0,
                                  double* scratch, size_t scratch_bytes) {
    double* const host[6] = {x, v, Fe, Fp, Jp, V0};
    size_t bytes = 0;
    int rc = download_bytes(sid, host, Fp != nullptr, &bytes);
    if (rc) return rc;
    if (bytes == 0) return MPM_OK;
    if (bytes > scratch_bytes) {
        set_error("mpm_scene_download_noalloc: the requested arrays need " + std::to_string(bytes) +
                  " bytes of staging, more than the library's " + std::to_string(scratch_bytes) +
                  "-byte static buffer; download fewer fields or use mpm_scene_download");
        return MPM_ERR_CAPACITY;
    }
    if ((rc = download_via(sid, host, scratch, true))) return rc;
    HIP_TRY(hipStreamSynchronize(stream_));
    return MPM_OK;
}

int Engine::scene_status(int sid, uint32_t* status, uint64_t* fb) {
    if (sid < 0 || sid >= (int)scenes_.size()) {
        set_error("bad scene id");
        return MPM_ERR_ARG;
    }
    uint32_t s = 0;
    unsigned long long c = 0;
    HIP_TRY(hipMemcpyAsync(&s, st_.scene_status + sid, 4, hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipMemcpyAsync(&c, st_.fp64_count + sid, 8, hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    if (status) *status = s;
    if (fb) *fb = c;
    return MPM_OK;
}

int Engine::scene_sums(int sid, double out[7]) {
    if (sid < 0 || sid >= (int)scenes_.size() || !scenes_[sid].used) {
        set_error("bad scene id");
        return MPM_ERR_ARG;
    }
    const int nb = 1024;
    StepArgs a = args(kSceneActive);
    launch_scene_sums(a, (uint32_t)sid, d_sums_, nb, stream_);
    std::vector<double> h((size_t)nb * 7);
    HIP_TRY(hipMemcpyAsync(h.data(), d_sums_, h.size() * 8, hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    for (int e = 0; e < 7; ++e) {
        double s = 0;
        for (int b = 0; b < nb; ++b) s += h[(size_t)b * 7 + e];
        out[e] = s;
    }
    return MPM_OK;
}

}  // namespace snowmpm

// ---------------------------------------------------------------------------
// C ABI
using snowmpm::Engine;
using snowmpm::set_error;


extern "C" {

const char* mpm_last_error(void) { return snowmpm::g_last_error.c_str(); }

int mpm_create(const mpm_config* cfg, mpm_engine** out) {
    if (!cfg || !out) {
        set_error("mpm_create: null argument");
        return MPM_ERR_ARG;
    }
    *out = nullptr;
    mpm_engine* e = new (std::nothrow) mpm_engine();
    if (!e) return MPM_ERR_OOM;
    int rc = e->eng.init(*cfg);
    if (rc) {
        delete e;
        return rc;
    }
    *out = e;
    return MPM_OK;
}

int mpm_destroy(mpm_engine* e) {
    if (!e) return MPM_ERR_ARG;
    delete e;
    return MPM_OK;
}

int mpm_scene_add(mpm_engine* e, const mpm_params* prm, const mpm_collider* colliders, int ncolliders, size_t n,
                  const double* x, const double* v, const double* m, const double* V0, const double* Fe,
                  const double* Fp, const double* E, int* scene_id) {
    if (!e || !prm || !scene_id) {
        set_error("mpm_scene_add: null argument");
        return MPM_ERR_ARG;
    }
    return e->eng.scene_add(*prm, colliders, ncolliders, n, x, v, m, V0, Fe, Fp, E, scene_id);
}

int mpm_scene_remove(mpm_engine* e, int scene) { return e ? e->eng.scene_remove(scene) : MPM_ERR_ARG; }

int mpm_clear(mpm_engine* e) { return e ? e->eng.clear() : MPM_ERR_ARG; }

int mpm_scene_count(mpm_engine* e, int scene, size_t* n) {
    if (!e || !n) return MPM_ERR_ARG;
    return e->eng.scene_count(scene, n);
}

int mpm_setup(mpm_engine* e) { return e ? e->eng.setup() : MPM_ERR_ARG; }

int mpm_step(mpm_engine* e, int nsteps, mpm_timings* t) { return e ? e->eng.step(nsteps, t) : MPM_ERR_ARG; }

int mpm_scene_download(mpm_engine* e, int scene, double* x, double* v, double* Fe, double* Fp, double* Jp,
                       double* V0) {
    return e ? e->eng.scene_download(scene, x, v, Fe, Fp, Jp, V0) : MPM_ERR_ARG;
}

int mpm_scene_status(mpm_engine* e, int scene, uint32_t* status, uint64_t* fb) {
    return e ? e->eng.scene_status(scene, status, fb) : MPM_ERR_ARG;
}

int mpm_scene_sums(mpm_engine* e, int scene, double out7[7]) {
    return (e && out7) ? e->eng.scene_sums(scene, out7) : MPM_ERR_ARG;
}

int mpm_scene_steps(mpm_engine* e, int scene, uint64_t* steps) {
    return (e && steps) ? e->eng.scene_steps(scene, steps) : MPM_ERR_ARG;
}

int mpm_kernel_times(mpm_engine* e, double ms[5], int* nsteps) {
    if (!e || !ms) return MPM_ERR_ARG;
    e->eng.kernel_times(ms, nsteps);
    return MPM_OK;
}

int mpm_slots(mpm_engine* e, size_t* used, size_t* total) {
    return (e && used && total) ? e->eng.slots(used, total) : MPM_ERR_ARG;
}

mpm_collider mpm_make_plane(const double point[3], const double normal[3], double friction) {
    mpm_collider c;
    std::memset(&c, 0, sizeof c);
    c.type = MPM_COLLIDER_PLANE;
    const double l = std::sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
    for (int a = 0; a < 3; ++a) {
        c.point[a] = point[a];
        c.normal[a] = normal[a] / l;  // normal.normalized(), collision.cpp:7
    }
    c.friction = friction;
    return c;
}

mpm_collider mpm_make_box(const double centre[3], const double half_extent[3], double friction) {
    mpm_collider c;
    std::memset(&c, 0, sizeof c);
    c.type = MPM_COLLIDER_BOX;
    for (int a = 0; a < 3; ++a) {
        c.point[a] = centre[a];
        c.half_extent[a] = half_extent[a];
    }
    c.friction = friction;
    return c;
}

mpm_collider mpm_make_sphere(const double centre[3], double radius, double friction) {
    mpm_collider c;
    std::memset(&c, 0, sizeof c);
    c.type = MPM_COLLIDER_SPHERE;
    for (int a = 0; a < 3; ++a) c.point[a] = centre[a];
    c.radius = radius;
    c.friction = friction;
    return c;
}

void mpm_free(void* p) { std::free(p); }

}  // extern "C"
