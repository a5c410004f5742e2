// mpm_serve.cpp -- graph-replayed steps, launch autotuning and the
// continuous-batching request loop of the engine (host C++20).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "../../include/snowmpm.h"
#include "mpm_engine.hpp"

namespace snowmpm {

// One captured graph per (buffer parities, permuting or not): the kernels read
// every count from device memory, so a graph stays valid while scenes come and
// go; it is re-captured only when the launch arguments or shapes change.
int Engine::ensure_graph(int permute) {
    const int k = (mov_ * 2 + con_) * 2 + permute;
    const StepArgs a = args(kSceneActive, permute);
    if (gexec_[k] && std::memcmp(&a, &gargs_[k], sizeof a) == 0 &&
        std::memcmp(&launch_, &glaunch_[k], sizeof launch_) == 0)
        return MPM_OK;
    if (gexec_[k]) {
        HIP_TRY(hipGraphExecDestroy(gexec_[k]));
        gexec_[k] = nullptr;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    launch_step(a, launch_, stream_, nullptr);
    HIP_TRY(hipStreamEndCapture(stream_, &g));
    HIP_TRY(hipGraphInstantiate(&gexec_[k], g, nullptr, nullptr, 0));
    HIP_TRY(hipGraphDestroy(g));
    gargs_[k] = a;
    glaunch_[k] = launch_;
    return MPM_OK;
}

int Engine::step_graph(int nsteps) {
    if (nsteps < 0) {
        set_error("nsteps < 0");
        return MPM_ERR_ARG;
    }
    int rc = setup();
    if (rc) return rc;
    for (int s = 0; s < nsteps; ++s) {
        const int permute = take_permute() ? 1 : 0;
        if ((rc = ensure_graph(permute))) return rc;
        HIP_TRY(hipGraphLaunch(gexec_[(mov_ * 2 + con_) * 2 + permute], stream_));
        advance(permute);
    }
    HIP_TRY(hipStreamSynchronize(stream_));
    if (nsteps > 0)
        for (auto& sh : scenes_) sh.pending_drop = false;
    return check_status();
}

int Engine::autotune(int spc, int chosen[4]) {
    if (spc < 1) spc = 1;
    int rc = setup();
    if (rc) return rc;
    struct Events {  // destroyed on every return path
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Events() {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } ev;
    HIP_TRY(hipEventCreate(&ev.e0));
    HIP_TRY(hipEventCreate(&ev.e1));
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    auto run_n = [&](const LaunchCfg& c, int nsteps, double* ms) -> int {
        launch_ = c;
        HIP_TRY(hipEventRecord(e0, stream_));
        for (int s = 0; s < nsteps; ++s) {
            const StepArgs a = args(kSceneActive, take_permute() ? 1 : 0);
            launch_step(a, launch_, stream_, nullptr);
            advance(a.permute);
        }
        HIP_TRY(hipEventRecord(e1, stream_));
        HIP_TRY(hipEventSynchronize(e1));
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        *ms = t / nsteps;
        return MPM_OK;
    };
    auto run = [&](const LaunchCfg& c, double* ms) { return run_n(c, spc, ms); };
    LaunchCfg best = launch_;
    double best_ms = 0;
    if ((rc = run(best, &best_ms))) return rc;  // warm-up
    if ((rc = run(best, &best_ms))) return rc;
    const int maxp = (int)std::min<size_t>((cfg_.particle_capacity + 255) / 256, (size_t)cus_ * 64);
    const int mult_w[] = {2, 4, 8, 16}, mult_p[] = {4, 8, 16, 32}, mult_b[] = {2, 4, 8, 16}, mult_k[] = {1, 2, 4};
    for (int field = 0; field < 4; ++field) {
        const int* mult = field == 0 ? mult_w : field == 1 ? mult_p : field == 2 ? mult_b : mult_k;
        const int nm = field == 3 ? 3 : 4;
        for (int k = 0; k < nm; ++k) {
            LaunchCfg c = best;
            int v = cus_ * mult[k];
            if (field == 0) c.grid_w = v;
            if (field == 1) c.grid_p = std::max(1, std::min(v, maxp));
            if (field == 2) c.grid_b = v;
            if (field == 3) c.grid_k = v;
            if (std::memcmp(&c, &best, sizeof c) == 0) continue;
            double ms = 0;
            if ((rc = run(c, &ms))) return rc;
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
    }
    // The re-sort period (results do not depend on it either): scenes whose
    // particles change blocks quickly gain from re-sorting more often.  Each
    // candidate is timed over 16 steps so that every one includes its re-sorts.
    {
        const int periods[] = {16, 8, 4, 2};
        int best_period = cfg_.permute_period;
        double best_pms = 0;
        for (int k = 0; k < 4; ++k) {
            cfg_.permute_period = periods[k];
            double ms = 0;
            if ((rc = run_n(best, 16, &ms))) return rc;
            if (k == 0 || ms < best_pms) {
                best_pms = ms;
                best_period = periods[k];
            }
        }
        cfg_.permute_period = best_period;
    }
    launch_ = best;
    drop_graphs();
    if (chosen) {
        chosen[0] = best.grid_p;
        chosen[1] = best.grid_b;
        chosen[2] = best.grid_w;
        chosen[3] = best.grid_k;
    }
    if (spc > 0)
        for (auto& sh : scenes_) sh.pending_drop = false;
    return check_status();
}

// Blocks a scene of n particles may occupy (home + neighbour blocks): a
// conservative bound for compact bodies at 8 particles per cell.
double Engine::footprint_blocks(size_t n) const { return 32.0 + 3.0 * (double)n / 256.0; }

int Engine::serve(mpm_request* R, int nreq, int mode, int max_chunk, mpm_serve_stats* st) {
    if (!R || nreq < 0 || !st) {
        set_error("mpm_serve: bad arguments");
        return MPM_ERR_ARG;
    }
    for (int k = 0; k < nreq; ++k)
        if (R[k].steps < 0 || R[k].ncolliders < 0 ||
            R[k].ncolliders > (int)(sizeof R[k].colliders / sizeof R[k].colliders[0])) {
            set_error("mpm_serve: a request has steps < 0 or more colliders than mpm_request holds");
            return MPM_ERR_ARG;
        }
    if (max_chunk <= 0) max_chunk = 25;
    std::memset(st, 0, sizeof *st);
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    auto wait_until = [&](double t) {
        for (;;) {
            const double d = t - now();
            if (d <= 0) return;
            if (d > 2e-3)
                std::this_thread::sleep_for(std::chrono::duration<double>(d - 1e-3));
            else
                std::this_thread::yield();
        }
    };
    std::vector<int> order(nreq);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return R[a].arrival_s < R[b].arrival_s; });
    int rc;
    double psteps = 0;
    int completed = 0, max_res = 0;

    auto add = [&](mpm_request& q, int* sid) {
        return scene_add(q.prm, q.colliders, q.ncolliders, q.n, q.x, q.v, q.m, nullptr, nullptr, nullptr, q.E, sid);
    };
    if (mode == 1) {
        for (int k = 0; k < nreq; ++k) {
            mpm_request& q = R[order[k]];
            wait_until(q.arrival_s);
            int sid;
            if ((rc = clear())) return rc;
            if ((rc = add(q, &sid))) return rc;
            q.t_admit_s = now();
            if ((rc = step(q.steps, nullptr))) return rc;
            if (q.out_x || q.out_v)
                if ((rc = scene_download(sid, q.out_x, q.out_v, nullptr, nullptr, nullptr, nullptr))) return rc;
            q.t_done_s = now();
            psteps += (double)q.n * q.steps;
            ++completed;
            max_res = 1;
        }
        if ((rc = clear())) return rc;
    } else {
        struct Act {
            int req, sid, left;
        };
        std::vector<Act> act;
        size_t next = 0;
        size_t res_n = 0, pend_n = 0;
        double res_blocks = 0;
        const double block_budget = 0.5 * (double)H_;
        while (completed < nreq) {
            const double tn = now();
            while (next < (size_t)nreq && R[order[next]].arrival_s <= tn) {
                mpm_request& q = R[order[next]];
                const double fb = footprint_blocks(q.n);
                bool fits = res_n + pend_n + q.n <= cfg_.particle_capacity &&
                            (act.empty() || res_blocks + fb <= block_budget);
                if (!fits && act.empty()) {
                    // Nothing resident: drop the retired scenes' particles now.
                    if ((rc = clear())) return rc;
                    pend_n = 0;
                    fits = q.n <= cfg_.particle_capacity;
                    if (!fits) {
                        set_error("mpm_serve: a request is larger than the particle capacity");
                        return MPM_ERR_CAPACITY;
                    }
                }
                if (!fits) break;
                int sid;
                if ((rc = add(q, &sid))) return rc;
                q.t_admit_s = now();
                act.push_back({order[next], sid, q.steps});
                res_n += q.n;
                res_blocks += fb;
                ++next;
            }
            if (act.empty()) {
                if (next < (size_t)nreq) wait_until(R[order[next]].arrival_s);
                continue;
            }
            max_res = std::max(max_res, (int)act.size());
            int K = max_chunk;
            for (const auto& a : act) K = std::min(K, a.left);
            if ((rc = step_graph(K))) return rc;
            pend_n = 0;
            psteps += (double)K * (double)res_n;
            for (size_t i = 0; i < act.size();) {
                act[i].left -= K;
                if (act[i].left > 0) {
                    ++i;
                    continue;
                }
                mpm_request& q = R[act[i].req];
                if (q.out_x || q.out_v)
                    if ((rc = scene_download(act[i].sid, q.out_x, q.out_v, nullptr, nullptr, nullptr, nullptr)))
                        return rc;
                if ((rc = scene_remove(act[i].sid))) return rc;
                q.t_done_s = now();
                res_n -= q.n;
                pend_n += q.n;
                res_blocks -= footprint_blocks(q.n);
                ++completed;
                act.erase(act.begin() + (long)i);
            }
        }
        if ((rc = clear())) return rc;
    }
    st->wall_s = now();
    st->completed = completed;
    st->max_resident = max_res;
    st->particle_steps = psteps;
    st->scenes_per_s = st->wall_s > 0 ? completed / st->wall_s : 0.0;
    std::vector<double> lat;
    for (int k = 0; k < nreq; ++k) lat.push_back(R[k].t_done_s - R[k].arrival_s);
    std::sort(lat.begin(), lat.end());
    if (!lat.empty()) {
        auto pct = [&](double p) {
            const double pos = p * (double)(lat.size() - 1);
            const size_t i = (size_t)pos;
            const double f = pos - (double)i;
            return i + 1 < lat.size() ? lat[i] * (1 - f) + lat[i + 1] * f : lat[i];
        };
        st->p50_s = pct(0.50);
        st->p99_s = pct(0.99);
        st->max_s = lat.back();
        st->mean_s = std::accumulate(lat.begin(), lat.end(), 0.0) / (double)lat.size();
    }
    return MPM_OK;
}

}  // namespace snowmpm


extern "C" {

int mpm_step_graph(mpm_engine* e, int nsteps) { return e ? e->eng.step_graph(nsteps) : MPM_ERR_ARG; }

int mpm_autotune(mpm_engine* e, int steps_per_config, int chosen[4]) {
    return e ? e->eng.autotune(steps_per_config, chosen) : MPM_ERR_ARG;
}

int mpm_permute_period(mpm_engine* e, int* period) {
    if (!e || !period) return MPM_ERR_ARG;
    *period = e->eng.permute_period();
    return MPM_OK;
}

int mpm_serve(mpm_engine* e, mpm_request* reqs, int nreq, int mode, int max_chunk, mpm_serve_stats* stats) {
    return e ? e->eng.serve(reqs, nreq, mode, max_chunk, stats) : MPM_ERR_ARG;
}

}  // extern "C"
