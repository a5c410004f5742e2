// mpm_serve.cpp -- graph-replayed steps, launch autotuning and the
// continuous-batching request loop of the engine (host C++20).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
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

int Engine::graph_begin(bool* host_driven) {
    *host_driven = implicit_active();  // the solve is host-driven: no graph
    return *host_driven ? MPM_OK : setup();
}

int Engine::graph_enqueue() {
    const int permute = take_permute() ? 1 : 0;
    const int rc = ensure_graph(permute);
    if (rc) return rc;
    HIP_TRY(hipGraphLaunch(gexec_[(mov_ * 2 + con_) * 2 + permute], stream_));
    advance(permute);
    return MPM_OK;
}

int Engine::graph_finish(int nsteps) {
    HIP_TRY(hipStreamSynchronize(stream_));
    if (nsteps > 0)
        for (auto& sh : scenes_) sh.pending_drop = false;
    return check_status();
}

int Engine::step_graph(int nsteps) {
    if (nsteps < 0) {
        set_error("nsteps < 0");
        return MPM_ERR_ARG;
    }
    bool host_driven = false;
    int rc = graph_begin(&host_driven);
    if (rc) return rc;
    if (host_driven) return step(nsteps, nullptr);
    for (int s = 0; s < nsteps; ++s)
        if ((rc = graph_enqueue())) return rc;
    return graph_finish(nsteps);
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
            const int r = one_step(take_permute() ? 1 : 0, nullptr, nullptr);
            if (r) return r;
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
    // Tile sizes first (P2G staging chunk, G2P workgroup size), then the
    // persistent grid sizes, each by coordinate descent.
    {
        const int chunks[] = {48, 32, 64}, threads[] = {128, 64, 256};
        for (int k = 0; k < 3; ++k) {
            LaunchCfg c = best;
            c.p2g_chunk = chunks[k];
            if (std::memcmp(&c, &best, sizeof c) == 0) continue;
            double ms = 0;
            if ((rc = run(c, &ms))) return rc;
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
        for (int k = 0; k < 3; ++k) {
            LaunchCfg c = best;
            c.g2p_threads = threads[k];
            if (std::memcmp(&c, &best, sizeof c) == 0) continue;
            double ms = 0;
            if ((rc = run(c, &ms))) return rc;
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
    }
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
    tiles_tuned_ = true;
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

// Device bytes per particle slot (both particle buffers with pid and sid,
// the four sort arrays, the FP64 fallback queue), per grid block (hash key,
// cell counts and offsets, particle range, neighbour and home-block indices,
// v/u, list entries) and per home block (the P2G partial sums of its 7^3
// node region).
static size_t bytes_per_particle(bool track) {
    const size_t nf = 3 + 3 + 9 + 1 + (track ? 9 : 0) + 1 + 1 + 1 + kStressCh;
    return 2 * (nf * 4 + 8) + 4 * 4 + 4 + 36;
}
constexpr size_t kBytesPerBlock = 8 + 2 * 64 * 4 + 4 + 4 + 8 * 4 + 8 * 4 + kVelCh * 64 * 4 + 4 + 4;
constexpr size_t kBytesPerHome = (size_t)kPartFloats * 4;
// Resident particles at which serving switches to the engine's own tile
// sizes, and back to the latency tiles (a gap, so the step graphs are not
// re-captured at every chunk near the threshold).
constexpr size_t kServeThroughputAbove = 4000000, kServeLatencyBelow = 3000000;

Footprint Engine::pools() const {
    return Footprint{cfg_.particle_capacity * bytes_per_particle(cfg_.track_fp != 0), (size_t)(H_ / 2) * kBytesPerBlock,
                     (size_t)grid_.max_home * kBytesPerHome};
}

// Blocks from the bounding box of the positions in 4^3-node blocks (home
// block of x: (floor(x/dx) - 1) div 4; a home block's nodes also reach the +1
// block per axis), with a 25% margin for a body that spreads as it moves.
Footprint Engine::footprint(size_t n, const double* x, double dx) const {
    Footprint f{n * bytes_per_particle(cfg_.track_fp != 0), 0, 0};
    if (n == 0 || !x || !(dx > 0)) return f;
    long lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    bool any = false;
    for (size_t i = 0; i < n; ++i) {
        long b[3];
        bool finite = true;
        for (int a = 0; a < 3; ++a) {
            const double g = x[i * 3 + a] / dx;
            if (!std::isfinite(g) || std::fabs(g) > 1e9) {
                finite = false;
                break;
            }
            const long base = (long)std::floor(g) - 1;
            b[a] = base >= 0 ? base / 4 : -((-base + 3) / 4);
        }
        if (!finite) continue;
        for (int a = 0; a < 3; ++a) {
            lo[a] = any ? std::min(lo[a], b[a]) : b[a];
            hi[a] = any ? std::max(hi[a], b[a]) : b[a];
        }
        any = true;
    }
    if (!any) return f;
    double home = 1.0, all = 1.0;
    for (int a = 0; a < 3; ++a) {
        home *= (double)(hi[a] - lo[a] + 1);
        all *= (double)(hi[a] - lo[a] + 2);
    }
    const double cap = 1e15;
    f.blocks = (size_t)std::min(cap, std::ceil(1.25 * all) + 8.0) * kBytesPerBlock;
    f.partials = (size_t)std::min(cap, std::ceil(1.25 * home) + 8.0) * kBytesPerHome;
    return f;
}

int Engine::serve(mpm_request* R, int nreq, int mode, int max_chunk, mpm_serve_stats* st) {
    if (!R || nreq < 0 || !st) {
        set_error("mpm_serve: bad arguments");
        return MPM_ERR_ARG;
    }
    for (int k = 0; k < nreq; ++k)
        if (R[k].steps < 0 || R[k].ncolliders < 0 || R[k].ncolliders > MPM_MAX_COLLIDERS) {
            set_error("mpm_serve: a request has steps < 0 or more than MPM_MAX_COLLIDERS colliders");
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
    std::vector<Footprint> fp(nreq);
    for (int k = 0; k < nreq; ++k) {
        R[k].status = 0;
        R[k].t_admit_s = R[k].t_done_s = 0.0;
        fp[k] = footprint(R[k].n, R[k].x, R[k].prm.dx);
        R[k].footprint_bytes = fp[k].total();
    }
    const Footprint pool = pools();
    int rc;
    double psteps = 0, max_bytes = 0, t_step = 0, t_admit = 0, t_retire = 0;
    int completed = 0, failed = 0, max_res = 0, chunks = 0, chunk_steps = 0;
    // Any error other than a scene's own status ends the call with the engine
    // emptied, so the caller can serve again.
    auto fail = [&](int code) {
        const std::string msg = g_last_error;
        (void)clear();
        set_error(msg);
        return code;
    };
    auto reject = [&](mpm_request& q) {
        q.status = MPM_STATUS_REJECTED;
        q.t_done_s = now();
        ++failed;
    };

    auto add = [&](mpm_request& q, int* sid) {
        return scene_add(q.prm, q.colliders, q.ncolliders, q.n, q.x, q.v, q.m, nullptr, nullptr, nullptr, q.E, sid);
    };
    // Tile sizes by load (results do not depend on them): while few
    // particles are resident, each step is bound by one home block's
    // latency, and 256-thread G2P workgroups and 64-particle P2G staging
    // chunks shorten it (p50 / p99 at 10-50 requests/s about 12% lower); at
    // saturation the engine's own tiles are faster.  Tiles the caller set or
    // mpm_autotune chose are kept; the engine's are restored on return.
    struct TileGuard {
        LaunchCfg& l;
        const int chunk, threads;
        ~TileGuard() {
            l.p2g_chunk = chunk;
            l.g2p_threads = threads;
        }
    } tile_guard{launch_, launch_.p2g_chunk, launch_.g2p_threads};
    const bool adapt_tiles = !cfg_.p2g_chunk && !cfg_.g2p_threads && !tiles_tuned_;
    bool latency_tiles = true;
    auto choose_tiles = [&](size_t resident) {
        if (!adapt_tiles) return;
        if (latency_tiles && resident > kServeThroughputAbove) latency_tiles = false;
        if (!latency_tiles && resident < kServeLatencyBelow) latency_tiles = true;
        launch_.p2g_chunk = latency_tiles ? 64 : tile_guard.chunk;
        launch_.g2p_threads = latency_tiles ? 256 : tile_guard.threads;
    };
    if (mode == 1) {
        for (int k = 0; k < nreq; ++k) {
            mpm_request& q = R[order[k]];
            wait_until(q.arrival_s);
            int sid;
            if ((rc = clear())) return fail(rc);
            if (q.n > cfg_.particle_capacity) {
                reject(q);
                continue;
            }
            if ((rc = add(q, &sid))) return fail(rc);
            q.t_admit_s = now();
            max_bytes = std::max(max_bytes, (double)fp[order[k]].total());
            choose_tiles(q.n);
            rc = step(q.steps, nullptr);
            if (rc == MPM_ERR_DEVICE) {
                uint32_t bits = 0;
                if ((rc = scene_status(sid, &bits, nullptr))) return fail(rc);
                q.status = bits ? bits : MPM_STATUS_NAN;
                q.t_done_s = now();
                ++failed;
                continue;
            }
            if (rc) return fail(rc);
            if (q.out_x || q.out_v)
                if ((rc = scene_download(sid, q.out_x, q.out_v, nullptr, nullptr, nullptr, nullptr))) return fail(rc);
            q.t_done_s = now();
            psteps += (double)q.n * q.steps;
            ++completed;
            max_res = 1;
        }
        if ((rc = clear())) return fail(rc);
    } else {
        struct Act {
            int req, sid, left;
        };
        std::vector<Act> act;
        size_t next = 0;
        // bytes reserved in each pool by resident requests, and particle
        // slots of retired scenes not yet compacted away
        Footprint res{0, 0, 0};
        size_t pend = 0;
        auto retire = [&](size_t i) {
            const Footprint& f = fp[act[i].req];
            res.particles -= f.particles;
            res.blocks -= f.blocks;
            res.partials -= f.partials;
            pend += f.particles;
            act.erase(act.begin() + (long)i);
        };
        while (completed + failed < nreq) {
            const double tn = now();
            double ta = tn;
            while (next < (size_t)nreq && R[order[next]].arrival_s <= tn) {
                const int r = order[next];
                mpm_request& q = R[r];
                const Footprint& f = fp[r];
                if (f.particles > pool.particles) {  // can never fit
                    reject(q);
                    ++next;
                    continue;
                }
                bool fits = res.particles + pend + f.particles <= pool.particles &&
                            res.blocks + f.blocks <= pool.blocks && res.partials + f.partials <= pool.partials;
                if (!fits && act.empty()) {
                    // Nothing resident: drop the retired scenes' particles now and
                    // admit the request alone (its block estimate is an upper bound).
                    if ((rc = clear())) return fail(rc);
                    pend = 0;
                    fits = true;
                }
                if (!fits) break;
                int sid;
                if ((rc = add(q, &sid))) return fail(rc);
                q.t_admit_s = now();
                act.push_back({r, sid, q.steps});
                res.particles += f.particles;
                res.blocks += f.blocks;
                res.partials += f.partials;
                max_bytes = std::max(max_bytes, (double)res.total());
                ++next;
            }
            const double ts = now();
            t_admit += ts - ta;
            if (act.empty()) {
                if (next < (size_t)nreq) wait_until(R[order[next]].arrival_s);
                continue;
            }
            max_res = std::max(max_res, (int)act.size());
            int K = max_chunk;
            size_t resident = 0;
            for (const auto& a : act) {
                K = std::min(K, a.left);
                resident += R[a.req].n;
            }
            choose_tiles(resident);
            rc = step_graph(K);
            if (rc && rc != MPM_ERR_DEVICE) return fail(rc);
            const double tr = now();
            t_step += tr - ts;
            ++chunks;
            chunk_steps += K;
            pend = 0;
            size_t res_n = 0;
            for (const auto& a : act) res_n += R[a.req].n;
            psteps += (double)K * (double)res_n;
            for (size_t i = 0; i < act.size();) {
                mpm_request& q = R[act[i].req];
                uint32_t bits = 0;
                if (rc == MPM_ERR_DEVICE) {
                    int rs = scene_status(act[i].sid, &bits, nullptr);
                    if (rs) return fail(rs);
                }
                act[i].left -= K;
                if (bits) {  // this request's scene failed: retire it, keep the others
                    q.status = bits;
                    if ((rc = scene_remove(act[i].sid))) return fail(rc);
                    q.t_done_s = now();
                    ++failed;
                    retire(i);
                    continue;
                }
                if (act[i].left > 0) {
                    ++i;
                    continue;
                }
                if (q.out_x || q.out_v)
                    if ((rc = scene_download(act[i].sid, q.out_x, q.out_v, nullptr, nullptr, nullptr, nullptr)))
                        return fail(rc);
                if ((rc = scene_remove(act[i].sid))) return fail(rc);
                q.t_done_s = now();
                ++completed;
                retire(i);
            }
            t_retire += now() - tr;
        }
        if ((rc = clear())) return fail(rc);
    }
    st->wall_s = now();
    st->completed = completed;
    st->failed = failed;
    st->max_resident = max_res;
    st->max_resident_bytes = max_bytes;
    st->time_step_s = t_step;
    st->time_admit_s = t_admit;
    st->time_retire_s = t_retire;
    st->chunks = chunks;
    st->chunk_steps = chunk_steps;
    st->particle_steps = psteps;
    st->scenes_per_s = st->wall_s > 0 ? completed / st->wall_s : 0.0;
    std::vector<double> lat;
    for (int k = 0; k < nreq; ++k)
        if (R[k].status == 0) lat.push_back(R[k].t_done_s - R[k].arrival_s);
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

int mpm_step_graph_multi(mpm_engine* const* es, int n, int nsteps) {
    using snowmpm::set_error;
    if (!es || n < 1 || n > 64 || nsteps < 0) {
        set_error("mpm_step_graph_multi: need 1..64 engines and nsteps >= 0");
        return MPM_ERR_ARG;
    }
    for (int k = 0; k < n; ++k) {
        if (!es[k] || es[k]->eng.device() != es[0]->eng.device()) {
            set_error("mpm_step_graph_multi: null engine, or engines on different devices");
            return MPM_ERR_ARG;
        }
        for (int j = 0; j < k; ++j)
            if (es[j] == es[k]) {
                set_error("mpm_step_graph_multi: an engine is listed twice");
                return MPM_ERR_ARG;
            }
    }
    // The first error is returned (its message prefixed with the engine's
    // index); every engine that began is still synchronised and checked.
    int first = MPM_OK;
    std::string msg;
    auto note = [&](int k, int rc) {
        if (rc && !first) {
            first = rc;
            msg = "engine " + std::to_string(k) + ": " + snowmpm::g_last_error;
        }
    };
    std::vector<char> host(n, 0), ok(n, 0);
    for (int k = 0; k < n; ++k) {
        bool hd = false;
        const int rc = es[k]->eng.graph_begin(&hd);
        note(k, rc);
        host[k] = hd;
        ok[k] = rc == MPM_OK;
    }
    // one step of every graph engine in turn, each on its own stream
    for (int s = 0; s < nsteps; ++s)
        for (int k = 0; k < n; ++k)
            if (ok[k] && !host[k]) {
                const int rc = es[k]->eng.graph_enqueue();
                note(k, rc);
                if (rc) ok[k] = 0;
            }
    // engines with a semi-implicit scene run their host-driven steps while
    // the graph engines' launches proceed on their streams
    for (int k = 0; k < n; ++k)
        if (ok[k] && host[k]) note(k, es[k]->eng.step(nsteps, nullptr));
    for (int k = 0; k < n; ++k)
        if (!host[k]) note(k, es[k]->eng.graph_finish(nsteps));
    if (first) set_error(msg);
    return first;
}

int mpm_autotune(mpm_engine* e, int steps_per_config, int chosen[4]) {
    return e ? e->eng.autotune(steps_per_config, chosen) : MPM_ERR_ARG;
}

int mpm_tile_sizes(mpm_engine* e, int* p2g_chunk, int* g2p_threads) {
    if (!e || !p2g_chunk || !g2p_threads) return MPM_ERR_ARG;
    e->eng.tile_sizes(p2g_chunk, g2p_threads);
    return MPM_OK;
}

int mpm_permute_period(mpm_engine* e, int* period) {
    if (!e || !period) return MPM_ERR_ARG;
    *period = e->eng.permute_period();
    return MPM_OK;
}

int mpm_serve(mpm_engine* e, mpm_request* reqs, int nreq, int mode, int max_chunk, mpm_serve_stats* stats) {
    return e ? e->eng.serve(reqs, nreq, mode, max_chunk, stats) : MPM_ERR_ARG;
}

int mpm_request_footprint(mpm_engine* e, const mpm_request* q, size_t* bytes, size_t pools[3]) {
    if (!e || !q || !bytes) return MPM_ERR_ARG;
    *bytes = e->eng.footprint(q->n, q->x, q->prm.dx).total();
    if (pools) {
        const snowmpm::Footprint p = e->eng.pools();
        pools[0] = p.particles;
        pools[1] = p.blocks;
        pools[2] = p.partials;
    }
    return MPM_OK;
}

}  // extern "C"
