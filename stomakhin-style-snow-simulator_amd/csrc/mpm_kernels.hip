// mpm_kernels.hip -- hand-written CDNA4 (gfx950) kernels of the MPM step.
//
// One step is the reference's step() (src/grid.cpp:359-450) restructured for
// the GPU:
//   k_bin       home block of each particle -> device hash, per-cell counts
//               (and the per-pass counters are reset)
//   k_list      compact the occupied hash slots into the home-block list
//   k_blocks    per home block: cell prefix sums, particle range, neighbour
//               blocks (appended to the list of all blocks)
//   k_scatter   counting-sort scatter of particle indices by (block, cell)
//   k_cellsort  order each cell by particle id (makes P2G order canonical)
//   k_p2g       P2G of mass, momentum and stress force (grid.cpp:119-129,
//               162-182), LDS staged; each home block writes the partial
//               sums of the 7^3 nodes its particles reach (no atomics)
//   k_grid      sums the <= 8 partials of each node in a fixed order, then
//               grid velocity + grid collisions (grid.cpp:195-243) fused
//   k_g2p       stress_update + G2P + particle collisions + advection
//               (grid.cpp:245-338) fused, also producing the next step's
//               force matrix -V0 P F_E^T so P2G needs no SVD
//   k_end       FP64 redo of the flagged plasticity updates, clear the hash
//               slots used this pass, publish the particle count
#include <hip/hip_runtime.h>

#include "mpm_device.hpp"
#include "mpm_kernels.hpp"
#include "mpm_math.hpp"

namespace snowmpm {

// ---------------------------------------------------------------------------
// Device hash (open addressing, linear probing, CAS insert).  The slot index
// is the block's storage index.
__device__ __forceinline__ uint32_t hash_slot(uint64_t k, uint32_t mask) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return (uint32_t)k & mask;
}

__device__ __forceinline__ uint32_t hash_insert(const GridDev& g, uint64_t key, bool& inserted) {
    uint32_t h = hash_slot(key, g.mask);
    inserted = false;
    for (uint32_t p = 0; p < g.max_probe; ++p) {
        uint64_t cur = __hip_atomic_load(&g.keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == key) return h;
        if (cur == kEmptyKey) {
            unsigned long long old = atomicCAS((unsigned long long*)&g.keys[h], (unsigned long long)kEmptyKey,
                                               (unsigned long long)key);
            if (old == kEmptyKey) {
                inserted = true;
                return h;
            }
            if (old == key) return h;
        }
        h = (h + 1) & g.mask;
    }
    return kInvalid;
}

__device__ __forceinline__ void flag(const StatusDev& st, uint32_t sid, uint32_t bits) {
    atomicOr(&st.scene_status[sid], bits);
}

__device__ __forceinline__ bool finite3(float a, float b, float c) {
    return isfinite(a) && isfinite(b) && isfinite(c);
}

// ---------------------------------------------------------------------------

__global__ void k_set_count(Counters* c, uint32_t n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) c->n_particles = n;
}

// Per particle: hash-insert the home block, count the particle into its cell.
// Consecutive particles mostly share a block and a cell (the buffer is sorted
// by the previous step), so both are aggregated per wave over runs of equal
// keys in consecutive lanes: one lane of each run inserts the block key, one
// lane of each run bumps the cell counter by the run's length, and the run's
// lanes take consecutive ranks.  A key that appears in several runs (particles
// out of sorted order) is inserted or counted once per run, which is still
// exact.  A wave takes two sets of 64 consecutive particles per pass, whose
// hash lookups and atomics are in flight together (set 0's runs are served by
// their first lane, set 1's by their last, so a lane seldom serves both), and
// the positions of the next pass are loaded meanwhile.
struct BinItem {
    uint64_t key;
    uint32_t l, sid;
    bool want;
};
// The scene fields binning needs, cached per wave for the scene of the
// wave's first particle (nearly every particle of a wave, the buffer being
// sorted by block).  Loaded from global memory after the previous pass's
// stores, they would make every pass wait for those stores' acknowledgement.
struct SceneCache {
    uint32_t sid;
    int flags;
    float inv_dx;
};
__device__ __forceinline__ BinItem bin_item(const StepArgs& a, const SceneCache& sc_c, uint32_t i, uint32_t n,
                                            uint32_t sid, float x0, float x1, float x2) {
    BinItem it{kEmptyKey, 0u, sid, false};
    if (i >= n || sid >= (uint32_t)kMaxScenes) return it;  // past the end, or a hole a re-sort left
    int flags;
    float inv_dx;
    if (sid == sc_c.sid) {
        flags = sc_c.flags;
        inv_dx = sc_c.inv_dx;
    } else {
        flags = a.scenes[sid].flags;
        inv_dx = a.scenes[sid].inv_dx;
    }
    if (!(flags & a.pass_flag)) return it;
    if (!finite3(x0, x1, x2)) {
        flag(a.st, sid, kStatusNaN);
        return it;
    }
    int b[3];
    float f[3];
    base_node(inv_dx, x0, x1, x2, b, f);
    const int bx = b[0] >> 2, by = b[1] >> 2, bz = b[2] >> 2;
    const int lim = kCoordBias - 2;
    if (bx < -lim || bx > lim || by < -lim || by > lim || bz < -lim || bz > lim) {
        flag(a.st, sid, kStatusRange);
        return it;
    }
    it.l = (uint32_t)(b[0] & 3) | ((uint32_t)(b[1] & 3) << 2) | ((uint32_t)(b[2] & 3) << 4);
    it.key = pack_key(sid, bx, by, bz);
    it.want = true;
    return it;
}
// Runs of equal keys among the lanes with want set.  heads: the lanes that
// differ from the lane below (or whose lane below has no key), tails: the
// lanes whose lane above starts a run or has no key.  Per lane, the first
// (hl) and last (tl) lane of its run; le / ge are the masks of the lanes at
// or below / at or above this lane.
struct Run {
    int hl, tl;
};
__device__ __forceinline__ int lowest_lane(uint64_t m) { return __ffsll((unsigned long long)m) - 1; }
__device__ __forceinline__ int highest_lane(uint64_t m) { return 63 - __clzll((long long)m); }
__device__ __forceinline__ Run bin_runs(bool head, bool want, uint64_t le, uint64_t ge) {
    const uint64_t H = __ballot(head), W = __ballot(want);
    const uint64_t T = W & ((H >> 1) | ~(W >> 1));
    return Run{highest_lane(H & le), lowest_lane(T & ge)};
}
__device__ __forceinline__ uint32_t lane_below(uint32_t v) { return __shfl_up(v, 1, 64); }

__global__ __launch_bounds__(256) void k_bin(StepArgs a) {
    const uint32_t n = a.cnt->n_particles;
    if (threadIdx.x == 0 && blockIdx.x == 0) {  // per-pass counters (no kernel of this pass has read them yet)
        a.cnt->n_home = 0;
        a.cnt->n_extra = 0;
        a.cnt->cursor = 0;
        a.cnt->n_fallback = 0;
    }
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t le = lane == 63 ? ~0ull : (2ull << lane) - 1ull, ge = ~((1ull << lane) - 1ull);
    const uint32_t stride = gridDim.x * blockDim.x * 2;
    uint32_t wbase = (blockIdx.x * blockDim.x + (threadIdx.x & ~63u)) * 2;
    uint32_t nsid[2] = {0u, 0u};
    float nx[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    auto load = [&](uint32_t base) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const uint32_t i = base + 64 * s + lane;
            if (i < n) {
                nsid[s] = a.cur.sid[i];
                nx[s][0] = comp(a.cur.x, a.cur, 0, i);
                nx[s][1] = comp(a.cur.x, a.cur, 1, i);
                nx[s][2] = comp(a.cur.x, a.cur, 2, i);
            }
        }
    };
    load(wbase);
    SceneCache scc;
    scc.sid = (uint32_t)__builtin_amdgcn_readfirstlane((int)nsid[0]);
    if (scc.sid >= (uint32_t)kMaxScenes) scc.sid = 0u;  // a hole: any scene serves as the cache
    scc.flags = a.scenes[scc.sid].flags;
    scc.inv_dx = a.scenes[scc.sid].inv_dx;
    for (; wbase < n; wbase += stride) {
        BinItem it[2];
#pragma unroll
        for (int s = 0; s < 2; ++s)
            it[s] = bin_item(a, scc, wbase + 64 * s + lane, n, nsid[s], nx[s][0], nx[s][1], nx[s][2]);
        load(wbase + stride);
        // block keys: one insert per run (set 0 by its first lane, set 1 by its last)
        Run kr[2];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const uint32_t plo = lane_below((uint32_t)it[s].key), phi = lane_below((uint32_t)(it[s].key >> 32));
            const bool head = it[s].want && (lane == 0 || (((uint64_t)phi << 32) | plo) != it[s].key);
            kr[s] = bin_runs(head, it[s].want, le, ge);
        }
        const int kl0 = kr[0].hl, kl1 = kr[1].tl;
        const bool do0 = it[0].want && kl0 == (int)lane, do1 = it[1].want && kl1 == (int)lane;
        uint32_t s0 = kInvalid, s1 = kInvalid;
        bool ins;
        if (do0 && do1) {
            s0 = hash_insert(a.g, it[0].key, ins);
            s1 = hash_insert(a.g, it[1].key, ins);
        } else if (do0 || do1) {
            const uint32_t sl = hash_insert(a.g, do0 ? it[0].key : it[1].key, ins);
            (do0 ? s0 : s1) = sl;
        }
        if (do0 && s0 == kInvalid) flag(a.st, it[0].sid, kStatusHashFull);
        if (do1 && s1 == kInvalid) flag(a.st, it[1].sid, kStatusHashFull);
        const uint32_t slot0 = __shfl(s0, it[0].want ? kl0 : (int)lane);
        const uint32_t slot1 = __shfl(s1, it[1].want ? kl1 : (int)lane);
        // cells: one counter bump per run of equal cell keys, by the run's length
        const bool h0 = it[0].want && slot0 != kInvalid, h1 = it[1].want && slot1 != kInvalid;
        const uint32_t ck0 = h0 ? (slot0 << 6) | it[0].l : kInvalid;
        const uint32_t ck1 = h1 ? (slot1 << 6) | it[1].l : kInvalid;
        const uint32_t pc0 = lane_below(ck0), pc1 = lane_below(ck1);
        const Run cr0 = bin_runs(h0 && (lane == 0 || pc0 != ck0), h0, le, ge);
        const Run cr1 = bin_runs(h1 && (lane == 0 || pc1 != ck1), h1, le, ge);
        const int cl0 = cr0.hl, cl1 = cr1.tl;
        const bool c0 = h0 && cl0 == (int)lane, c1 = h1 && cl1 == (int)lane;
        uint32_t b0 = 0u, b1 = 0u;
        if (c0 && c1) {
            b0 = atomicAdd(&a.g.cellcnt[ck0], (uint32_t)(cr0.tl - cr0.hl + 1));
            b1 = atomicAdd(&a.g.cellcnt[ck1], (uint32_t)(cr1.tl - cr1.hl + 1));
        } else if (c0 || c1) {
            const uint32_t b = atomicAdd(&a.g.cellcnt[c0 ? ck0 : ck1],
                                         (uint32_t)(c0 ? cr0.tl - cr0.hl + 1 : cr1.tl - cr1.hl + 1));
            (c0 ? b0 : b1) = b;
        }
        b0 = __shfl(b0, h0 ? cl0 : (int)lane);
        b1 = __shfl(b1, h1 ? cl1 : (int)lane);
        const uint32_t i0 = wbase + lane, i1 = wbase + 64 + lane;
        if (i0 < n) {
            a.bin.pkey[i0] = ck0;
            a.bin.prank[i0] = h0 ? b0 + (lane - (uint32_t)cr0.hl) : 0u;
        }
        if (i1 < n) {
            a.bin.pkey[i1] = ck1;
            a.bin.prank[i1] = h1 ? b1 + (lane - (uint32_t)cr1.hl) : 0u;
        }
    }
}

// Compacts the occupied hash slots into the list of home blocks (after
// k_bin every occupied slot is a home block).  One atomic per 4096-slot tile.
__global__ __launch_bounds__(1024) void k_list(StepArgs a) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t base_s;
    uint32_t* list = a.g.home_list;
    uint32_t* counter = &a.cnt->n_home;
    const uint32_t H = a.g.mask + 1;
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (uint32_t tile = blockIdx.x * 4096u; tile < H; tile += gridDim.x * 4096u) {
        const uint32_t s0 = tile + threadIdx.x * 4;
        uint32_t occ = 0, cnt = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (s0 + k < H && a.g.keys[s0 + k] != kEmptyKey) {
                occ |= 1u << k;
                ++cnt;
            }
        uint32_t incl = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, 64);
            if (lane >= (uint32_t)o) incl += t;
        }
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t run = 0;
            for (int k = 0; k < 16; ++k) {
                const uint32_t v = wsum[k];
                wsum[k] = run;
                run += v;
            }
            base_s = run ? atomicAdd(counter, run) : 0u;
        }
        __syncthreads();
        uint32_t pos = base_s + wsum[wid] + incl - cnt;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (occ & (1u << k)) list[pos++] = s0 + k;
        __syncthreads();
    }
}

// Sorted ranges of home blocks start on 32-particle (128-byte) boundaries
// when the holes fit: at most 31 per home block, within the slots the
// buffers hold beyond the capacity (hole_room), so the binned particles
// (at most the capacity) and their holes fit the slots, and the holes a
// re-sort leaves never take room an added scene within the capacity needs.
// Then every range G2P, k_scatter and k_cellsort write covers whole lines of
// each field, not parts of lines that a neighbouring range's workgroup also
// writes; after a re-sort the buffers keep that layout, with holes (scene id
// kInvalid) between the blocks.  The decision is per pass and the same in
// every kernel that asks.
__device__ __forceinline__ bool block_ranges_aligned(const StepArgs& a, uint32_t nh) {
    return a.align_blocks && 31ull * nh <= (uint64_t)a.hole_room;
}

// Per home block (one wave): exclusive scan of the 64 cell counts, particle
// range, and the 8 blocks its nodes live in (the block itself + {0,1}^3).
// A workgroup of 16 waves takes one particle range for its 16 blocks.  The
// list of all blocks is the home list followed by the neighbour blocks this
// kernel inserts (each appended by the lane whose insert created it).
__global__ __launch_bounds__(1024) void k_blocks(StepArgs a) {
    __shared__ uint32_t tot[16], ext[16];
    __shared__ uint32_t base_s, ext_s;
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const uint32_t nh = a.cnt->n_home;
    const uint32_t stride = gridDim.x * 16;
    const bool pad = block_ranges_aligned(a, nh);
    // The next block's slot, cell counts and key are loaded while this block
    // is handled (its hash inserts and counter updates wait on memory).
    uint32_t nslot = 0u, nc = 0u;
    uint64_t nkey = 0ull;
    auto records = [&](uint32_t bb) {
        nslot = a.g.home_list[bb];
        nc = a.g.cellcnt[(size_t)nslot * 64 + lane];
        if (lane > 0 && lane < 8) nkey = a.g.keys[nslot];
    };
    if (blockIdx.x * 16 + wid < nh) records(blockIdx.x * 16 + wid);
    for (uint32_t b0 = blockIdx.x * 16; b0 < nh; b0 += stride) {
        const uint32_t b = b0 + wid;
        const bool valid = b < nh;
        const uint32_t slot = valid ? nslot : 0u;
        const uint32_t c = nc;
        const uint64_t key = nkey;
        const uint32_t bn = b + stride;
        const uint32_t pslot = bn < nh ? a.g.home_list[bn] : 0u;  // the next block's slot, in flight meanwhile
        uint32_t total = 0;
        uint32_t ns = slot;
        bool ins = false;
        if (valid) {
            uint32_t s = c;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t t = __shfl_up(s, o, 64);
                if (lane >= (uint32_t)o) s += t;
            }
            a.g.celloff[(size_t)slot * 64 + lane] = s - c;
            total = __shfl(s, 63, 64);
            if (lane < 8) {
                if (lane != 0) {
                    uint32_t sid;
                    int bx, by, bz;
                    unpack_key(key, sid, bx, by, bz);
                    ns = hash_insert(a.g, pack_key(sid, bx + (int)(lane & 1), by + (int)((lane >> 1) & 1), bz + (int)(lane >> 2)), ins);
                    if (ns == kInvalid) {
                        flag(a.st, sid, kStatusHashFull);
                        ns = slot;  // keep indices in range; the pass is flagged as failed
                        ins = false;
                    }
                } else {
                    a.g.all_list[b] = slot;
                }
                a.g.nb[(size_t)slot * 8 + lane] = ns;
                // The block H + t learns that home block H (home-list index b)
                // holds partial sums for it; only H writes entry t of H + t.
                if (b < a.g.max_home) a.g.rev[(size_t)ns * 8 + lane] = b;
            }
        }
        if (bn < nh) {  // the next block's cell counts and key
            nslot = pslot;
            nc = a.g.cellcnt[(size_t)pslot * 64 + lane];
            if (lane > 0 && lane < 8) nkey = a.g.keys[pslot];
        }
        // Neighbour blocks created here are appended to the all-blocks list
        // with one counter update per workgroup (a counter bumped by every
        // wave serialises at one L2 channel).
        const uint64_t imask = __ballot(ins);
        if (lane == 0) {
            tot[wid] = pad ? (total + 31u) & ~31u : total;
            ext[wid] = (uint32_t)__popcll(imask);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t run = 0, erun = 0;
            for (int k = 0; k < 16; ++k) {
                const uint32_t v = tot[k], e = ext[k];
                tot[k] = run;
                ext[k] = erun;
                run += v;
                erun += e;
            }
            // one 64-bit atomic for both counters (n_extra low, cursor high;
            // neither sum reaches 2^32, so no carry crosses between them)
            const unsigned long long old =
                (run | erun) ? atomicAdd(reinterpret_cast<unsigned long long*>(&a.cnt->n_extra),
                                         ((unsigned long long)run << 32) | erun)
                             : 0ull;
            ext_s = (uint32_t)old;
            base_s = (uint32_t)(old >> 32);
        }
        __syncthreads();
        if (valid && lane == 0) {
            a.g.blkoff[slot] = base_s + tot[wid];
            a.g.blkcnt[slot] = total;
        }
        if (ins) a.g.all_list[nh + ext_s + ext[wid] + (uint32_t)__popcll(imask & lt_mask)] = ns;
        __syncthreads();
    }
}

// Counting-sort scatter.  Each thread handles 4 particles per pass with all
// their loads issued together, so the dependent chain key -> block/cell
// offsets -> store overlaps across them.
__global__ __launch_bounds__(256) void k_scatter(StepArgs a) {
    const uint32_t n = a.cnt->n_particles;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += 4 * stride) {
        uint32_t key[4], rank[4], pid[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t i = i0 + k * stride;
            key[k] = i < n ? a.bin.pkey[i] : kInvalid;
            rank[k] = i < n ? a.bin.prank[i] : 0u;
            pid[k] = i < n ? a.cur.pid[i] : 0u;
        }
        uint32_t dst[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t slot = key[k] >> 6, l = key[k] & 63;
            dst[k] = key[k] == kInvalid ? 0u : a.g.blkoff[slot] + a.g.celloff[(size_t)slot * 64 + l] + rank[k];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (key[k] != kInvalid) {
                a.bin.sidx[dst[k]] = i0 + k * stride;
                a.bin.spid[dst[k]] = pid[k];
            }
    }
}

// Insertion sort of each cell's segment by particle id: the P2G summation
// order inside a cell then depends only on the particles, not on the order
// in which atomics handed out ranks.  One wave per block; the block's range is
// sorted in LDS (segments are almost sorted already: the buffer keeps the
// previous step's order) and written back.
constexpr int kSortCap = 1024;
__global__ __launch_bounds__(256) void k_cellsort(StepArgs a) {
    __shared__ uint32_t kp[4][kSortCap];
    __shared__ uint32_t ki[4][kSortCap];
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint32_t nh = a.cnt->n_home;
    const uint32_t stride = gridDim.x * 4;
    uint32_t b = blockIdx.x * 4 + wid;
    // The next block's records are loaded while this block is sorted.
    uint32_t noff = 0, nnb = 0, ncnt = 0, nc0 = 0;
    auto records = [&](uint32_t bb) {
        const uint32_t slot = a.g.home_list[bb];
        noff = a.g.blkoff[slot];
        nnb = a.g.blkcnt[slot];
        ncnt = a.g.cellcnt[(size_t)slot * 64 + lane];
        nc0 = a.g.celloff[(size_t)slot * 64 + lane];
    };
    if (b < nh) records(b);
    for (; b < nh; b += stride) {
        const uint32_t off = noff, nb = nnb, cnt = ncnt, c0 = nc0;
        if (b + stride < nh) records(b + stride);
        if (nb <= (uint32_t)kSortCap) {
            uint32_t* P = kp[wid];
            uint32_t* I = ki[wid];
            // all loads of a batch of 8 x 64 particles in flight before any
            // is written to LDS (one memory round trip per batch)
            for (uint32_t t0 = 0; t0 < nb; t0 += 8 * 64) {
                uint32_t pv[8], iv[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const uint32_t t = t0 + lane + 64 * k;
                    pv[k] = t < nb ? a.bin.spid[off + t] : 0u;
                    iv[k] = t < nb ? a.bin.sidx[off + t] : 0u;
                }
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const uint32_t t = t0 + lane + 64 * k;
                    if (t < nb) {
                        P[t] = pv[k];
                        I[t] = iv[k];
                    }
                }
            }
            __asm__ __volatile__("s_waitcnt lgkmcnt(0)" ::: "memory");
            bool moved = false;
            for (uint32_t k = 1; k < cnt; ++k) {
                const uint32_t kpk = P[c0 + k], kik = I[c0 + k];
                uint32_t j = k;
                while (j > 0 && P[c0 + j - 1] > kpk) {
                    P[c0 + j] = P[c0 + j - 1];
                    I[c0 + j] = I[c0 + j - 1];
                    --j;
                }
                P[c0 + j] = kpk;
                I[c0 + j] = kik;
                moved |= j != k;
            }
            __asm__ __volatile__("s_waitcnt lgkmcnt(0)" ::: "memory");
            // Most blocks arrive sorted (the buffer keeps the previous order
            // and few particles change cells): only reordered blocks are
            // written back.
            if (__ballot(moved)) {
                for (uint32_t t = lane; t < nb; t += 64) {
                    a.bin.spid[off + t] = P[t];
                    a.bin.sidx[off + t] = I[t];
                }
            }
            __asm__ __volatile__("s_waitcnt lgkmcnt(0)" ::: "memory");
        } else {
            const uint32_t s0 = off + c0;
            for (uint32_t k = 1; k < cnt; ++k) {
                const uint32_t kpk = a.bin.spid[s0 + k], kik = a.bin.sidx[s0 + k];
                uint32_t j = k;
                while (j > 0 && a.bin.spid[s0 + j - 1] > kpk) {
                    a.bin.spid[s0 + j] = a.bin.spid[s0 + j - 1];
                    a.bin.sidx[s0 + j] = a.bin.sidx[s0 + j - 1];
                    --j;
                }
                a.bin.spid[s0 + j] = kpk;
                a.bin.sidx[s0 + j] = kik;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// P2G on the matrix cores.  One wave per home block, no workgroup barriers.
// For the particles of one cell the 4x4x4 stencil nodes are fixed, and the
// reference sums (P2G grid.cpp:119-129, force_grid_accumulation
// grid.cpp:174-182)
//   m_i   += w m,   p_i += w m v,
//   f_i,c += S_c0 dwx wy wz + S_c1 wx dwy wz + S_c2 wx wy dwz   (S = -V0 P F^T)
// are, for a group of 4 particles of the cell, sums of outer products: with
// rows (a, d) = the 4x4 x/z nodes and columns (b, ch) = y node x channel,
//   C_M[(a,d),(b,ch)] += sum_k (wx_a wz_d)[k] (wy_b q_ch)[k],  q = (m, m v)
//   C_F[(a,d),(b,c)]  += sum_k (dwx_a wz_d)(wy_b S_c0) + (wx_a wz_d)(dwy_b S_c1)
//                        + (wx_a dwz_d)(wy_b S_c2)
// i.e. four v_mfma_f32_16x16x4f32 per group (K = the 4 particles).  Channel
// c = 3 of C_F is unused.  Lane l supplies A[l%16][l/16] and B[l/16][l%16]
// and holds D[4(l/16)+r][l%16], r = 0..3.  Rows split as (a, d) = (i&3, i>>2)
// and columns as (b, ch) = (j&3, j>>2), so lane l serves particle k = l/16 of
// the group with x node a = y node b = l&3 and z node d = channel
// ch = (l>>2)&3, and holds the x nodes r = 0..3 of node (., l&3, l/16),
// channel (l>>2)&3.
//
// A wave stages its block's particles (sorted: cell, then particle id) CHUNK
// (16, 32 or 64) at a time into LDS records (per-axis B-spline weights and
// derivatives of grid.cpp:24-58 in grid units, m, m v, S/dx): a lane fetches
// its x/y terms with one b128; the six z/channel terms are the same for the 4
// lanes of a quad, which fetch them with one b64 each and share them by DPP
// quad broadcasts.  A group past the end of a cell reads an all-zero record.
// The fields of the next chunk are loaded into registers while the current
// one is accumulated.  The cells of an x row (cx = 0..3 at fixed cy, cz) are
// consecutive in sorted order; their tiles, shifted by cx, are summed in
// registers (x nodes 0..6) and added to the node region once per x row.
//
// The node region is held in registers, 4 node planes Z at a time: lane
// (g, n, ch) = (l/16, l&3, (l>>2)&3) holds plane Z = zb + g (zb: the lowest
// plane not yet written), y nodes Y = n and n + 4 and channel ch of the mass
// sums and of the force sums, x nodes 0..6 (28 floats).  Cells come in
// z-major order and a row of cells at cz adds to planes cz..cz+3, the planes
// of its own lane groups once zb = cz; its y nodes cy + b reach the lanes
// that hold them by a rotation of each quad by cy (DPP).  When rows reach cz
// every plane below is complete: plane zb is written out as the block's
// partial sums for that plane (through a 384-float LDS plane buffer whose
// padding stays 0, so every 128-byte line is written whole), its registers
// are cleared and the region moves down one lane group.  All 7 planes of
// every home block are written.  With no ring in LDS a wave needs only its
// staged records and the plane buffer, so occupancy is bound by registers.
// Summation order: the particles of a cell form groups of 4 counted from the
// cell's first particle, and staging chunks end only at cell boundaries (or,
// inside a cell longer than the chunk, after a whole number of groups), so no
// group spans two chunks; the MFMAs of cells at even and at odd cx accumulate
// into two separate register sets, added in a fixed order when the row is
// deposited; rows are added to the region in sorted order.  Every float sum
// is therefore the same for any chunk size.
// A home block H's 7x7x7 node region splits into the pieces that fall into
// the blocks H + t, t = (tx, ty, tz) in {0,1}^3: 4 nodes along an axis with
// bit 0, 3 with bit 1.  part stores a home block's sums plane by plane
// (plane Z = 4 tz + lz of the region: kPlaneFloats = 384 floats), each plane
// as its 4 xy pieces txy = tx + 2 ty (at plane_piece(txy): each piece padded
// to a whole number of 128-byte lines), channel-major, nodes x-fastest.
// P2G writes every plane whole, padding included, so no line is written in
// part.
__host__ __device__ constexpr int piece_n(int t, int axis) { return ((t >> axis) & 1) ? 3 : 4; }
__host__ __device__ constexpr int piece_base(int t) {
    int b = 0;
    for (int u = 0; u < t; ++u) b += piece_n(u, 0) * piece_n(u, 1) * piece_n(u, 2);
    return b;
}
static_assert(piece_base(8) == kRegion, "pieces tile the 7^3 region");
__host__ __device__ constexpr int xy_base(int txy) {
    int b = 0;
    for (int u = 0; u < txy; ++u) b += piece_n(u, 0) * piece_n(u, 1);
    return b;
}
static_assert(xy_base(4) == 49, "xy pieces tile a 7x7 plane");
// Offset of xy piece txy within a plane of partial sums (pieces of
// 7 nx ny floats, each padded to a multiple of 32 floats).
__host__ __device__ constexpr int plane_piece(int txy) {
    int b = 0;
    for (int u = 0; u < txy; ++u) b += (kGridCh * piece_n(u, 0) * piece_n(u, 1) + 31) / 32 * 32;
    return b;
}
static_assert(plane_piece(4) == kPlaneFloats, "padded pieces fill a plane");

constexpr int kP2GRec = 40;                // floats per staged particle record
// record layout (floats):
//   [4a .. 4a+3]         a = 0..3: wx_a, dwx_a, wy_a, dwy_a
//   [16+6c .. 16+6c+5]   c = 0..3: wz_c, dwz_c, q_c, S_c0, S_c1, S_c2
//                        (q = m, m v0, m v1, m v2; S_3* = 0)
// Lane (quad member) n reads the pair [16+6c+2n, +1]; member 3's pair is
// never used (it overlaps the next record, or the 2 floats of slack).
constexpr int kP2GRecC = 16;
// staging floats of a CHUNK: the records, the all-zero record, 2 floats of slack
template <int CHUNK>
constexpr int p2g_pst() { return (CHUNK + 1) * kP2GRec + 2; }
constexpr int kP2GFlushQ = kPlaneFloats / 128;  // float pairs of a plane per lane
static_assert(kPlaneFloats % 128 == 0, "whole float pairs of a plane per lane");
// Offsets in a plane of partial sums of the x/y node pieces (tx, ty) =
// (0,0), (1,0), (0,1), (1,1) for channel c, y node ly of the piece:
// 4 x nodes (tx = 0) from a 16-byte boundary, 3 (tx = 1) from a 4-byte one.
__device__ __forceinline__ int plane_off(int txy, int c, int ly) {
    const int nx = (txy & 1) ? 3 : 4, ny = (txy & 2) ? 3 : 4;
    return plane_piece(txy) + c * nx * ny + ly * nx;
}

typedef float p2g_v4f __attribute__((ext_vector_type(4)));

// Value of lane L of this lane's quad (DPP quad_perm [L, L, L, L]).
template <int L>
__device__ __forceinline__ float quad_bcast(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), L * 0x55, 0xF, 0xF, true));
}

// A home block's records as a P2G wave needs them: lane l holds cell l's
// sorted range within the block.
struct P2GHeader {
    uint64_t key;
    uint32_t off;   // first sorted position
    int nb;         // particles in the block
    int cbeg, cend; // cell (lane) range within the block
};
__device__ __forceinline__ void p2g_header(const StepArgs& a, uint32_t slot, int lane, P2GHeader& h) {
    h.key = a.g.keys[slot];
    h.off = (uint32_t)__builtin_amdgcn_readfirstlane((int)a.g.blkoff[slot]);
    h.nb = __builtin_amdgcn_readfirstlane((int)a.g.blkcnt[slot]);
    h.cbeg = (int)a.g.celloff[(size_t)slot * 64 + lane];
    h.cend = h.cbeg + (int)a.g.cellcnt[(size_t)slot * 64 + lane];
}

// The fields P2G reads of one particle, loaded into registers one chunk
// ahead of their use so the global-load latency hides under the MFMAs.  With
// FORCE (the force differential of the semi-implicit solve) only x and the
// general 3x3 matrix G = -V0 dP F^T are read.
struct P2GRaw {
    float x[3], v[3], m, S[9];
};
template <bool FORCE>
__device__ __forceinline__ void p2g_load(const StepArgs& a, uint32_t i, P2GRaw& r) {
#pragma unroll
    for (int e = 0; e < 3; ++e) r.x[e] = comp(a.cur.x, a.cur, e, i);
    if (FORCE) {
#pragma unroll
        for (int e = 0; e < 9; ++e) r.S[e] = at(a.imp.G + (size_t)e * a.cur.stride, i);
        return;
    }
#pragma unroll
    for (int e = 0; e < 3; ++e) r.v[e] = comp(a.cur.v, a.cur, e, i);
    r.m = at(a.cur.m, i);
#pragma unroll
    for (int e = 0; e < kStressCh; ++e) r.S[e] = comp(a.cur.S, a.cur, e, i);
}

// Stage a particle as record rec: weights of the three axes, m, m v, S / dx
// (FORCE: zero mass and momentum, row c of G / dx in place of row c of S).
template <bool FORCE>
__device__ __forceinline__ void p2g_stage(const P2GRaw& r, float inv_dx, float* rec) {
    const float g0 = r.x[0] * inv_dx, g1 = r.x[1] * inv_dx, g2 = r.x[2] * inv_dx;
    const float m = FORCE ? 0.0f : r.m;
    const float q1 = FORCE ? 0.0f : m * r.v[0], q2 = FORCE ? 0.0f : m * r.v[1], q3 = FORCE ? 0.0f : m * r.v[2];
    float S[3][3];
    if (FORCE) {
#pragma unroll
        for (int e = 0; e < 9; ++e) S[e / 3][e % 3] = r.S[e] * inv_dx;
    } else {
        S[0][0] = r.S[0] * inv_dx, S[1][1] = r.S[1] * inv_dx, S[2][2] = r.S[2] * inv_dx;
        S[0][1] = S[1][0] = r.S[3] * inv_dx, S[0][2] = S[2][0] = r.S[4] * inv_dx, S[1][2] = S[2][1] = r.S[5] * inv_dx;
    }
    float wx[4], dwx[4], wy[4], dwy[4], wz[4], dwz[4];
    bspline(g0 - floorf(g0), wx, dwx);
    bspline(g1 - floorf(g1), wy, dwy);
    bspline(g2 - floorf(g2), wz, dwz);
    float4* r4 = reinterpret_cast<float4*>(rec);
#pragma unroll
    for (int e = 0; e < 4; ++e) r4[e] = make_float4(wx[e], dwx[e], wy[e], dwy[e]);
    float2* r2 = reinterpret_cast<float2*>(rec + kP2GRecC);
    r2[0] = make_float2(wz[0], dwz[0]);
    r2[1] = make_float2(m, S[0][0]);
    r2[2] = make_float2(S[0][1], S[0][2]);
    r2[3] = make_float2(wz[1], dwz[1]);
    r2[4] = make_float2(q1, S[1][0]);
    r2[5] = make_float2(S[1][1], S[1][2]);
    r2[6] = make_float2(wz[2], dwz[2]);
    r2[7] = make_float2(q2, S[2][0]);
    r2[8] = make_float2(S[2][1], S[2][2]);
    r2[9] = make_float2(wz[3], dwz[3]);
    r2[10] = make_float2(q3, 0.0f);
    r2[11] = make_float2(0.0f, 0.0f);
}

// Value of lane (n - R) & 3 of this lane's quad, n = this lane's place in it
// (DPP quad_perm [-R, 1-R, 2-R, 3-R] mod 4).
template <int R>
__device__ __forceinline__ float quad_rot(float v) {
    constexpr int ctl = ((4 - R) & 3) | (((5 - R) & 3) << 2) | (((6 - R) & 3) << 4) | (((7 - R) & 3) << 6);
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctl, 0xF, 0xF, true));
}

// Adds an x row of cells at cy = R (its sums sm, sf for y node R + n, x
// nodes 0..6, in quad member n) to the region: quad member n holds y nodes
// n (slot 0) and n + 4 (slot 1); y node R + ((n - R) & 3) is n for n >= R.
// (Value selects, not branches: a store through a selected slot would keep
// the region out of registers.)
template <int R, bool MASS>
__device__ __forceinline__ void p2g_row_add(float (&rm)[2][7], float (&rf)[2][7], const float (&sm)[7],
                                            const float (&sf)[7], int n) {
    const bool s0 = n >= R;
#pragma unroll
    for (int x = 0; x < 7; ++x) {
        const float tm = R ? quad_rot<R>(sm[x]) : sm[x];
        const float tf = R ? quad_rot<R>(sf[x]) : sf[x];
        if (R == 0) {
            if (MASS) rm[0][x] += tm;
            rf[0][x] += tf;
        } else {
            if (MASS) {
                const float a0 = rm[0][x] + tm, a1 = rm[1][x] + tm;
                rm[0][x] = s0 ? a0 : rm[0][x];
                rm[1][x] = s0 ? rm[1][x] : a1;
            }
            const float b0 = rf[0][x] + tf, b1 = rf[1][x] + tf;
            rf[0][x] = s0 ? b0 : rf[0][x];
            rf[1][x] = s0 ? rf[1][x] : b1;
        }
    }
}

// Length of the staging chunk that starts at sorted position c of a block of
// nb particles whose cell l starts at cbeg (lane l) (wave-uniform): up to
// CHUNK particles, ending at the last cell boundary that
// fits, or after CHUNK particles (a whole number of groups) when the cell at
// c is longer than that.  c is a cell start or such a split point.
template <int CHUNK>
__device__ __forceinline__ int p2g_chunk_len(int cbeg, int nb, int c) {
    const int end = c + CHUNK;
    if (end >= nb) return nb - c;
    const uint64_t starts = __ballot(cbeg <= end);  // cells are in order: a prefix of the lanes
    const int l = 63 - __clzll((long long)starts);
    const int b = __builtin_amdgcn_readlane(cbeg, l);
    return b > c ? b - c : CHUNK;
}

// At most 128 VGPRs, 4 waves per SIMD (the region in registers would
// otherwise take ~134 and allow only 3; the few spills are outside the
// per-cell loop).
template <int CHUNK, bool FORCE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_p2g(StepArgs a) {
    static_assert(CHUNK <= 64 && CHUNK % 4 == 0, "one staged particle per lane, whole groups");
    constexpr int kPst = p2g_pst<CHUNK>();
    __shared__ __align__(16) float plane[kPlaneFloats];
    __shared__ __align__(16) float pst[kPst];
    const int lane = (int)threadIdx.x;
    const int kq = lane >> 4, n = lane & 15, na = n & 3, nb = n >> 2;
    const uint32_t nh = a.cnt->n_home;
    const uint32_t G = gridDim.x;
    uint32_t blk = blockIdx.x;
    if (blk >= nh) return;
    // the plane buffer's padding stays 0; the zero record
    for (int e = lane; e < kPlaneFloats / 4; e += 64) reinterpret_cast<float4*>(plane)[e] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (lane < kP2GRec + 2) pst[CHUNK * kP2GRec + lane] = 0.0f;
    // the node region: [y slot][x] of this lane's plane, y nodes and channel
    float rm[2][7], rf[2][7];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int x = 0; x < 7; ++x) rm[t][x] = rf[t][x] = 0.0f;

    // Chunk pipeline over this wave's blocks blk, blk + G, ...: while chunk k
    // is accumulated, the fields of chunk k + 1 and the sorted indices of
    // chunk k + 2 are loaded.  Headers are kept for blocks blk, blk + G, blk + 2G
    // (hdr 0, 1, 2); a chunk is (hdr, start, len).
    P2GHeader h{}, hn{}, hnn{};
    p2g_header(a, a.g.home_list[blk], lane, h);
    if (blk + G < nh) p2g_header(a, a.g.home_list[blk + G], lane, hn);
    if (blk + 2 * G < nh) p2g_header(a, a.g.home_list[blk + 2 * G], lane, hnn);
    struct Pos {
        int hdr, start, len;
    };
    // (every branch on p.hdr is wave-uniform and names the header it uses: a
    // header picked by a run-time index would live in scratch memory)
    auto exists = [&](int k) { return blk + (uint32_t)k * G < nh; };
    auto after = [&](const Pos& p) -> Pos {  // p.hdr is 0 or 1
        const int e = p.start + p.len;
        if (p.hdr == 0) {
            if (e < h.nb) return Pos{0, e, p2g_chunk_len<CHUNK>(h.cbeg, h.nb, e)};
            return Pos{1, 0, exists(1) ? p2g_chunk_len<CHUNK>(hn.cbeg, hn.nb, 0) : 0};
        }
        if (exists(1) && e < hn.nb) return Pos{1, e, p2g_chunk_len<CHUNK>(hn.cbeg, hn.nb, e)};
        return Pos{2, 0, exists(2) ? p2g_chunk_len<CHUNK>(hnn.cbeg, hnn.nb, 0) : 0};
    };
    // This lane's particle index in chunk p.  The load is issued
    // unconditionally (sorted position 0 when the lane has no particle there)
    // and the flag says whether it is used, so no branch forces a wait on the
    // loads in flight.
    bool ok;
    auto chunk_index = [&](const Pos& p) -> uint32_t {
        ok = exists(p.hdr) && lane < p.len;
        const uint32_t o0 = h.off, o1 = hn.off, o2 = hnn.off;  // read all, then select values
        const uint32_t off = p.hdr == 0 ? o0 : (p.hdr == 1 ? o1 : o2);
        return a.bin.sidx[ok ? off + p.start + lane : 0u];
    };
    Pos cur{0, 0, p2g_chunk_len<CHUNK>(h.cbeg, h.nb, 0)};
    Pos nxt = after(cur);
    P2GRaw raw{};
    {
        const uint32_t i0 = chunk_index(cur);
        if (ok) p2g_load<FORCE>(a, i0, raw);
    }
    uint32_t inext = chunk_index(nxt);
    bool inext_ok = ok;

    // operands of the group at chunk position p of a cell ending at hi (the
    // all-zero record for the lanes past the end)
    const char* lane_x = reinterpret_cast<const char*>(pst + kq * kP2GRec + 4 * na);
    const char* lane_c = reinterpret_cast<const char*>(pst + kq * kP2GRec + kP2GRecC + 6 * nb + 2 * na);
    const int zero_b = (CHUNK - kq) * kP2GRec * 4;  // bytes from record kq to the zero record
    auto fetch = [&](int p, int hi, float4& x, float2& c) {
        const int b = kq < hi - p ? p * (kP2GRec * 4) : zero_b;  // hi - p and p * 160 are wave-uniform
        x = *reinterpret_cast<const float4*>(lane_x + b);  // wx_a dwx_a wy_b dwy_b
        c = *reinterpret_cast<const float2*>(lane_c + b);
    };

    for (; blk < nh; blk += G) {
        uint32_t sid;
        int bx, by, bz;
        unpack_key(h.key, sid, bx, by, bz);
        const float inv_dx = a.scenes[sid].inv_dx;
        const bool write = blk < a.g.max_home;
        if (!write && lane == 0) flag(a.st, sid, kStatusOverflow);
        float* out = a.g.part + (size_t)(write ? blk : 0u) * kPartFloats;
        // Writes the plane lane group g holds as plane Z of the block's
        // partial sums: the group's lanes put it into the plane buffer, then
        // every lane stores 3 float pairs of it (whole 128-byte lines).
        auto write_plane = [&](int g, int Z) __attribute__((always_inline)) {
            if (kq == g) {
                // plane buffer offsets of this lane's nodes: y node na (ty = 0)
                // and na + 4 (ty = 1), x nodes 0..3 (tx = 0) and 4..6 (tx = 1);
                // mass channel nb, force channel 4 + nb.  (Formed here from a
                // lane id the compiler cannot see through, so they are not
                // held in registers across the block.)
                int ln = lane;
                __asm__ __volatile__("" : "+v"(ln));
                const int la = ln & 3, lb = (ln >> 2) & 3;
                const int pm00 = plane_off(0, lb, la), pm10 = plane_off(1, lb, la);
                const int pm01 = plane_off(2, lb, la), pm11 = plane_off(3, lb, la);
                const int pf00 = plane_off(0, 4 + lb, la), pf10 = plane_off(1, 4 + lb, la);
                const int pf01 = plane_off(2, 4 + lb, la), pf11 = plane_off(3, 4 + lb, la);
                *reinterpret_cast<float4*>(plane + pm00) = make_float4(rm[0][0], rm[0][1], rm[0][2], rm[0][3]);
                plane[pm10] = rm[0][4], plane[pm10 + 1] = rm[0][5], plane[pm10 + 2] = rm[0][6];
                if (na < 3) {
                    *reinterpret_cast<float4*>(plane + pm01) = make_float4(rm[1][0], rm[1][1], rm[1][2], rm[1][3]);
                    plane[pm11] = rm[1][4], plane[pm11 + 1] = rm[1][5], plane[pm11 + 2] = rm[1][6];
                }
                if (nb < 3) {
                    *reinterpret_cast<float4*>(plane + pf00) = make_float4(rf[0][0], rf[0][1], rf[0][2], rf[0][3]);
                    plane[pf10] = rf[0][4], plane[pf10 + 1] = rf[0][5], plane[pf10 + 2] = rf[0][6];
                    if (na < 3) {
                        *reinterpret_cast<float4*>(plane + pf01) =
                            make_float4(rf[1][0], rf[1][1], rf[1][2], rf[1][3]);
                        plane[pf11] = rf[1][4], plane[pf11 + 1] = rf[1][5], plane[pf11 + 2] = rf[1][6];
                    }
                }
            }
            float2 v[kP2GFlushQ];  // (a wave's LDS operations complete in order)
#pragma unroll
            for (int q = 0; q < kP2GFlushQ; ++q) v[q] = reinterpret_cast<const float2*>(plane)[lane + 64 * q];
            float2* po = reinterpret_cast<float2*>(out + Z * kPlaneFloats);
            if (write)
#pragma unroll
                for (int q = 0; q < kP2GFlushQ; ++q) po[lane + 64 * q] = v[q];
        };
        // Plane zb is complete: write it, clear it and move the region down
        // one lane group (group 3 then holds the empty plane zb + 4).
        int zb = 0;
        auto advance = [&]() __attribute__((always_inline)) {
            write_plane(0, zb);
            const int rot_src = ((lane + 16) & 63) << 2;  // ds_bpermute address: lane group g + 1
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int x = 0; x < 7; ++x) {
                    const float m = kq == 0 ? 0.0f : rm[t][x], f = kq == 0 ? 0.0f : rf[t][x];
                    rm[t][x] = __int_as_float(__builtin_amdgcn_ds_bpermute(rot_src, __float_as_int(m)));
                    rf[t][x] = __int_as_float(__builtin_amdgcn_ds_bpermute(rot_src, __float_as_int(f)));
                }
            ++zb;
        };
        // x-row accumulators, C_M and C_F: cells at even cx add to x nodes
        // 0..5 (aE), cells at odd cx to x nodes 1..6 (aO, index x - 1)
        float amE[6], afE[6], amO[6], afO[6];
#pragma unroll
        for (int e = 0; e < 6; ++e) amE[e] = afE[e] = amO[e] = afO[e] = 0.0f;
        int row = 0;  // current x row (cy + 4 cz)
        int l = 0;    // next cell to visit
        auto deposit_row = [&]() __attribute__((always_inline)) {
            const int cy = row & 3, cz = row >> 2;
            // rows have reached cz: planes below it are complete (cz - zb <= 3)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (zb < cz) advance();
            const float sm[7] = {amE[0],         amE[1] + amO[0], amE[2] + amO[1], amE[3] + amO[2],
                                 amE[4] + amO[3], amE[5] + amO[4], amO[5]};
            const float sf[7] = {afE[0],         afE[1] + afO[0], afE[2] + afO[1], afE[3] + afO[2],
                                 afE[4] + afO[3], afE[5] + afO[4], afO[5]};
            if (cy == 0)
                p2g_row_add<0, !FORCE>(rm, rf, sm, sf, na);
            else if (cy == 1)
                p2g_row_add<1, !FORCE>(rm, rf, sm, sf, na);
            else if (cy == 2)
                p2g_row_add<2, !FORCE>(rm, rf, sm, sf, na);
            else
                p2g_row_add<3, !FORCE>(rm, rf, sm, sf, na);
#pragma unroll
            for (int e = 0; e < 6; ++e) amE[e] = afE[e] = amO[e] = afO[e] = 0.0f;
        };
        for (;;) {  // the chunks of this block
            const int c0 = cur.start, cn = cur.len;
            if (lane < cn) p2g_stage<FORCE>(raw, inv_dx, pst + lane * kP2GRec);
            // fields of chunk k + 1, then the indices of chunk k + 2
            if (inext_ok) p2g_load<FORCE>(a, inext, raw);
            const Pos nn = after(nxt);
            inext = chunk_index(nn);
            inext_ok = ok;

            // cells l.. that reach into this chunk, in sorted order.  The
            // cell code is written out for each x position cx = l & 3 in the
            // row, so its tile lands on x nodes cx..cx+3 of the row
            // accumulators without data-dependent register selection.
            bool more = true;
            while (more && l < 64) {
#pragma unroll
                for (int cx = 0; cx < 4; ++cx) {
                    if (!more || (l & 3) != cx) continue;
                    if (cx == 0 && (l >> 2) != row) {  // a new x row: the rows before it are complete
                        deposit_row();
                        row = l >> 2;
                    }
                    const int cb = __builtin_amdgcn_readlane(h.cbeg, l) - c0;
                    const int ce = __builtin_amdgcn_readlane(h.cend, l) - c0;
                    if (cb >= cn) {  // starts in a later chunk
                        more = false;
                        continue;
                    }
                    const int lo = max(cb, 0), hi = min(ce, cn);
                    if (lo < hi) {
                        float4 x0, x1;
                        float2 q0, q1;
                        fetch(lo, hi, x0, q0);
                        fetch(lo + 4, hi, x1, q1);
                        __builtin_amdgcn_sched_barrier(0);  // both groups' reads in flight before the MFMAs
                        // the MFMAs accumulate straight into the row registers of
                        // x nodes cx..cx+3 (an even-aligned quad of aE or aO, as
                        // MFMA operands must be)
                        const int o = cx & ~1;
                        p2g_v4f cm, cf;
                        if (cx & 1) {
                            cm = p2g_v4f{amO[o], amO[o + 1], amO[o + 2], amO[o + 3]};
                            cf = p2g_v4f{afO[o], afO[o + 1], afO[o + 2], afO[o + 3]};
                        } else {
                            cm = p2g_v4f{amE[o], amE[o + 1], amE[o + 2], amE[o + 3]};
                            cf = p2g_v4f{afE[o], afE[o + 1], afE[o + 2], afE[o + 3]};
                        }
                        auto group = [&](const float4& rx, const float2& rc) {
                            const float wz = quad_bcast<0>(rc.x), dwz = quad_bcast<0>(rc.y);
                            const float q = quad_bcast<1>(rc.x), s0 = quad_bcast<1>(rc.y);
                            const float s1 = quad_bcast<2>(rc.x), s2 = quad_bcast<2>(rc.y);
                            const float wxz = rx.x * wz;  // wx_a wz_d
                            if (!FORCE) cm = __builtin_amdgcn_mfma_f32_16x16x4f32(wxz, rx.z * q, cm, 0, 0, 0);
                            cf = __builtin_amdgcn_mfma_f32_16x16x4f32(rx.y * wz, rx.z * s0, cf, 0, 0, 0);
                            cf = __builtin_amdgcn_mfma_f32_16x16x4f32(wxz, rx.w * s1, cf, 0, 0, 0);
                            cf = __builtin_amdgcn_mfma_f32_16x16x4f32(rx.x * dwz, rx.z * s2, cf, 0, 0, 0);
                        };
                        // the first two groups (a cell piece of <= 4 particles needs
                        // only the first: the second would read the zero record);
                        // longer pieces loop
                        group(x0, q0);
                        if (hi - lo > 4) group(x1, q1);
                        for (int p = lo + 8; p < hi; p += 4) {
                            fetch(p, hi, x0, q0);
                            group(x0, q0);
                        }
                        if (cx & 1) {
#pragma unroll
                            for (int r = 0; r < 4; ++r) amO[o + r] = cm[r], afO[o + r] = cf[r];
                        } else {
#pragma unroll
                            for (int r = 0; r < 4; ++r) amE[o + r] = cm[r], afE[o + r] = cf[r];
                        }
                    }
                    if (ce > cn)  // continues in the next chunk
                        more = false;
                    else
                        ++l;
                }
            }
            // advance the chunk pipeline: chunk k + 1 becomes current
            const bool last = nxt.hdr != cur.hdr;  // chunk k + 1 belongs to the next block
            cur = nxt;
            nxt = nn;
            if (last) break;
        }
        deposit_row();
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (zb < 3) advance();
        // the planes of the last rows: zb..zb+3 = 3..6 in lane groups 0..3
#pragma unroll
        for (int g = 0; g < 4; ++g) write_plane(g, 3 + g);
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int x = 0; x < 7; ++x) rm[t][x] = rf[t][x] = 0.0f;
        // shift the headers and the chunk positions to the next block
        h = hn;
        hn = hnn;
        if (blk + 3 * G < nh) p2g_header(a, a.g.home_list[blk + 3 * G], lane, hnn);
        cur.hdr -= 1;
        nxt.hdr -= 1;
    }
}

// ---------------------------------------------------------------------------
// Grid update: each node's mass, momentum and force are the sum, in the fixed
// order t = 0..7, of the partials of the home blocks B - t that reach it;
// then grid_velocity (grid.cpp:195-204) and grid_collisions
// (grid.cpp:233-243) in FP64, plus the PIC/FLIP blend u that G2P needs:
// u = (1-alpha) v + alpha [m>0] (v - p/m)   (grid.cpp:312-324).
// One wave per block.  With mass_only (setup) only the mass is produced, in
// the first velocity channel.
__global__ __launch_bounds__(256) void k_grid(StepArgs a) {
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int lx = (int)(lane & 3), ly = (int)((lane >> 2) & 3), lz = (int)(lane >> 4);
    const uint32_t na = a.cnt->n_home + a.cnt->n_extra;
    const uint32_t stride = gridDim.x * 4;
    uint32_t b = blockIdx.x * 4 + wid;
    // The next block's slot, key and home-block indices are loaded while
    // this block sums its partials.
    uint32_t nslot = 0, nrev = kInvalid;
    uint64_t nkey = 0;
    if (b < na) {
        nslot = a.g.all_list[b];
        nkey = a.g.keys[nslot];
        nrev = lane < 8 ? a.g.rev[(size_t)nslot * 8 + lane] : kInvalid;
    }
    for (; b < na; b += stride) {
        const uint32_t slot = nslot, revl = nrev;
        uint32_t sid;
        int bx, by, bz;
        unpack_key(nkey, sid, bx, by, bz);
        if (b + stride < na) {
            nslot = a.g.all_list[b + stride];
            nkey = a.g.keys[nslot];
            nrev = lane < 8 ? a.g.rev[(size_t)nslot * 8 + lane] : kInvalid;
        }
        const SceneDev& sc = a.scenes[sid];
        float q[kGridCh];
#pragma unroll
        for (int c = 0; c < kGridCh; ++c) q[c] = 0.0f;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            // node (lx, ly, lz) of piece t of home block hb: plane 4 tz + lz of
            // its region, xy piece t & 3 (at plane_piece)
            const int nx = piece_n(t, 0), ny = piece_n(t, 1), nz = piece_n(t, 2), nxy = nx * ny;
            const uint32_t hb = (uint32_t)__builtin_amdgcn_readlane((int)revl, t);
            if (hb != kInvalid && lx < nx && ly < ny && lz < nz) {
                const float* src = a.g.part + (size_t)hb * kPartFloats + (4 * (t >> 2) + lz) * kPlaneFloats +
                                   plane_piece(t & 3) + ly * nx + lx;
#pragma unroll
                for (int c = 0; c < kGridCh; ++c) q[c] += src[c * nxy];
            }
        }
        float* out = a.g.vu + (size_t)slot * kVelCh * 64;
        if (a.mass_only) {
            out[lane] = q[0];
            continue;
        }
        if (!isfinite(q[0] + q[1] + q[2] + q[3] + q[4] + q[5] + q[6])) flag(a.st, sid, kStatusNaN);
        double v[3] = {0.0, 0.0, 0.0}, dv[3] = {0.0, 0.0, 0.0};
        if (q[0] > 0.0f) {
            const double im = 1.0 / (double)q[0];
            const double dt = sc.dt;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const double vold = (double)q[1 + r] * im;
                v[r] = vold + dt * ((double)q[4 + r] * im + (double)sc.gravity[r]);
                dv[r] = -vold;
            }
        }
        const double dx = sc.dx;
        const double xn0 = dx * (double)(4 * bx + lx), xn1 = dx * (double)(4 * by + ly), xn2 = dx * (double)(4 * bz + lz);
        const double vs0 = v[0], vs1 = v[1], vs2 = v[2];
        collide<double>(sc, xn0, xn1, xn2, v[0], v[1], v[2]);
        if (a.imp.gm) {
            // semi-implicit solve: the unknowns are the nodes with mass of
            // scenes with beta > 0 that no collider changed
            const bool held = v[0] != vs0 || v[1] != vs1 || v[2] != vs2;
            a.imp.gm[(size_t)slot * 64 + lane] = (q[0] > 0.0f && !held && sc.beta > 0.0f) ? q[0] : 0.0f;
        }
        const double al = sc.alpha;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const double flip = (q[0] > 0.0f) ? (v[r] + dv[r]) : 0.0;
            out[r * 64 + lane] = (float)v[r];
            out[(3 + r) * 64 + lane] = (float)((1.0 - al) * v[r] + al * flip);
        }
    }
}

// ---------------------------------------------------------------------------
// Plasticity (grid.cpp:277-285) and the next step's stress
// (grid.cpp:162-176) from one factorisation.  The trial elastic gradient
// Fh = F_E + dt L F_E is factored through the symmetric eigenproblem
// Fh^T Fh = V diag(s^2) V^T.  Clamping s to [1-theta_c, 1+theta_s] gives
// F_E' = U diag(s_c) V^T = Fh V diag(s_c/s) V^T, J_P' = J_P / prod(s_c/s),
// F_P' = V diag(s/s_c) V^T F_P, and, with U = Fh V diag(1/s),
// -V0 P(F_E') F_E'^T = -V0 U diag(2 mu (s_c-1) s_c + lambda (J-1) J) U^T.
// Inverted, singular or badly conditioned trial gradients (cond > kappa) are
// redone in FP64 with an explicit unsigned SVD.
struct PlastOut {
    float Fe[9];
    float Jp;
    float S[6];
};

__device__ __noinline__ void plastic_fp64(const float Fe[9], const float L[9], float dt, float Jp_old,
                                          const float* Fp_old, float* Fp_new, bool clampit, float tc, float ts,
                                          float mu0, float la0, float xi, float V0, PlastOut& o) {
    double F[9], Fh[9];
    for (int i = 0; i < 9; ++i) F[i] = Fe[i];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += (double)L[r * 3 + k] * F[k * 3 + c];
            Fh[r * 3 + c] = F[r * 3 + c] + (double)dt * s;
        }
    double U[9], S[3], V[9];
    svd3_fp64(Fh, U, S, V);
    double sc[3];
    for (int k = 0; k < 3; ++k) {
        sc[k] = S[k];
        if (clampit) sc[k] = fmin(fmax(S[k], 1.0 - (double)tc), 1.0 + (double)ts);
    }
    double Fn[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += U[r * 3 + k] * sc[k] * V[c * 3 + k];
            Fn[r * 3 + c] = s;
        }
    const double detFh = det3<double>(Fh), detFn = det3<double>(Fn);
    double Jp = (double)Jp_old * (detFn != 0.0 ? detFh / detFn : 0.0);
    if (Fp_old) {
        // F_P' = F_E'^{-1} Fh F_P   (grid.cpp:285)
        double inv[9];
        const double id = 1.0 / detFn;
        inv[0] = (Fn[4] * Fn[8] - Fn[5] * Fn[7]) * id;
        inv[1] = (Fn[2] * Fn[7] - Fn[1] * Fn[8]) * id;
        inv[2] = (Fn[1] * Fn[5] - Fn[2] * Fn[4]) * id;
        inv[3] = (Fn[5] * Fn[6] - Fn[3] * Fn[8]) * id;
        inv[4] = (Fn[0] * Fn[8] - Fn[2] * Fn[6]) * id;
        inv[5] = (Fn[2] * Fn[3] - Fn[0] * Fn[5]) * id;
        inv[6] = (Fn[3] * Fn[7] - Fn[4] * Fn[6]) * id;
        inv[7] = (Fn[1] * Fn[6] - Fn[0] * Fn[7]) * id;
        inv[8] = (Fn[0] * Fn[4] - Fn[1] * Fn[3]) * id;
        double T[9], P[9];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k) s += inv[r * 3 + k] * Fh[k * 3 + c];
                T[r * 3 + c] = s;
            }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k) s += T[r * 3 + k] * (double)Fp_old[k * 3 + c];
                P[r * 3 + c] = s;
            }
        for (int i = 0; i < 9; ++i) Fp_new[i] = (float)P[i];
        Jp = det3<double>(P);
    }
    const double h = exp((double)xi * (1.0 - Jp));
    const double mu = (double)mu0 * h, la = (double)la0 * h;
    const double J = detFn;
    double q[3];
    for (int k = 0; k < 3; ++k) q[k] = 2.0 * mu * (sc[k] - 1.0) * sc[k] + la * (J - 1.0) * J;
    const int rr[6] = {0, 1, 2, 0, 0, 1}, cc[6] = {0, 1, 2, 1, 2, 2};
    for (int e = 0; e < 6; ++e) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += q[k] * U[rr[e] * 3 + k] * U[cc[e] * 3 + k];
        o.S[e] = (float)(-(double)V0 * s);
    }
    for (int i = 0; i < 9; ++i) o.Fe[i] = (float)Fn[i];
    o.Jp = (float)Jp;
}

// Returns true when the FP64 path was (or, with defer, must be) taken.
template <bool TRACK>
__device__ __forceinline__ bool plastic_update(const float (&Fe)[9], const float (&L)[9], float dt, float Jp_old,
                                               const float (&Fp_old)[9], float (&Fp_new)[9], bool clampit,
                                               float tc, float ts, float mu0, float la0, float xi, float V0,
                                               float kappa, PlastOut& o, bool defer = false) {
    float Fh[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float s = L[r * 3 + 0] * Fe[0 * 3 + c] + L[r * 3 + 1] * Fe[1 * 3 + c] + L[r * 3 + 2] * Fe[2 * 3 + c];
            Fh[r * 3 + c] = Fe[r * 3 + c] + dt * s;
        }
    float A[3][3], V[3][3], lam[3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = r; c < 3; ++c) {
            const float s = Fh[0 * 3 + r] * Fh[0 * 3 + c] + Fh[1 * 3 + r] * Fh[1 * 3 + c] + Fh[2 * 3 + r] * Fh[2 * 3 + c];
            A[r][c] = s;
            A[c][r] = s;
        }
    sym_eig3<float, 4>(A, V, lam);
    float s[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) s[k] = sqrt_(fmaxf(lam[k], 0.0f));
    const float detFh = det3<float>(Fh);
    const float smin = fminf(s[0], fminf(s[1], s[2])), smax = fmaxf(s[0], fmaxf(s[1], s[2]));
    if (!(detFh > 0.0f) || !(smin > 0.0f) || !(smax <= kappa * smin)) {
        if (!defer) plastic_fp64(Fe, L, dt, Jp_old, TRACK ? Fp_old : nullptr, Fp_new, clampit, tc, ts, mu0, la0, xi, V0, o);
        return true;
    }
    float sc[3], cr[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        sc[k] = clampit ? fminf(fmaxf(s[k], 1.0f - tc), 1.0f + ts) : s[k];
        cr[k] = sc[k] * rcp_(s[k]);
    }
    const bool plastic = (cr[0] != 1.0f) || (cr[1] != 1.0f) || (cr[2] != 1.0f);
    const float prodc = cr[0] * cr[1] * cr[2];
    if (plastic) {
        float W[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                W[r][c] = cr[0] * V[r][0] * V[c][0] + cr[1] * V[r][1] * V[c][1] + cr[2] * V[r][2] * V[c][2];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                o.Fe[r * 3 + c] = Fh[r * 3 + 0] * W[0][c] + Fh[r * 3 + 1] * W[1][c] + Fh[r * 3 + 2] * W[2][c];
        if (TRACK) {
            float Wi[3][3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    Wi[r][c] = V[r][0] * V[c][0] * rcp_(cr[0]) + V[r][1] * V[c][1] * rcp_(cr[1]) +
                               V[r][2] * V[c][2] * rcp_(cr[2]);
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    Fp_new[r * 3 + c] = Wi[r][0] * Fp_old[0 * 3 + c] + Wi[r][1] * Fp_old[1 * 3 + c] + Wi[r][2] * Fp_old[2 * 3 + c];
        }
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) o.Fe[i] = Fh[i];
        if (TRACK) {
#pragma unroll
            for (int i = 0; i < 9; ++i) Fp_new[i] = Fp_old[i];
        }
    }
    o.Jp = Jp_old * rcp_(prodc);
    const float J = detFh * prodc;
    const float h = expf(xi * (1.0f - o.Jp));
    const float mu = mu0 * h, la = la0 * h;
    float U[3][3], q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float is = rcp_(s[k]);
#pragma unroll
        for (int r = 0; r < 3; ++r) U[r][k] = (Fh[r * 3 + 0] * V[0][k] + Fh[r * 3 + 1] * V[1][k] + Fh[r * 3 + 2] * V[2][k]) * is;
        q[k] = 2.0f * mu * (sc[k] - 1.0f) * sc[k] + la * (J - 1.0f) * J;
    }
    const float nv = -V0;
    o.S[0] = nv * (q[0] * U[0][0] * U[0][0] + q[1] * U[0][1] * U[0][1] + q[2] * U[0][2] * U[0][2]);
    o.S[1] = nv * (q[0] * U[1][0] * U[1][0] + q[1] * U[1][1] * U[1][1] + q[2] * U[1][2] * U[1][2]);
    o.S[2] = nv * (q[0] * U[2][0] * U[2][0] + q[1] * U[2][1] * U[2][1] + q[2] * U[2][2] * U[2][2]);
    o.S[3] = nv * (q[0] * U[0][0] * U[1][0] + q[1] * U[0][1] * U[1][1] + q[2] * U[0][2] * U[1][2]);
    o.S[4] = nv * (q[0] * U[0][0] * U[2][0] + q[1] * U[0][1] * U[2][1] + q[2] * U[0][2] * U[2][2]);
    o.S[5] = nv * (q[0] * U[1][0] * U[2][0] + q[1] * U[1][1] * U[2][1] + q[2] * U[1][2] * U[2][2]);
    return false;
}

// Elastic trial step.  If the Gershgorin discs of B = Fh Fh^T lie inside
// [(1-theta_c)^2, (1+theta_s)^2] (and inside [0.9, 1.1]), no singular value of
// Fh leaves the clamp interval, so the plasticity update is the identity on
// F_E and J_P (F_E' = Fh, as U Sigma V^T = Fh), and the stress needs no
// factorisation: P(F) F^T = 2 mu (B - sqrt(B)) + lambda (J-1) J I with
// B - sqrt(B) = E/2 + E^2/8 - E^3/16 + 5E^4/128 - 7E^5/256 + O(E^6), E = B - I
// (|E| <= 0.1 bounds the truncation below 2e-8 relative).  E is symmetric and
// every Horner factor is a polynomial in E, so all products are symmetric.
struct Sym {
    float a00, a11, a22, a01, a02, a12;
};
__device__ __forceinline__ Sym sym_mul(const Sym& A, const Sym& B) {
    Sym C;
    C.a00 = A.a00 * B.a00 + A.a01 * B.a01 + A.a02 * B.a02;
    C.a11 = A.a01 * B.a01 + A.a11 * B.a11 + A.a12 * B.a12;
    C.a22 = A.a02 * B.a02 + A.a12 * B.a12 + A.a22 * B.a22;
    C.a01 = A.a00 * B.a01 + A.a01 * B.a11 + A.a02 * B.a12;
    C.a02 = A.a00 * B.a02 + A.a01 * B.a12 + A.a02 * B.a22;
    C.a12 = A.a01 * B.a02 + A.a11 * B.a12 + A.a12 * B.a22;
    return C;
}

__device__ __forceinline__ bool elastic_trial(const float Fh[9], float tc, float ts, Sym& E) {
    E.a00 = Fh[0] * Fh[0] + Fh[1] * Fh[1] + Fh[2] * Fh[2] - 1.0f;
    E.a11 = Fh[3] * Fh[3] + Fh[4] * Fh[4] + Fh[5] * Fh[5] - 1.0f;
    E.a22 = Fh[6] * Fh[6] + Fh[7] * Fh[7] + Fh[8] * Fh[8] - 1.0f;
    E.a01 = Fh[0] * Fh[3] + Fh[1] * Fh[4] + Fh[2] * Fh[5];
    E.a02 = Fh[0] * Fh[6] + Fh[1] * Fh[7] + Fh[2] * Fh[8];
    E.a12 = Fh[3] * Fh[6] + Fh[4] * Fh[7] + Fh[5] * Fh[8];
    const float r0 = fabsf(E.a01) + fabsf(E.a02), r1 = fabsf(E.a01) + fabsf(E.a12), r2 = fabsf(E.a02) + fabsf(E.a12);
    const float lo = fminf(E.a00 - r0, fminf(E.a11 - r1, E.a22 - r2));
    const float hi = fmaxf(E.a00 + r0, fmaxf(E.a11 + r1, E.a22 + r2));
    const float c = 1.0f - tc, t = 1.0f + ts;
    return lo >= fmaxf(c * c - 1.0f, -0.1f) && hi <= fminf(t * t - 1.0f, 0.1f) && det3<float>(Fh) > 0.0f;
}

__device__ __forceinline__ void elastic_stress(const Sym& E, float J, float mu, float la, float V0, float S[6]) {
    Sym T;
    T.a00 = T.a11 = T.a22 = 5.0f / 128.0f;
    T.a01 = T.a02 = T.a12 = 0.0f;
    const float c5 = -7.0f / 256.0f;
    T.a00 += c5 * E.a00;
    T.a11 += c5 * E.a11;
    T.a22 += c5 * E.a22;
    T.a01 = c5 * E.a01;
    T.a02 = c5 * E.a02;
    T.a12 = c5 * E.a12;
    const float cs[3] = {-1.0f / 16.0f, 1.0f / 8.0f, 0.5f};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T = sym_mul(E, T);
        T.a00 += cs[k];
        T.a11 += cs[k];
        T.a22 += cs[k];
    }
    const Sym f = sym_mul(E, T);  // B - sqrt(B)
    const float two_mu = 2.0f * mu, iso = la * (J - 1.0f) * J, nv = -V0;
    S[0] = nv * (two_mu * f.a00 + iso);
    S[1] = nv * (two_mu * f.a11 + iso);
    S[2] = nv * (two_mu * f.a22 + iso);
    S[3] = nv * (two_mu * f.a01);
    S[4] = nv * (two_mu * f.a02);
    S[5] = nv * (two_mu * f.a12);
}


// ---------------------------------------------------------------------------
// G2P grid staging: neighbour block t of a home block (t = 0..7) is staged by
// a group of 32 threads, each holding 12 of its 6 x 64 floats, so one slot
// lookup per thread and block suffices and the loads are 128-byte runs.
constexpr int kG2PStage = kVelCh * 64 / 32;
static_assert(kG2PStage * 32 == kVelCh * 64, "staging split");
// Threads per G2P workgroup (T = 64, 128 or 256, chosen by the autotuner):
// one workgroup per home block, so a block's particles take ceil(count / T)
// passes; smaller T leaves fewer idle lanes in the last pass for the
// 100-500-particle blocks of uneven scenes, larger T fewer staging rounds.

__device__ __forceinline__ void g2p_stage_load(const StepArgs& a, uint32_t ts, int l32, float (&r)[kG2PStage]) {
    const float* src = a.g.vu + (size_t)ts * kVelCh * 64;
#pragma unroll
    for (int i = 0; i < kG2PStage; ++i) r[i] = src[l32 + 32 * i];
}

// Staged node layout: node n = Z * 64 + Y * 8 + X of the 8x8x8 nodes has
// (v0, v1, v2, u0) at va[n] and (u1, u2) at ub[n], so a gather reads one b128
// and one b64 per node.
__device__ __forceinline__ void g2p_stage_store(float4* va, float2* ub, int tg, int l32, const float (&r)[kG2PStage]) {
    float* fa = reinterpret_cast<float*>(va);
    float* fb = reinterpret_cast<float*>(ub);
#pragma unroll
    for (int i = 0; i < kG2PStage; ++i) {
        const int e = l32 + 32 * i, c = e >> 6, nn = e & 63;
        const int X = (tg & 1) * 4 + (nn & 3), Y = ((tg >> 1) & 1) * 4 + ((nn >> 2) & 3), Z = (tg >> 2) * 4 + (nn >> 4);
        const int node = Z * 64 + Y * 8 + X;
        if (c < 4)
            fa[node * 4 + c] = r[i];
        else
            fb[node * 2 + c - 4] = r[i];
    }
}

// Fused stress_update + G2P + particle collisions + advection.  One workgroup
// per home block (the grid covers the home blocks, so the dispatcher balances
// them over the CUs) stages the 8x8x8 nodes of its 8 blocks (v and u) in LDS,
// then each thread updates one particle and writes it to the next buffer at
// its sorted position.
// Four waves per SIMD: the register budget (<= 128 VGPRs) is met without
// spills, and eight workgroups per CU hide each other's staging and particle
// loads.
template <int kG2PThreads, bool TRACK, bool PERMUTE, bool PROJECT>
__global__ __launch_bounds__(kG2PThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_g2p(StepArgs a) {
    static_assert(kG2PThreads % 32 == 0 && 8 % (kG2PThreads / 32) == 0, "neighbour blocks per 32-thread group");
    __shared__ float4 va[512];  // [Z][Y][X]: v0 v1 v2 u0
    __shared__ float2 ub[512];  // [Z][Y][X]: u1 u2
    // The block's scene (parameters and colliders), copied to LDS once per
    // block.  Read from global memory after the particle stores, every
    // scene or collider access would be a vector load whose wait
    // (vmcnt(0)) also waits for those stores to be acknowledged; G2P spent
    // ~90 us of 0.45 ms on the 8M slab in such waits in the collider loop.
    __shared__ SceneDev ssc;
    const uint32_t nh = a.cnt->n_home;
    const bool pad = PERMUTE && block_ranges_aligned(a, nh);
    // Threads 32 tg .. 32 tg + 31 stage neighbour blocks tg, tg + T/32, ...
    // (T threads per workgroup).
    constexpr int kGroups = kG2PThreads / 32;
    const uint32_t G = gridDim.x;
    const int tg = (int)(threadIdx.x >> 5), l32 = (int)(threadIdx.x & 31);
    for (uint32_t blk = blockIdx.x; blk < nh; blk += G) {
        const uint32_t slot = a.g.home_list[blk];
        uint32_t sid;
        int bx, by, bz;
        unpack_key(a.g.keys[slot], sid, bx, by, bz);
        // The scene record is loaded into registers with the staging loads and
        // written to LDS after the barrier (kSW words per thread).
        constexpr int kSceneWords = (int)(sizeof(SceneDev) / 4);
        constexpr int kSW = (kSceneWords + kG2PThreads - 1) / kG2PThreads;
        const int* ssrc = reinterpret_cast<const int*>(&a.scenes[sid]);
        int sw[kSW];
#pragma unroll
        for (int k = 0; k < kSW; ++k) {
            const int t = (int)threadIdx.x + k * kG2PThreads;
            sw[k] = t < kSceneWords ? ssrc[t] : 0;
        }
        {
            float stg[8 / kGroups][kG2PStage];
#pragma unroll
            for (int k = 0; k < 8 / kGroups; ++k)
                g2p_stage_load(a, a.g.nb[(size_t)slot * 8 + tg + k * kGroups], l32, stg[k]);
            __syncthreads();  // the previous block's gathers are done
#pragma unroll
            for (int k = 0; k < 8 / kGroups; ++k) g2p_stage_store(va, ub, tg + k * kGroups, l32, stg[k]);
        }
        {  // after the barrier above: no thread still reads the previous block's copy
            int* dst = reinterpret_cast<int*>(&ssc);
#pragma unroll
            for (int k = 0; k < kSW; ++k) {
                const int t = (int)threadIdx.x + k * kG2PThreads;
                if (t < kSceneWords) dst[t] = sw[k];
            }
        }
        const SceneDev& sc = ssc;
        if (threadIdx.x < 64) a.g.cellcnt[(size_t)slot * 64 + threadIdx.x] = 0u;
        __syncthreads();
        const uint32_t off = a.g.blkoff[slot], cnt = a.g.blkcnt[slot];
        const float inv_dx = sc.inv_dx, dt = sc.dt, alpha = sc.alpha;
        // The next particle's sorted index is loaded before this particle's
        // stores are issued, so waiting for it does not wait for them.
        const uint32_t jend = off + cnt;
        uint32_t i_n = off + threadIdx.x < jend ? at(a.bin.sidx, off + threadIdx.x) : 0u;
        for (uint32_t j = off + threadIdx.x; j < jend; j += kG2PThreads) {
            const uint32_t i = i_n;
            if (j + kG2PThreads < jend) i_n = at(a.bin.sidx, j + kG2PThreads);
            const float x0 = comp(a.cur.x, a.cur, 0, i), x1 = comp(a.cur.x, a.cur, 1, i), x2 = comp(a.cur.x, a.cur, 2, i);
            float v0 = comp(a.cur.v, a.cur, 0, i), v1 = comp(a.cur.v, a.cur, 1, i), v2 = comp(a.cur.v, a.cur, 2, i);
            int b[3];
            float f[3];
            base_node(sc, x0, x1, x2, b, f);
            int lx = b[0] - 4 * bx, ly = b[1] - 4 * by, lz = b[2] - 4 * bz;
            if ((unsigned)lx > 3u || (unsigned)ly > 3u || (unsigned)lz > 3u) {
                flag(a.st, sid, kStatusRange);
                lx = min(max(lx, 0), 3);
                ly = min(max(ly, 0), 3);
                lz = min(max(lz, 0), 3);
            }
            // The deformation state is loaded before the gather so its latency
            // hides under the gather.
            float Fe[9];
#pragma unroll
            for (int e = 0; e < 9; ++e) Fe[e] = comp(a.cur.F, a.cur, e, i);
            const float Jp = at(a.cur.Jp, i);
            const float V0 = at(a.cur.V0, i);
            float mu0, la0;
            lame(sc, a.cur, i, mu0, la0);
            float wx[4], dwx[4], wy[4], dwy[4], wz[4], dwz[4];
            bspline(f[0], wx, dwx);
            bspline(f[1], wy, dwy);
            bspline(f[2], wz, dwz);
            float pv0 = 0.f, pv1 = 0.f, pv2 = 0.f;
            float L[9];
#pragma unroll
            for (int e = 0; e < 9; ++e) L[e] = 0.f;
#pragma unroll 1
            for (int dd = 0; dd < 4; ++dd) {
                const float wzd = dd == 0 ? wz[0] : (dd == 1 ? wz[1] : (dd == 2 ? wz[2] : wz[3]));
                const float dwzd = dd == 0 ? dwz[0] : (dd == 1 ? dwz[1] : (dd == 2 ? dwz[2] : dwz[3]));
#pragma unroll
                for (int bb = 0; bb < 4; ++bb) {
                    const int rowb = (lz + dd) * 64 + (ly + bb) * 8 + lx;
                    float S0 = 0.f, S1 = 0.f, S2 = 0.f, X0 = 0.f, X1 = 0.f, X2 = 0.f, U0 = 0.f, U1 = 0.f, U2 = 0.f;
#pragma unroll
                    for (int aa = 0; aa < 4; ++aa) {
                        const float4 gv = va[rowb + aa];
                        const float2 gu = ub[rowb + aa];
                        const float gv0 = gv.x, gv1 = gv.y, gv2 = gv.z, gu0 = gv.w, gu1 = gu.x, gu2 = gu.y;
                        S0 += wx[aa] * gv0;
                        S1 += wx[aa] * gv1;
                        S2 += wx[aa] * gv2;
                        X0 += dwx[aa] * gv0;
                        X1 += dwx[aa] * gv1;
                        X2 += dwx[aa] * gv2;
                        U0 += wx[aa] * gu0;
                        U1 += wx[aa] * gu1;
                        U2 += wx[aa] * gu2;
                    }
                    const float yz = wy[bb] * wzd, dyz = dwy[bb] * wzd, ydz = wy[bb] * dwzd;
                    pv0 += yz * U0;
                    pv1 += yz * U1;
                    pv2 += yz * U2;
                    L[0] += yz * X0;
                    L[1] += dyz * S0;
                    L[2] += ydz * S0;
                    L[3] += yz * X1;
                    L[4] += dyz * S1;
                    L[5] += ydz * S1;
                    L[6] += yz * X2;
                    L[7] += dyz * S2;
                    L[8] += ydz * S2;
                }
            }
#pragma unroll
            for (int e = 0; e < 9; ++e) L[e] *= inv_dx;

            float Fp_old[9], Fp_new[9];
            constexpr bool track = TRACK;
#pragma unroll
            for (int e = 0; e < 9; ++e) Fp_old[e] = TRACK ? comp(a.cur.Fp, a.cur, e, i) : 0.f;
            PlastOut po;
            float Fh[9];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    Fh[r * 3 + c] = Fe[r * 3 + c] + dt * (L[r * 3 + 0] * Fe[0 * 3 + c] + L[r * 3 + 1] * Fe[1 * 3 + c] +
                                                          L[r * 3 + 2] * Fe[2 * 3 + c]);
            Sym Es;
            const bool elastic = elastic_trial(Fh, sc.theta_c, sc.theta_s, Es);
            bool fb = false;
            // Outputs go to sorted position j of the next buffers on a
            // permuting pass, else to the particle's own index i.
            const PBuf& o = a.nxt;
            const uint32_t w = PERMUTE ? j : i;
            // J_P and F_P change only in the plastic update (a deferred
            // particle's are written by the FP64 fix-up): outside re-sort steps
            // the output view shares their buffer with the input, so they are
            // written only where the update changed them (stored here, in the
            // branch that forms them, to keep their registers short-lived).
            if (elastic) {
                const float h = expf(sc.hardening * (1.0f - Jp));
#pragma unroll
                for (int e = 0; e < 9; ++e) {
                    po.Fe[e] = Fh[e];
                    Fp_new[e] = Fp_old[e];
                }
                po.Jp = Jp;
                if (PERMUTE) {
                    at(o.Jp, w) = Jp;
                    if (track)
#pragma unroll
                        for (int e = 0; e < 9; ++e) comp(o.Fp, o, e, w) = Fp_old[e];
                }
                elastic_stress(Es, det3<float>(Fh), mu0 * h, la0 * h, V0, po.S);
            } else {
                fb = plastic_update<TRACK>(Fe, L, dt, Jp, Fp_old, Fp_new, true, sc.theta_c, sc.theta_s, mu0, la0,
                                           sc.hardening, V0, a.kappa, po, true);
                if (PERMUTE || !fb) {
                    at(o.Jp, w) = fb ? Jp : po.Jp;
                    if (track)
#pragma unroll
                        for (int e = 0; e < 9; ++e) comp(o.Fp, o, e, w) = fb ? Fp_old[e] : Fp_new[e];
                }
            }
            if (fb) {
                // Redone in FP64 at the start of k_end; keep finite placeholders meanwhile.
                const uint32_t q = atomicAdd(&a.cnt->n_fallback, 1u);
                a.fb.pos[q] = j;
#pragma unroll
                for (int e = 0; e < 9; ++e) {
                    a.fb.L[(size_t)q * 9 + e] = L[e];
                    po.Fe[e] = Fe[e];
                }
                po.Jp = Jp;
#pragma unroll
                for (int e = 0; e < 6; ++e) po.S[e] = 0.f;
                if (track)
#pragma unroll
                    for (int e = 0; e < 9; ++e) Fp_new[e] = Fp_old[e];
            }

            // Permuting pass: the particle moves to sorted position j of the
            // next buffers with all its fields.  Otherwise the moving fields
            // are written at the particle's own index i into the other buffer
            // of the moving fields (which never aliases the input).  The
            // deformation state is stored first, which frees its registers
            // before the velocity update.
#pragma unroll
            for (int e = 0; e < 9; ++e) comp(o.F, o, e, w) = po.Fe[e];
#pragma unroll
            for (int e = 0; e < 6; ++e) comp(o.S, o, e, w) = po.S[e];

            // v_p = (1-alpha) v_pic + alpha (v_p + sum w (v_i - p_i/m_i))  (grid.cpp:321-324)
            v0 = alpha * v0 + pv0;
            v1 = alpha * v1 + pv1;
            v2 = alpha * v2 + pv2;
            collide<float>(sc, x0, x1, x2, v0, v1, v2);           // grid.cpp:327-333, pre-advection x
            float nx0 = x0 + dt * v0, nx1 = x1 + dt * v1, nx2 = x2 + dt * v2;  // grid.cpp:335-338
            if (PROJECT && sc.nproject) project<float>(sc, nx0, nx1, nx2);
            if (!finite3(nx0, nx1, nx2) || !finite3(v0, v1, v2) || !isfinite(po.Jp)) flag(a.st, sid, kStatusNaN);

            comp(o.x, o, 0, w) = nx0;
            comp(o.x, o, 1, w) = nx1;
            comp(o.x, o, 2, w) = nx2;
            comp(o.v, o, 0, w) = v0;
            comp(o.v, o, 1, w) = v1;
            comp(o.v, o, 2, w) = v2;
            if (PERMUTE) {
                at(a.nxt.m, j) = at(a.cur.m, i);
                at(a.nxt.V0, j) = V0;
                if (a.copy_E) at(a.nxt.E, j) = at(a.cur.E, i);
                at(a.nxt.pid, j) = at(a.cur.pid, i);
                at(a.nxt.sid, j) = sid;
            }
        }
        if (PERMUTE && pad) {  // the slots up to the next 32-particle boundary stay holes
            const uint32_t hole = jend + threadIdx.x;
            if (hole < off + ((cnt + 31u) & ~31u)) at(a.nxt.sid, hole) = kInvalid;
        }
    }
}

// FP64 redo of the plasticity update for the particles G2P deferred
// (inverted, singular or cond(trial F_E) > kappa).  Reads the pre-step
// state through the sorted index (G2P left the old F_E, J_P and F_P there
// for these particles) and writes the post-step state where G2P wrote.
__device__ __forceinline__ void fp64_fixup(const StepArgs& a) {
    const uint32_t nf = a.cnt->n_fallback;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nf; q += gridDim.x * blockDim.x) {
        const uint32_t j = a.fb.pos[q];
        const uint32_t i = a.bin.sidx[j];
        const uint32_t sid = a.cur.sid[i];
        const SceneDev& sc = a.scenes[sid];
        float Fe[9], L[9], Fp_old[9], Fp_new[9];
        for (int e = 0; e < 9; ++e) {
            Fe[e] = a.cur.F[e][i];
            L[e] = a.fb.L[(size_t)q * 9 + e];
        }
        const bool track = a.cur.Fp[0] != nullptr;
        if (track)
            for (int e = 0; e < 9; ++e) Fp_old[e] = a.cur.Fp[e][i];
        float mu0, la0;
        lame(sc, a.cur, i, mu0, la0);
        PlastOut po;
        const PBuf& o = a.nxt;
        const uint32_t w = a.permute ? j : i;
        plastic_fp64(Fe, L, sc.dt, a.cur.Jp[i], track ? Fp_old : nullptr, Fp_new, true, sc.theta_c, sc.theta_s, mu0,
                     la0, sc.hardening, a.cur.V0[i], po);
        for (int e = 0; e < 9; ++e) o.F[e][w] = po.Fe[e];
        o.Jp[w] = po.Jp;
        for (int e = 0; e < 6; ++e) o.S[e][w] = po.S[e];
        if (track)
            for (int e = 0; e < 9; ++e) o.Fp[e][w] = Fp_new[e];
        if (!isfinite(po.Jp)) flag(a.st, sid, kStatusNaN);
        atomicAdd(&a.st.fp64_count[sid], 1ull);
    }
}

// End of a pass: the FP64 fix-ups, then clear the hash slots used this pass;
// after a permuting pass publish the new (compacted) particle count.  (The
// fix-ups touch only particle buffers, the clearing only the grid.)
__global__ __launch_bounds__(256) void k_end(StepArgs a, int publish) {
    fp64_fixup(a);
    const uint32_t na = a.cnt->n_home + a.cnt->n_extra;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < na; b += gridDim.x * blockDim.x) {
        const uint32_t slot = a.g.all_list[b];
        a.g.keys[slot] = kEmptyKey;
        uint4* r = reinterpret_cast<uint4*>(a.g.rev + (size_t)slot * 8);
        r[0] = make_uint4(kInvalid, kInvalid, kInvalid, kInvalid);
        r[1] = make_uint4(kInvalid, kInvalid, kInvalid, kInvalid);
    }
    if (publish && blockIdx.x == 0 && threadIdx.x == 0) a.cnt->n_particles = a.cnt->cursor;
}

// ---------------------------------------------------------------------------
// Setup (grid.cpp:340-357): after a mass-only P2G, rho_p = sum_i m_i w_ip / dx^3
// and V_p0 = m_p / rho_p (calculate_volume, grid.cpp:138-160); then the
// particle's force matrix for the first step.  Writes in place.
__global__ __launch_bounds__(256) void k_setup_gather(StepArgs a) {
    __shared__ float ms[512];
    const uint32_t nh = a.cnt->n_home;
    for (uint32_t blk = blockIdx.x; blk < nh; blk += gridDim.x) {
        const uint32_t slot = a.g.home_list[blk];
        uint32_t sid;
        int bx, by, bz;
        unpack_key(a.g.keys[slot], sid, bx, by, bz);
        const SceneDev& sc = a.scenes[sid];
        for (int t = threadIdx.x; t < 8 * 64; t += 256) {
            const int tgt = t >> 6, nn = t & 63;
            const uint32_t ts = a.g.nb[(size_t)slot * 8 + tgt];
            const int X = (tgt & 1) * 4 + (nn & 3), Y = ((tgt >> 1) & 1) * 4 + ((nn >> 2) & 3),
                      Z = (tgt >> 2) * 4 + (nn >> 4);
            ms[Z * 64 + Y * 8 + X] = a.g.vu[((size_t)ts * kVelCh + 0) * 64 + nn];
        }
        if (threadIdx.x < 64) a.g.cellcnt[(size_t)slot * 64 + threadIdx.x] = 0u;
        __syncthreads();
        const uint32_t off = a.g.blkoff[slot], cnt = a.g.blkcnt[slot];
        const float idx3 = sc.inv_dx * sc.inv_dx * sc.inv_dx;
        for (uint32_t j = off + threadIdx.x; j < off + cnt; j += 256) {
            const uint32_t i = a.bin.sidx[j];
            int b[3];
            float f[3];
            base_node(sc, a.cur.x[0][i], a.cur.x[1][i], a.cur.x[2][i], b, f);
            const int lx = min(max(b[0] - 4 * bx, 0), 3), ly = min(max(b[1] - 4 * by, 0), 3),
                      lz = min(max(b[2] - 4 * bz, 0), 3);
            float wx[4], dwx[4], wy[4], dwy[4], wz[4], dwz[4];
            bspline(f[0], wx, dwx);
            bspline(f[1], wy, dwy);
            bspline(f[2], wz, dwz);
            float rho = 0.f;
#pragma unroll
            for (int dd = 0; dd < 4; ++dd)
#pragma unroll
                for (int bb = 0; bb < 4; ++bb) {
                    const int rowb = (lz + dd) * 64 + (ly + bb) * 8 + lx;
                    float s = 0.f;
#pragma unroll
                    for (int aa = 0; aa < 4; ++aa) s += wx[aa] * ms[rowb + aa];
                    rho += wy[bb] * wz[dd] * s;
                }
            rho *= idx3;
            const float V0 = a.cur.m[i] / rho;
            a.cur.V0[i] = V0;
            float Fe[9], L[9];
#pragma unroll
            for (int e = 0; e < 9; ++e) {
                Fe[e] = a.cur.F[e][i];
                L[e] = 0.f;
            }
            float mu0, la0;
            lame(sc, a.cur, i, mu0, la0);
            PlastOut po;
            float dummy[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, dummy2[9];
            plastic_update<false>(Fe, L, 0.f, a.cur.Jp[i], dummy, dummy2, false, 0.f, 0.f, mu0, la0, sc.hardening,
                                  V0, a.kappa, po);
#pragma unroll
            for (int e = 0; e < 6; ++e) a.cur.S[e][i] = po.S[e];
        }
        __syncthreads();
    }
}

// Force matrix from the stored F_E (force_constitutive, grid.cpp:162-176) for
// particles [first, first+n) of the current buffer (after an upload).
__global__ __launch_bounds__(256) void k_init_stress(StepArgs a, uint32_t first, uint32_t n) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const uint32_t i = first + t;
        const SceneDev& sc = a.scenes[a.cur.sid[i]];
        float Fe[9], L[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            Fe[e] = a.cur.F[e][i];
            L[e] = 0.f;
        }
        float mu0, la0;
        lame(sc, a.cur, i, mu0, la0);
        PlastOut po;
        float dummy[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, dummy2[9];
        plastic_update<false>(Fe, L, 0.f, a.cur.Jp[i], dummy, dummy2, false, 0.f, 0.f, mu0, la0, sc.hardening,
                              a.cur.V0[i], a.kappa, po);
#pragma unroll
        for (int e = 0; e < 6; ++e) a.cur.S[e][i] = po.S[e];
    }
}

// Gathers scene `sid`'s particles into pid order (download staging).
// A new scene's particles into the FP32 buffers at first .. first + n - 1
// (pid = index in the scene).  J_P = det F_P is formed in double exactly as
// written (no contraction), F_P kept only with the full-F_P layout.
__global__ __launch_bounds__(256) void k_put_scene(StepArgs a, uint32_t first, uint32_t n, uint32_t sid, SceneIn in) {
#pragma clang fp contract(off)
    const bool track = a.cur.Fp[0] != nullptr;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t d = first + i;
        for (int e = 0; e < 3; ++e) {
            comp(a.cur.x, a.cur, e, d) = (float)in.x[(size_t)i * 3 + e];
            comp(a.cur.v, a.cur, e, d) = (float)in.v[(size_t)i * 3 + e];
        }
        double P[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        if (in.Fp)
            for (int e = 0; e < 9; ++e) P[e] = in.Fp[(size_t)i * 9 + e];
        for (int e = 0; e < 9; ++e) {
            comp(a.cur.F, a.cur, e, d) = in.Fe ? (float)in.Fe[(size_t)i * 9 + e] : (e % 4 == 0 ? 1.0f : 0.0f);
            if (track) comp(a.cur.Fp, a.cur, e, d) = (float)P[e];
        }
        const double jp = P[0] * (P[4] * P[8] - P[5] * P[7]) - P[1] * (P[3] * P[8] - P[5] * P[6]) +
                          P[2] * (P[3] * P[7] - P[4] * P[6]);
        at(a.cur.Jp, d) = (float)jp;
        at(a.cur.m, d) = (float)in.m[i];
        at(a.cur.V0, d) = in.V0 ? (float)in.V0[i] : 0.0f;
        at(a.cur.E, d) = in.E ? (float)in.E[i] : 0.0f;
        at(a.cur.pid, d) = i;
        at(a.cur.sid, d) = sid;
    }
}

// A scene's particles in their original order, in the reference's host
// layout (doubles; x and v 3 per particle, F_E and F_P 9 row-major, J_P and
// V_p0 1), written straight into the requested outputs (null: not wanted),
// so the host copies exactly the bytes the caller asked for.
__global__ __launch_bounds__(256) void k_gather_scene(StepArgs a, uint32_t sid, uint32_t nscene, SceneOut o) {
    const uint32_t n = a.cnt->n_particles;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (at(a.cur.sid, i) != sid) continue;
        const uint32_t p = at(a.cur.pid, i);
        if (p >= nscene) continue;
        if (o.x)
            for (int e = 0; e < 3; ++e) o.x[(size_t)p * 3 + e] = comp(a.cur.x, a.cur, e, i);
        if (o.v)
            for (int e = 0; e < 3; ++e) o.v[(size_t)p * 3 + e] = comp(a.cur.v, a.cur, e, i);
        if (o.Fe)
            for (int e = 0; e < 9; ++e) o.Fe[(size_t)p * 9 + e] = comp(a.cur.F, a.cur, e, i);
        if (o.Fp)
            for (int e = 0; e < 9; ++e) o.Fp[(size_t)p * 9 + e] = comp(a.cur.Fp, a.cur, e, i);
        if (o.Jp) o.Jp[p] = at(a.cur.Jp, i);
        if (o.V0) o.V0[p] = at(a.cur.V0, i);
    }
}

__global__ __launch_bounds__(256) void k_fill_zero(double* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// FP64 conservation sums of one scene: sum m, sum m x (3), sum m v (3).
__global__ __launch_bounds__(256) void k_scene_sums(StepArgs a, uint32_t sid, double* out7) {
    __shared__ double red[7][256];
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    const uint32_t n = a.cnt->n_particles;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (a.cur.sid[i] != sid) continue;
        const double m = a.cur.m[i];
        s[0] += m;
        for (int e = 0; e < 3; ++e) {
            s[1 + e] += m * (double)a.cur.x[e][i];
            s[4 + e] += m * (double)a.cur.v[e][i];
        }
    }
    for (int e = 0; e < 7; ++e) red[e][threadIdx.x] = s[e];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int e = 0; e < 7; ++e) red[e][threadIdx.x] += red[e][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int e = 0; e < 7; ++e) out7[blockIdx.x * 7 + e] = red[e][0];
}

// ---------------------------------------------------------------------------
// Host launch wrappers.

// P2G runs one single-wave workgroup per home block when the grid covers
// them all: the dispatcher then balances the blocks across the CUs, which a
// persistent loop over blk, blk + G, ... does not (a workgroup that starts
// late and takes a second block becomes the tail).  Workgroups past the last
// home block exit at once.
inline int p2g_grid(const LaunchCfg& c) { return c.grid_w * 4 > c.grid_h ? c.grid_w * 4 : c.grid_h; }
// The grid kernel likewise gets a wave for every block (home blocks and
// their neighbours, about 1.4x the home blocks for a dense slab).
inline int grid_grid(const LaunchCfg& c) { return c.grid_b > c.grid_h / 2 ? c.grid_b : c.grid_h / 2; }
// G2P: a workgroup per home block as well.
inline int g2p_grid(const LaunchCfg& c) { return c.grid_w > c.grid_h ? c.grid_w : c.grid_h; }

// P2G with the staging chunk the launch configuration names (16, 32, 48 or
// 64).  48 is the default: 48 of a wave's 64 lanes stage particles and the
// records and plane buffer (9.2 KB of LDS) still let 16 waves share a CU.
inline void launch_p2g(const StepArgs& a, const LaunchCfg& c, hipStream_t s) {
    if (c.p2g_chunk == 16)
        k_p2g<16, false><<<p2g_grid(c), 64, 0, s>>>(a);
    else if (c.p2g_chunk == 48)
        k_p2g<48, false><<<p2g_grid(c), 64, 0, s>>>(a);
    else if (c.p2g_chunk == 64)
        k_p2g<64, false><<<p2g_grid(c), 64, 0, s>>>(a);
    else
        k_p2g<32, false><<<p2g_grid(c), 64, 0, s>>>(a);
}
void launch_p2g_force(const StepArgs& a, const LaunchCfg& c, hipStream_t s) {
    if (c.p2g_chunk == 16)
        k_p2g<16, true><<<p2g_grid(c), 64, 0, s>>>(a);
    else if (c.p2g_chunk == 48)
        k_p2g<48, true><<<p2g_grid(c), 64, 0, s>>>(a);
    else if (c.p2g_chunk == 64)
        k_p2g<64, true><<<p2g_grid(c), 64, 0, s>>>(a);
    else
        k_p2g<32, true><<<p2g_grid(c), 64, 0, s>>>(a);
}

// G2P with or without the position projection (a branch per particle only
// when some scene asks for it), at the launch configuration's workgroup size.
template <int T, bool TRACK, bool PERMUTE>
void launch_g2p_t(const StepArgs& a, const LaunchCfg& c, hipStream_t s) {
    if (a.project)
        k_g2p<T, TRACK, PERMUTE, true><<<g2p_grid(c), T, 0, s>>>(a);
    else
        k_g2p<T, TRACK, PERMUTE, false><<<g2p_grid(c), T, 0, s>>>(a);
}
template <bool TRACK, bool PERMUTE>
void launch_g2p(const StepArgs& a, const LaunchCfg& c, hipStream_t s) {
    if (c.g2p_threads == 64)
        launch_g2p_t<64, TRACK, PERMUTE>(a, c, s);
    else if (c.g2p_threads == 256)
        launch_g2p_t<256, TRACK, PERMUTE>(a, c, s);
    else
        launch_g2p_t<128, TRACK, PERMUTE>(a, c, s);
}

// G2P with the buffer layout's variant.
inline void launch_g2p_any(const StepArgs& a, const LaunchCfg& c, hipStream_t s) {
    if (a.cur.Fp[0]) {
        if (a.permute)
            launch_g2p<true, true>(a, c, s);
        else
            launch_g2p<true, false>(a, c, s);
    } else {
        if (a.permute)
            launch_g2p<false, true>(a, c, s);
        else
            launch_g2p<false, false>(a, c, s);
    }
}

void launch_step(const StepArgs& a, const LaunchCfg& c, hipStream_t s, hipEvent_t* ev) {
    if (ev) (void)hipEventRecord(ev[0], s);
    k_bin<<<c.grid_p, 256, 0, s>>>(a);
    k_list<<<c.grid_l, 1024, 0, s>>>(a);
    k_blocks<<<c.grid_k, 1024, 0, s>>>(a);
    k_scatter<<<c.grid_p, 256, 0, s>>>(a);
    k_cellsort<<<c.grid_b, 256, 0, s>>>(a);
    if (ev) (void)hipEventRecord(ev[1], s);
    launch_p2g(a, c, s);
    if (ev) (void)hipEventRecord(ev[2], s);
    k_grid<<<grid_grid(c), 256, 0, s>>>(a);
    if (ev) (void)hipEventRecord(ev[3], s);
    launch_g2p_any(a, c, s);
    if (ev) (void)hipEventRecord(ev[4], s);
    k_end<<<c.grid_b, 256, 0, s>>>(a, a.permute);
    if (ev) (void)hipEventRecord(ev[5], s);
}

void launch_step_head(const StepArgs& a, const LaunchCfg& c, hipStream_t s) {
    k_bin<<<c.grid_p, 256, 0, s>>>(a);
    k_list<<<c.grid_l, 1024, 0, s>>>(a);
    k_blocks<<<c.grid_k, 1024, 0, s>>>(a);
    k_scatter<<<c.grid_p, 256, 0, s>>>(a);
    k_cellsort<<<c.grid_b, 256, 0, s>>>(a);
    launch_p2g(a, c, s);
    k_grid<<<grid_grid(c), 256, 0, s>>>(a);
}

void launch_step_tail(const StepArgs& a, const LaunchCfg& c, hipStream_t s) {
    launch_g2p_any(a, c, s);
    k_end<<<c.grid_b, 256, 0, s>>>(a, a.permute);
}

void launch_setup(const StepArgs& a, const LaunchCfg& c, hipStream_t s) {
    k_bin<<<c.grid_p, 256, 0, s>>>(a);
    k_list<<<c.grid_l, 1024, 0, s>>>(a);
    k_blocks<<<c.grid_k, 1024, 0, s>>>(a);
    k_scatter<<<c.grid_p, 256, 0, s>>>(a);
    k_cellsort<<<c.grid_b, 256, 0, s>>>(a);
    launch_p2g(a, c, s);
    StepArgs am = a;
    am.mass_only = 1;
    k_grid<<<grid_grid(c), 256, 0, s>>>(am);
    k_setup_gather<<<c.grid_w, 256, 0, s>>>(a);
    k_end<<<c.grid_b, 256, 0, s>>>(a, 0);
}

void launch_init_stress(const StepArgs& a, uint32_t first, uint32_t n, hipStream_t s) {
    if (n == 0) return;
    const uint32_t g = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
    k_init_stress<<<g, 256, 0, s>>>(a, first, n);
}

void launch_set_count(Counters* c, uint32_t n, hipStream_t s) { k_set_count<<<1, 64, 0, s>>>(c, n); }

void launch_put_scene(const StepArgs& a, const LaunchCfg& c, uint32_t first, uint32_t n, uint32_t sid,
                      const SceneIn& in, hipStream_t s) {
    const uint32_t g = (n + 255) / 256 < (uint32_t)c.grid_p ? (n + 255) / 256 : (uint32_t)c.grid_p;
    k_put_scene<<<g, 256, 0, s>>>(a, first, n, sid, in);
}

void launch_gather_scene(const StepArgs& a, const LaunchCfg& c, uint32_t sid, uint32_t nscene, const SceneOut& o,
                         hipStream_t s) {
    k_gather_scene<<<c.grid_p, 256, 0, s>>>(a, sid, nscene, o);
}

void launch_fill_zero(double* p, size_t n, hipStream_t s) {
    if (n == 0) return;
    const size_t g = (n + 255) / 256 < 2048 ? (n + 255) / 256 : 2048;
    k_fill_zero<<<(unsigned)g, 256, 0, s>>>(p, n);
}

void launch_scene_sums(const StepArgs& a, uint32_t sid, double* out, int nblocks, hipStream_t s) {
    k_scene_sums<<<nblocks, 256, 0, s>>>(a, sid, out);
}


}  // namespace snowmpm
