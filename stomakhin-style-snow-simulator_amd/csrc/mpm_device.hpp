// mpm_device.hpp -- device-side data layout of the MI355X MPM step engine.
//
// Particles live in HBM as FP32 structure-of-arrays, double buffered.  G2P
// reads the current state through the block-sorted index and writes the
// fields a step changes to the other buffer at the same index; every few
// steps it instead writes whole particles in sorted order, so the buffers stay
// grouped by 4^3-node home block.  The grid is block sparse: a device hash maps a
// packed (scene, bx, by, bz) key to a slot, and the slot index is the block's
// storage index in every per-block array (no indirection, no compaction).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace snowmpm {

constexpr int kMaxColliders = 8;
constexpr int kMaxScenes = 8191;            // scene id 8191 is reserved (empty key)
constexpr uint64_t kEmptyKey = ~0ull;
constexpr int kCoordBits = 17;              // per axis, biased
constexpr int kCoordBias = 1 << 16;         // block coordinate range +-65534
constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr int kBlockNodes = 64;             // 4x4x4 nodes per block
constexpr int kGridCh = 7;                  // P2G channels: m, p(3), f(3)
constexpr int kRegion = 343;                // 7x7x7 nodes a home block's particles reach
// A home block's P2G partial sums, node plane by node plane: each of the 7
// planes holds its 4 xy pieces (4x4, 3x4, 4x3, 3x3 nodes) x 7 channels, each
// piece padded to a multiple of 32 floats so that it starts on a 128-byte
// line: 128 + 96 + 96 + 64 = 384 floats per plane.
constexpr int kPlaneFloats = 384;
constexpr int kPartFloats = 7 * kPlaneFloats;  // per home block
constexpr int kVelCh = 6;                   // FP32 outputs: v(3), u(3) (u = PIC/FLIP blend)
constexpr int kStressCh = 6;                // symmetric force matrix -V0 P F^T

// Status bits (per scene, OR-ed by kernels).
constexpr uint32_t kStatusNaN = 0x1;
constexpr uint32_t kStatusOverflow = 0x2;   // block list capacity exceeded
constexpr uint32_t kStatusHashFull = 0x4;
constexpr uint32_t kStatusRange = 0x8;

constexpr int kProjectFlag = 0x1;  // ColliderDev::flags, = MPM_COLLIDER_PROJECT

struct ColliderDev {
    int type;   // 0 plane, 1 box, 2 sphere
    int flags;  // kProjectFlag: project particles that end a step inside
    float point[3];
    float normal[3];
    float half_extent[3];
    float radius;
    float velocity[3];
    float friction;
};

struct SceneDev {
    float dx, inv_dx, dt;
    float gravity[3];
    float mu0, lambda0, nu;
    float mu_per_E, la_per_E;                // mu0 = E mu_per_E, lambda0 = E la_per_E
    float hardening, theta_c, theta_s, alpha;
    int use_E;                               // per-particle Young's modulus present
    int ncoll;
    int flags;                               // kSceneActive | kSceneSetup
    int nproject;                            // colliders with the projection flag
    float beta;                              // semi-implicit ratio (0: explicit step)
    float cr_tol2;                           // conjugate residual: stop at |r|_M^2 <= cr_tol2 |b|_M^2
    int cr_maxit;                            // and after at most this many iterations
    ColliderDev coll[kMaxColliders];
};
static_assert(sizeof(SceneDev) % 4 == 0, "scene records are copied as 32-bit words");
constexpr int kSceneActive = 1;
constexpr int kSceneSetup = 2;

// One particle buffer (SoA, FP32).  Fp is null unless the engine tracks the
// full plastic deformation gradient; otherwise J_P = det(F_P) carries it.
struct PBuf {
    float* x[3];
    float* v[3];
    float* F[9];   // elastic deformation gradient F_E, row major
    float* Jp;     // det F_P
    float* Fp[9];  // optional full F_P
    float* m;
    float* V0;
    float* E;      // per-particle Young's modulus (used if scene.use_E)
    float* S[kStressCh];  // force matrix for the next P2G: S00 S11 S22 S01 S02 S12
    uint32_t* pid; // particle index within its scene (download order, sort tie-break)
    uint32_t* sid; // scene id
    uint32_t stride;  // elements from one component of x, v, F, Fp, S to the next
};

// The fields every step changes (x v F_E S) and the others (m V0 E pid sid,
// and J_P F_P, which only the plastic update changes and an ordinary step
// updates in place) of a particle view may come from different buffers.
inline PBuf compose_pbuf(const PBuf& moving, const PBuf& constant) {
    PBuf p = moving;
    p.Jp = constant.Jp;
    for (int e = 0; e < 9; ++e) p.Fp[e] = constant.Fp[e];
    p.m = constant.m;
    p.V0 = constant.V0;
    p.E = constant.E;
    p.pid = constant.pid;
    p.sid = constant.sid;
    return p;
}

struct Counters {
    uint32_t n_particles;  // particles in the current buffer
    uint32_t n_home;       // blocks holding particles
    uint32_t n_extra;      // neighbour blocks k_blocks inserted (all_list = home list, then these)
    uint32_t cursor;       // particles binned this pass
    uint32_t n_fallback;   // particles deferred to the FP64 fix-up this pass
    uint32_t global_status;
    uint32_t pad_[2];
};
// k_blocks bumps n_extra and cursor with one 64-bit atomic.
static_assert(offsetof(Counters, n_extra) % 8 == 0 && offsetof(Counters, cursor) == offsetof(Counters, n_extra) + 4,
              "n_extra and cursor form one aligned 64-bit word");

struct GridDev {
    uint64_t* keys;        // [H] hash keys, kEmptyKey when free
    uint32_t mask;         // H - 1
    uint32_t max_probe;
    uint32_t* cellcnt;     // [H*64] particles per home cell (zero between passes)
    uint32_t* celloff;     // [H*64] exclusive offsets within the block
    uint32_t* blkoff;      // [H] first sorted position of the block
    uint32_t* blkcnt;      // [H] particles in the block
    uint32_t* nb;          // [H*8] slots of the block + {0,1}^3 neighbours
    float* part;           // [max_home*7*343] each home block's P2G partial sums
    uint32_t* rev;         // [H*8] home-list index of block - {0,1}^3, or kInvalid
    uint32_t max_home;     // capacity of part (home blocks per pass)
    uint32_t pad_;
    float* vu;             // [H*6*64] grid velocity and PIC/FLIP blend
    uint32_t* home_list;   // [H]
    uint32_t* all_list;    // [H] home blocks (same order as home_list), then the other neighbour blocks
};

struct BinDev {
    // [S]: one entry per particle slot of a buffer (S = capacity + the room
    // for the block-aligned layout's holes), by particle index or sorted position
    uint32_t* pkey;        // [S] slot<<6 | cell, or kInvalid
    uint32_t* prank;       // [S] rank within cell
    uint32_t* sidx;        // [S] sorted position -> particle index
    uint32_t* spid;        // [S] sorted position -> pid (sort key inside a cell)
};

struct FallbackDev {
    uint32_t* pos;         // [capacity] sorted positions whose update is redone in FP64
    float* L;              // [capacity*9] their velocity gradients
};

// Semi-implicit solve state (allocated on the first step of a scene with
// beta > 0).  Node vectors share vu's layout: [slot][channel 3][node 64].
constexpr int kCRVecs = 5;  // X, R, P, AR, AP
struct CRScene {            // per scene, in FP64
    double bb, rr, rho, den, alpha, beta;
    int active, iters;
};
struct ImpDev {
    float* gm;              // [H*64] node mass of the unknowns (0: massless or held by a collider)
    float* vec[kCRVecs];    // [H*3*64] each
    float* G;               // [9][S] per-particle -V0 dP F^T for the force differential
    uint32_t* ord;          // [nblocks] slots of all blocks in key order
    uint64_t* skey;         // [nblocks] their keys
    double* dp;             // [nblocks*4] per-block dot-product partials (key order)
    uint32_t* seg;          // [2*kMaxScenes] per scene: first, end index into ord
    CRScene* cr;            // [kMaxScenes]
    double* hist;           // [kMaxScenes*kCRHist] |r|_M per iteration of the last solve
    uint32_t* nactive;      // [kCRHist + 1] scenes still iterating after each reduction of a solve
};
constexpr int kCRHist = 257;

struct StatusDev {
    uint32_t* scene_status;          // [kMaxScenes]
    unsigned long long* fp64_count;  // [kMaxScenes] FP64 SVD fallbacks
};

__host__ __device__ inline uint64_t pack_key(uint32_t sid, int bx, int by, int bz) {
    return ((uint64_t)sid << 51) | ((uint64_t)(uint32_t)(bx + kCoordBias) << 34) |
           ((uint64_t)(uint32_t)(by + kCoordBias) << 17) | (uint64_t)(uint32_t)(bz + kCoordBias);
}
__host__ __device__ inline void unpack_key(uint64_t k, uint32_t& sid, int& bx, int& by, int& bz) {
    const uint64_t m = (1ull << kCoordBits) - 1;
    sid = (uint32_t)(k >> 51);
    bx = (int)((k >> 34) & m) - kCoordBias;
    by = (int)((k >> 17) & m) - kCoordBias;
    bz = (int)(k & m) - kCoordBias;
}

}  // namespace snowmpm
