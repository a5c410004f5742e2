/*
 * snowmpm.h -- C ABI of the MI355X MPM step engine (libsnowmpm.so).
 *
 * Drop-in boundary for the reference solver loop.  Every entry point returns
 * an int status (MPM_OK = 0) unless stated; plain pointers and sizes only.
 * Host arrays are double precision in the reference's per-particle layout
 * (Eigen::Vector3d / row-major Matrix3d per particle, contiguous).
 *
 * Reference interfaces replaced (paths relative to the reference repo):
 *   setup()        include/snow/grid.hpp:96,  src/grid.cpp:340-357
 *   step()         include/snow/grid.hpp:97,  src/grid.cpp:359-450
 *   make_plane()   include/snow/collision.hpp:14, src/collision.cpp:5-27
 *   StepTimings    include/snow/debug.hpp:3-15
 *   Particle       include/snow/particle.hpp:5-28 (state layout of the arrays)
 *   import_*()     include/snow/import.hpp:22-25, src/import.cpp:11-137
 *   export_particles() include/snow/export.hpp:8, src/export.cpp:10-22
 */
#ifndef SNOWMPM_H
#define SNOWMPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define MPM_OK 0
#define MPM_ERR_ARG (-1)      /* bad argument (null pointer, bad id, bad size) */
#define MPM_ERR_HIP (-2)      /* HIP runtime error */
#define MPM_ERR_OOM (-3)      /* device allocation failed */
#define MPM_ERR_CAPACITY (-4) /* particle capacity or scene table exhausted */
#define MPM_ERR_STATE (-5)    /* call not valid in the engine's state */
#define MPM_ERR_DEVICE (-6)   /* a kernel raised a status bit (see mpm_scene_status) */
#define MPM_ERR_IO (-7)       /* file could not be read/written or was malformed */

/* device status bits, per scene */
#define MPM_STATUS_NAN 0x1u       /* non-finite particle state */
#define MPM_STATUS_OVERFLOW 0x2u  /* more occupied blocks than the P2G partial-sum buffer holds */
#define MPM_STATUS_HASH_FULL 0x4u /* block hash table full */
#define MPM_STATUS_RANGE 0x8u     /* particle outside the block-coordinate range */
#define MPM_STATUS_REJECTED 0x100u /* serving: request larger than the engine's particle pool (host-side) */

typedef struct mpm_engine mpm_engine;

/* The scalar arguments of step() (grid.hpp:97), plus nu for per-particle E. */
typedef struct {
    double gravity[3];
    double dt;              /* time_step */
    double dx;              /* grid_spacing */
    double mu0;             /* mu_0 */
    double lambda0;         /* lambda_0 */
    double hardening;       /* hardening_coefficient (xi) */
    double max_compression; /* theta_c */
    double max_stretch;     /* theta_s */
    double alpha;           /* FLIP ratio */
    double nu;              /* Poisson ratio used with per-particle Young's moduli */
    /* Semi-implicit velocity solve, declared but never defined by the
     * reference (grid.hpp:77, 81, 89-92): beta = 0 is the explicit step;
     * beta > 0 solves (I + beta dt^2 M^-1 H) v = v* by conjugate residuals
     * until |r|_M <= cr_tolerance |v*|_M or cr_max_iterations (<= 256). */
    double beta;
    double cr_tolerance;    /* 0 = 1e-6 */
    int cr_max_iterations;  /* 0 = 100 */
} mpm_params;

/* POD collider (replaces CollisionBody, collision.hpp:6-12).  type 0 is the
 * reference's plane (make_plane, collision.cpp:5-27); 1 a solid axis-aligned
 * box; 2 a solid sphere.  velocity: the body's constant velocity.
 * flags: MPM_COLLIDER_PROJECT moves a particle that ends a step inside the
 * body (phi < 0 after advection) onto its surface, x -= phi n.  It has no
 * counterpart in the reference, whose collisions change velocities only
 * (grid.cpp:327-338); the factories leave flags 0, the reference behaviour. */
#define MPM_COLLIDER_PLANE 0
#define MPM_COLLIDER_BOX 1
#define MPM_COLLIDER_SPHERE 2
#define MPM_COLLIDER_PROJECT 0x1
typedef struct {
    int type;
    int flags;
    double point[3];
    double normal[3];
    double half_extent[3];
    double radius;
    double velocity[3];
    double friction;
} mpm_collider;

/* StepTimings (debug.hpp:3-15), milliseconds, accumulated with +=.  Fused
 * kernels report under the first reference stage they absorb:
 *   stencil       <- block binning, sort and hash clear
 *   p2g           <- P2G + force_constitutive + force_grid_accumulation
 *   grid_velocity <- grid_velocity + grid_collisions
 *   g2p           <- stress_update + G2P + particle_collisions + position_update
 * and the absorbed stages read 0. */
typedef struct {
    double stencil, p2g, forces_constitutive, forces_accumulative, grid_velocity, grid_collisions, stress_update,
        g2p, particle_collisions, position_update;
} mpm_timings;

typedef struct {
    size_t particle_capacity; /* total particles of all resident scenes, < 2^30 */
    int hash_log2;            /* block hash slots = 2^hash_log2; 0 = sized from capacity */
    int track_fp;             /* 1: keep the full plastic gradient F_P; 0: keep J_P = det F_P */
    int device;               /* HIP device ordinal */
    int permute_period;       /* steps between full re-sorts of the particle buffers; 0 = 16.
                                 Results do not depend on it (steps in between update the
                                 particles at their index); 1 re-sorts every step.  Fast-moving
                                 scenes run ~3% faster at 4, scenes at rest ~3% slower */
    double kappa;             /* FP64 fallback when cond(trial F_E) > kappa; 0 = 1.5 */
    int p2g_chunk;            /* particles P2G stages per pass: 16, 32, 48 or 64; 0 = 48 */
    int g2p_threads;          /* G2P workgroup size: 64, 128 or 256; 0 = 128.
                                 Results depend on neither; mpm_autotune chooses both */
    int dense_blocks;         /* 0 (default): the buffers hold capacity + capacity/4 slots and
                                 each home block's particles take a range that starts on a
                                 32-particle (128-byte) boundary whenever the holes (at most 31
                                 per home block) fit in the extra quarter; a re-sort leaves them
                                 in the buffers.  1: capacity slots, ranges packed end to end.
                                 Results do not depend on it */
} mpm_config;

/* ---- engine lifetime ---- */
int mpm_create(const mpm_config* cfg, mpm_engine** out);
int mpm_destroy(mpm_engine* e);
const char* mpm_last_error(void);

/* ---- scenes (a scene = one reference std::vector<Particle> + colliders) ----
 * x, v: n*3; m: n; V0: n or NULL (then mpm_setup computes it, as setup());
 * Fe, Fp: n*9 row-major or NULL (identity, particle.hpp:24-25);
 * E: per-particle Young's modulus or NULL (use prm->mu0/lambda0). */
int mpm_scene_add(mpm_engine* e, const mpm_params* prm, const mpm_collider* colliders, int ncolliders, size_t n,
                  const double* x, const double* v, const double* m, const double* V0, const double* Fe,
                  const double* Fp, const double* E, int* scene_id);
int mpm_scene_remove(mpm_engine* e, int scene);
/* Drop every scene and empty the particle buffer at once. */
int mpm_clear(mpm_engine* e);
int mpm_scene_count(mpm_engine* e, int scene, size_t* n);

/* setup(): V_p0 for every scene added without V0 (grid.cpp:340-357). */
int mpm_setup(mpm_engine* e);

/* nsteps calls of step() on every active scene (grid.cpp:359-450).
 * timings may be NULL; when given, per-stage times are added to it. */
int mpm_step(mpm_engine* e, int nsteps, mpm_timings* timings);

/* Copy a scene's state back in its original particle order.  Any output may be
 * NULL.  Fp is only available with track_fp; Jp = det F_P always is. */
int mpm_scene_download(mpm_engine* e, int scene, double* x, double* v, double* Fe, double* Fp, double* Jp,
                       double* V0);

/* mpm_scene_download without device allocation.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid allocation): the arrays are staged
 * through a static device buffer of *max_bytes (32 MiB) that each GPU holds
 * for the life of the process, so a request needing more staging (26 doubles
 * per particle with every field, 3 with x alone) fails with
 * MPM_ERR_CAPACITY.  Calls that overlap on one GPU share that buffer and
 * race.  max_bytes may be NULL. */
int mpm_scene_download_noalloc(mpm_engine* e, int scene, double* x, double* v, double* Fe, double* Fp, double* Jp,
                               double* V0, size_t* max_bytes);

/* Status bits OR-ed since the scene was added, and FP64 SVD fallback count. */
int mpm_scene_status(mpm_engine* e, int scene, uint32_t* status, uint64_t* fp64_fallbacks);

/* FP64 sums on the device: [sum m, sum m x (3), sum m v (3)]. */
int mpm_scene_sums(mpm_engine* e, int scene, double out7[7]);

/* Average per-kernel milliseconds of the last timed mpm_step call:
 * [bin+sort, p2g, grid, g2p(+fp64 fix-up), end]; *nsteps = steps averaged. */
int mpm_kernel_times(mpm_engine* e, double ms[5], int* nsteps);

/* Particle slots of the buffers (no counterpart in the reference; a
 * diagnostic of the block-aligned layout, mpm_config.dense_blocks): *used =
 * slots in use up to the last particle, the holes between home blocks and a
 * removed scene's particles (until the next step compacts them) included;
 * *total = slots per buffer. */
int mpm_slots(mpm_engine* e, size_t* used, size_t* total);

/* Issue-rate probe (no counterpart in the reference; a measurement for the
 * P2G bound): on `device`, waves_per_simd waves per SIMD each run
 * `iterations` iterations of 4 independent v_mfma_f32_16x16x4f32 (mode 0),
 * of 4 * valu_per_mfma FP32 FMAs (mode 1, valu_per_mfma in 1, 2, 4, 8, 16),
 * or of both interleaved (mode 2).  out[mode] = shader-clock cycles per
 * iteration of the slowest wave, out[3 + mode] = kernel microseconds. */
int mpm_probe_issue(int device, int waves_per_simd, int valu_per_mfma, int iterations, double out[6]);

/* HBM streaming probe (no counterpart in the reference; a measurement for
 * the G2P bound): on `device`, mode 0 copies 1 GiB as float4, mode 1 reads
 * 19 and writes 22 FP32 arrays of 8,000,000 elements (G2P's particle streams
 * on the 8M slab) with a 128-thread workgroup per 512 consecutive elements,
 * mode 2 the same with one element per thread.  out[0] = microseconds per
 * launch (mean of 10), out[1] = bytes read + written per launch,
 * out[2] = TB/s. */
int mpm_probe_stream(int device, int mode, double out[3]);

/* The same steps replayed from a captured hipGraph (one graph per buffer
 * parity, re-captured only when the launch setup changes); the device-side
 * particle/block counts make one graph valid for any set of resident scenes.
 * Status is checked once at the end. */
int mpm_step_graph(mpm_engine* e, int nsteps);

/* mpm_step_graph for n engines of one device (1 <= n <= 64), stepped
 * together: each step launches every engine's graph on that engine's own
 * stream, so their kernels can share the GPU (a memory-bound stage of one
 * engine beside a compute-bound stage of another).  Engines with a
 * semi-implicit scene take their host-driven steps meanwhile.  Every engine
 * is synchronised and checked before returning; the first error is returned,
 * its message prefixed with "engine <index>: ".  Each scene's results are
 * those of mpm_step_graph on its own engine.  No counterpart in the
 * reference (its solver steps one particle set). */
int mpm_step_graph_multi(mpm_engine* const* engines, int n, int nsteps);

/* Chooses the tile sizes (P2G staging chunk 16/32/64, G2P workgroup size
 * 64/128/256), then the persistent grid sizes of the step kernels, then the
 * re-sort period among 16, 8, 4 and 2 (16 steps per candidate), by timing
 * real steps (results depend on none of them, so tuning steps are ordinary
 * steps).  steps_per_config >= 1; chosen (4 ints, may be NULL) receives
 * {particle grid, wave-per-block grid, workgroup-per-block grid, k_blocks grid}. */
int mpm_autotune(mpm_engine* e, int steps_per_config, int chosen[4]);

/* The engine's current tile sizes (mpm_config's, or the ones mpm_autotune
 * chose).  No counterpart in the reference. */
int mpm_tile_sizes(mpm_engine* e, int* p2g_chunk, int* g2p_threads);

/* The engine's current re-sort period (mpm_config.permute_period, or the one
 * mpm_autotune chose).  No counterpart in the reference. */
int mpm_permute_period(mpm_engine* e, int* period);

/* Conjugate-residual statistics of the scene's last semi-implicit solve:
 * iterations taken and |r|_M after 0..iterations of them (at most cap
 * values written; residual may be NULL).  0 iterations for beta = 0. */
int mpm_cr_stats(mpm_engine* e, int scene, int* iterations, double* residual, int cap);

/* Steps scene `scene` has taken since it was added (setup not counted).
 * No counterpart in the reference, whose step() takes no scene handle. */
int mpm_scene_steps(mpm_engine* e, int scene, uint64_t* steps);

/* ---- serving: continuous batching of scene requests ----
 * Each request brings its particles (host arrays, V0 computed by setup) and a
 * step count; it is admitted once it has arrived (arrival_s after the call
 * starts) and its device footprint fits (mpm_request_footprint: particle
 * slots, grid blocks and P2G partial sums, in bytes of the engine's
 * pools), steps with all resident scenes in graph-replayed chunks, and is
 * downloaded and retired as soon as its steps are done.  A request whose
 * scene raises a device status bit is retired at the end of that chunk with
 * the bits in `status` (its outputs are not written); the others go on.
 * mode 0: continuous batching; mode 1: the plain per-scene GPU step, one
 * request at a time in arrival order. */
#define MPM_MAX_COLLIDERS 8
typedef struct {
    double arrival_s;
    int steps;
    int ncolliders;   /* <= MPM_MAX_COLLIDERS */
    size_t n;
    const double* x;
    const double* v;
    const double* m;
    const double* E; /* may be NULL */
    mpm_params prm;
    mpm_collider colliders[MPM_MAX_COLLIDERS];
    double* out_x; /* n*3, may be NULL */
    double* out_v; /* n*3, may be NULL */
    double t_admit_s; /* filled in */
    double t_done_s;  /* filled in */
    uint32_t status;  /* filled in: 0, or the device status bits that retired the request */
    size_t footprint_bytes; /* filled in: device bytes reserved for the request */
} mpm_request;

typedef struct {
    double wall_s;
    double scenes_per_s;
    double p50_s, p99_s, mean_s, max_s; /* completion latency = t_done - arrival, completed requests */
    double particle_steps;
    int completed;
    int max_resident;
    int failed;              /* requests retired with a device status */
    double max_resident_bytes; /* peak sum of the resident requests' footprints */
    /* where the loop's wall time went (mode 0): stepping (graph replays and
     * the set-up of newly admitted scenes, to the end of each chunk),
     * admitting (uploads) and retiring (status reads, downloads, removals);
     * chunks and the steps they took together */
    double time_step_s, time_admit_s, time_retire_s;
    int chunks;
    int chunk_steps;
} mpm_serve_stats;

/* Device bytes request q would hold in engine e: n particle slots (both
 * particle buffers, sort and fallback arrays), its grid blocks (hash slot,
 * cell tables, neighbour and velocity arrays) and its home blocks' P2G
 * partial sums.  Blocks are estimated from the bounding box of q's positions
 * in 4^3-node blocks, with a 25% margin for the body spreading as it moves.
 * pools (may be NULL) receives the engine's three pool sizes in bytes:
 * particles, blocks (at the hash table's 50% load limit), partial sums. */
int mpm_request_footprint(mpm_engine* e, const mpm_request* q, size_t* bytes, size_t pools[3]);

int mpm_serve(mpm_engine* e, mpm_request* reqs, int nreq, int mode, int max_chunk, mpm_serve_stats* stats);

/* ---- collider factories (collision.cpp:5-27 and the box/sphere bodies) ---- */
mpm_collider mpm_make_plane(const double point[3], const double normal[3], double friction);
mpm_collider mpm_make_box(const double centre[3], const double half_extent[3], double friction);
mpm_collider mpm_make_sphere(const double centre[3], double radius, double friction);

/* ---- host I/O in the reference formats ----
 * groups.csv: group_id,vx,vy,vz,mass; particles.csv: group_id,x,y,z
 * (import.cpp:37-137).  On success *n is the particle count and x, v, m are
 * malloc'ed (release with mpm_free).  Errors return MPM_ERR_IO and the
 * reference's message is in mpm_last_error(). */
int mpm_import_particles(const char* groups_csv, const char* particles_csv, size_t* n, double** x, double** v,
                         double** m);
/* output/frame_%06d.csv writer (export.cpp:10-22): header x,y,z, rows x,-z,y. */
int mpm_export_frame(const char* directory, int frame_no, size_t n, const double* x);
void mpm_free(void* p);

/* ---- synthetic workloads (seeded, identical for oracle and GPU) ----
 * kind 0: snow cube (64,000 particles); 1: two tori; 2: slab; 3: one scene of
 * the batched distribution (index selects it).  scale < 1 shrinks the shape
 * (same spacing) for parity-sized runs; 1 = the full configuration.
 * Outputs are malloc'ed: x, v (n*3), m, E (n; E may be NULL if unused).
 * prm / colliders receive the configuration's parameters. */
int mpm_scene_generate(int kind, uint64_t seed, int index, double scale, size_t* n, double** x, double** v,
                       double** m, double** E, mpm_params* prm, mpm_collider* colliders, int* ncolliders);

/* ---- seeding: lattice particles inside a closed triangle mesh ----
 * Replaces seed_mesh / point_inside_mesh / ray_inside / world_bounds of the
 * reference's Blender add-on (blender/seed_export.py:43-131).  verts: nverts*3
 * coordinates in the mesh's frame (rounded to float32, as Blender stores
 * them); tris: ntris*3 vertex indices (polygons already triangulated);
 * spacing > 0 (rounded to float32, like the add-on's spacing property).
 * Lattice points min + spacing/2 + k spacing per axis inside the mesh's
 * bounds are kept when at least two of three fixed rays cross the surface an
 * odd number of times; the inside test runs on GPU `device`.  On success *n is
 * the count and *x (n*3, malloc'ed, release with mpm_free; NULL when n = 0)
 * the points as float32 values, x outer and z inner, in the mesh's frame. */
int mpm_seed_mesh(int device, const double* verts, size_t nverts, const int32_t* tris, size_t ntris, double spacing,
                  size_t* n, double** x);

#ifdef __cplusplus
}
#endif
#endif
