/*
 * snow_oracle.h -- FP64 CPU oracle for the Stomakhin-style MPM step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or the timed CPU baseline.  The product path (libsnowmpm.so) never
 * links, loads or calls it.
 *
 * This is a from-scratch restatement of the reference solver
 * (src/grid.cpp:15-450, src/collision.cpp:5-27, include/snow/ headers) in plain
 * C++17 without Eigen.  Every function cites the reference lines it follows.
 * The reference itself cannot be built here (it needs Eigen3, which is not
 * installed), so this restatement is pinned by the reference's own unit-test
 * known answers (tests/snow_tests.cpp), by closed forms and invariants, and
 * by SVD/polar golden vectors produced with numpy/scipy.
 *
 * Layout: all structs are POD so that Python (ctypes) can fill them.
 */
#ifndef SNOW_ORACLE_H
#define SNOW_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Particle, include/snow/particle.hpp:5-28 (+ optional per-particle material).
 * F_P and F_E are row-major 3x3.  mu0/lambda0 < 0 means "use the step's
 * scalar mu_0/lambda_0" exactly as the reference does. */
typedef struct {
    double x[3];
    double v[3];
    double m;
    double V0;
    double Fp[9];
    double Fe[9];
    double mu0;
    double lambda0;
} orc_particle;

/* Collider descriptor.  type 0 = plane (collision.cpp:5-27), 1 = solid
 * axis-aligned box, 2 = solid sphere.  velocity is the collider's constant
 * velocity (the reference plane returns zero, collision.cpp:19-22).
 * flags bit 0: after position_update, move particles inside the body onto
 * its surface (an extension; 0 is the reference behaviour). */
typedef struct {
    int type;
    int flags;
    double point[3];       /* plane: point on plane; box/sphere: centre */
    double normal[3];      /* plane: normal (normalised by orc_make_plane) */
    double half_extent[3]; /* box */
    double radius;         /* sphere */
    double velocity[3];
    double friction;
} orc_collider;

/* Step parameters, the scalar arguments of step() (grid.hpp:97). */
typedef struct {
    double gravity[3];
    double dt;
    double dx;
    double mu0;
    double lambda0;
    double hardening;       /* xi */
    double max_compression; /* theta_c */
    double max_stretch;     /* theta_s */
    double alpha;           /* FLIP ratio */
    /* Semi-implicit velocity solve (declared, never defined, by the reference:
     * grid.hpp:77, 81, 89-92).  beta = 0: the explicit step. */
    double beta;            /* implicit ratio in [0, 1] */
    double cr_tolerance;    /* stop when |r|_M <= cr_tolerance |b|_M */
    int cr_max_iterations;
} orc_params;

/* Conjugate-residual statistics of the last step's solve. */
typedef struct {
    int iterations;
    int nodes;              /* unknown nodes (mass > 0, not held by a collider) */
    double residual[257];   /* |r|_M after 0..iterations iterations */
} orc_cr_stats;

/* StepTimings, include/snow/debug.hpp:3-15 (milliseconds, accumulated). */
typedef struct {
    double stencil, p2g, forces_constitutive, forces_accumulative,
        grid_velocity, grid_collisions, stress_update, g2p,
        particle_collisions, position_update;
} orc_timings;

/* ---- scalar kernels (grid.cpp:24-58) ---- */
double orc_N(double a, double b);
double orc_N_dash(double grid_node, double particle_coordinate);

/* ---- 3x3 linear algebra used in place of Eigen ---- */
double orc_det3(const double A[9]);
int orc_inv3(const double A[9], double Ainv[9]);
/* Unsigned SVD A = U diag(S) V^T, S >= 0 sorted descending (Eigen JacobiSVD
 * convention).  Returns the number of Jacobi sweeps used. */
int orc_svd3(const double A[9], double U[9], double S[3], double V[9]);

/* ---- collider factory and helper (collision.cpp:5-27, grid.cpp:206-231) */
orc_collider orc_make_plane(const double point[3], const double normal[3], double friction);
orc_collider orc_make_box(const double centre[3], const double half_extent[3], double friction);
orc_collider orc_make_sphere(const double centre[3], double radius, double friction);
double orc_collider_phi(const orc_collider* c, const double x[3]);
void orc_collision_helper(const double x[3], const double v[3], const orc_collider* c, double out[3]);

/* ---- sparse grid handle (std::unordered_map, grid.hpp:72) ---- */
typedef struct orc_grid orc_grid;
orc_grid* orc_grid_new(void);
void orc_grid_free(orc_grid* g);
size_t orc_grid_size(const orc_grid* g);
/* Insert-or-overwrite one node (like operator[] then assignment). */
void orc_grid_set(orc_grid* g, int i, int j, int k, double mass,
                  const double momentum[3], const double velocity[3], const double force[3]);
/* Returns 1 and fills outputs if present, else 0 (like find()). */
int orc_grid_get(const orc_grid* g, int i, int j, int k, double* mass,
                 double momentum[3], double velocity[3], double force[3]);
/* Dump all nodes: idx[n*3], mass[n], momentum[n*3], velocity[n*3], force[n*3]. */
size_t orc_grid_dump(const orc_grid* g, int* idx, double* mass, double* momentum,
                     double* velocity, double* force);

/* ---- per-stage functions (grid.hpp:75-95) ---- */
void orc_build_stencil(const orc_particle* p, double dx, int idx[64 * 3], double w[64], double grad[64 * 3]);
void orc_p2g(orc_grid* g, const orc_particle* p, double dx);
void orc_calculate_volume(orc_grid* g, orc_particle* p, double dx);
void orc_force_constitutive(const orc_particle* p, double mu0, double lambda0, double xi, double P[9]);
void orc_force_calculation(orc_grid* g, const orc_particle* p, double dx, double mu0, double lambda0, double xi);
void orc_grid_velocity(orc_grid* g, const double gravity[3], double dt);
void orc_grid_collisions(orc_grid* g, const orc_collider* c, int nc, double dx);
void orc_stress_update(const orc_grid* g, orc_particle* p, double dx, double dt, double theta_c, double theta_s);
void orc_g2p(const orc_grid* g, orc_particle* p, double dx, double alpha);
void orc_particle_collisions(orc_particle* p, const orc_collider* c, int nc);
void orc_position_update(orc_particle* p, double dt);
/* Optional projection after position_update (colliders with flags bit 0). */
void orc_position_projection(orc_particle* p, const orc_collider* c, int nc);

/* ---- semi-implicit solve (grid.hpp:77, 81, 89-92; restated design) ----
 * orc_constitutive_differential: dP = d(P)/dF : dF of the corotated model
 * with hardening at particle p.  orc_apply_A: y = A q on the unknown nodes of
 * a grid built by orc_step(..., grid_out) (node order = orc_krylov_nodes). */
void orc_constitutive_differential(const orc_particle* p, const double dF[9], double mu0, double lambda0, double xi,
                                   double dP[9]);
size_t orc_krylov_nodes(const orc_grid* g, int* idx);
void orc_apply_A(orc_grid* g, const orc_particle* ps, size_t n, const orc_params* prm, const double* q, double* y);
void orc_last_cr_stats(orc_cr_stats* out);

/* ---- whole-path entry points (grid.cpp:340-450) ---- */
void orc_setup(orc_particle* ps, size_t n, double dx);
/* One reference step.  nthreads <= 0 keeps the OpenMP default.  If grid_out
 * is non-NULL it receives the step's grid (after grid collisions). */
void orc_step(orc_particle* ps, size_t n, const orc_collider* c, int nc,
              const orc_params* prm, orc_timings* t, int nthreads, orc_grid* grid_out);
/* Run nsteps steps (convenience for parity runs). */
void orc_run(orc_particle* ps, size_t n, const orc_collider* c, int nc,
             const orc_params* prm, int nsteps, orc_timings* t, int nthreads);

/* ---- mesh seeding (blender/seed_export.py:43-131), seed_oracle.cpp ----
 * verts: nverts x 3 (stored as float32, like Blender's mesh), tris: ntris x 3
 * vertex indices.  orc_seed_axis: lattice coordinates of one axis (returns
 * the count; writes at most cap).  orc_seed_direction: ray direction k.
 * orc_seed_ray_hits: hits counted along ray k from point p (-1: bad input).
 * orc_seed_mesh: the seeded points (x outer, z inner), count returned, at
 * most cap written to out (n x 3); -1 on bad input. */
size_t orc_seed_axis(float lo, float hi, double spacing, float* out, size_t cap);
void orc_seed_direction(int k, float d[3]);
int orc_seed_ray_hits(const double* verts, size_t nverts, const int* tris, size_t ntris, const double p[3], int k);
long orc_seed_mesh(const double* verts, size_t nverts, const int* tris, size_t ntris, double spacing, double* out,
                   size_t cap);

#ifdef __cplusplus
}
#endif
#endif
