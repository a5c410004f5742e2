// snow_oracle.cpp -- FP64 CPU oracle for the Stomakhin-style MPM step.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker / CPU baseline.  Never linked
// into the product library.
//
// A restatement, written from scratch, of the reference solver's algorithm:
//   src/grid.cpp:15-450      (all stages, setup, step)
//   src/collision.cpp:5-27   (make_plane)
//   include/snow/grid.hpp, particle.hpp, collision.hpp, debug.hpp (types)
// with Eigen replaced by small hand-written 3x3 routines (determinant,
// adjugate inverse and a one-sided Jacobi SVD with Eigen's unsigned,
// descending convention).  The parallel split follows the reference exactly:
// OpenMP on the stencil build, stress update and G2P loops (grid.cpp:364,
// 414, 424); P2G, forces, grid operations, particle collisions and advection
// are serial.  The sparse grid is a std::unordered_map rebuilt every step
// (grid.cpp:361), so timing this file is timing the reference's algorithm.
//
// Extensions beyond the reference (documented, used only where the workload
// needs them): box and sphere colliders, constant collider velocity, and an
// optional per-particle (mu0, lambda0).

#include "snow_oracle.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V3 {
    double a[3];
    double& operator[](int i) { return a[i]; }
    double operator[](int i) const { return a[i]; }
};

inline V3 v3(double x, double y, double z) { return V3{{x, y, z}}; }
inline V3 add(const V3& p, const V3& q) { return v3(p[0] + q[0], p[1] + q[1], p[2] + q[2]); }
inline V3 sub(const V3& p, const V3& q) { return v3(p[0] - q[0], p[1] - q[1], p[2] - q[2]); }
inline V3 scl(double s, const V3& p) { return v3(s * p[0], s * p[1], s * p[2]); }
inline double dot(const V3& p, const V3& q) { return p[0] * q[0] + p[1] * q[1] + p[2] * q[2]; }
inline double norm(const V3& p) { return std::sqrt(dot(p, p)); }

// Row-major 3x3.
struct M3 {
    double a[9];
};
inline M3 m3_identity() {
    M3 r{};
    r.a[0] = r.a[4] = r.a[8] = 1.0;
    return r;
}
inline M3 m3_mul(const M3& A, const M3& B) {
    M3 r{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += A.a[i * 3 + k] * B.a[k * 3 + j];
            r.a[i * 3 + j] = s;
        }
    return r;
}
inline M3 m3_transpose(const M3& A) {
    M3 r{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.a[i * 3 + j] = A.a[j * 3 + i];
    return r;
}
inline V3 m3_apply(const M3& A, const V3& x) {
    return v3(A.a[0] * x[0] + A.a[1] * x[1] + A.a[2] * x[2],
              A.a[3] * x[0] + A.a[4] * x[1] + A.a[5] * x[2],
              A.a[6] * x[0] + A.a[7] * x[1] + A.a[8] * x[2]);
}

double det3(const double* A) {
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}

// Adjugate / determinant inverse (Eigen uses the same cofactor form for 3x3).
M3 inv3(const M3& M) {
    const double* A = M.a;
    M3 r{};
    double c00 = A[4] * A[8] - A[5] * A[7];
    double c01 = A[5] * A[6] - A[3] * A[8];
    double c02 = A[3] * A[7] - A[4] * A[6];
    double d = A[0] * c00 + A[1] * c01 + A[2] * c02;
    double id = 1.0 / d;
    r.a[0] = c00 * id;
    r.a[1] = (A[2] * A[7] - A[1] * A[8]) * id;
    r.a[2] = (A[1] * A[5] - A[2] * A[4]) * id;
    r.a[3] = c01 * id;
    r.a[4] = (A[0] * A[8] - A[2] * A[6]) * id;
    r.a[5] = (A[2] * A[3] - A[0] * A[5]) * id;
    r.a[6] = c02 * id;
    r.a[7] = (A[1] * A[6] - A[0] * A[7]) * id;
    r.a[8] = (A[0] * A[4] - A[1] * A[3]) * id;
    return r;
}

V3 cross(const V3& p, const V3& q) {
    return v3(p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]);
}

// One-sided (Hestenes) Jacobi SVD of a 3x3 matrix.  Produces the unsigned
// decomposition A = U diag(S) V^T with S >= 0 sorted descending, U and V
// orthogonal (either may be a reflection), which is the convention of
// Eigen::JacobiSVD used at grid.cpp:168 and 279.  Any correct unsigned SVD
// yields the same U*V^T (polar factor) and the same U*clamp(S)*V^T, because
// both are matrix functions of A; only degenerate (singular) inputs leave a
// choice, resolved here by completing U with a cross product.
int svd3(const double* Ain, double* Uo, double* So, double* Vo) {
    double W[3][3];  // columns of the working matrix: W[col][row]
    double Vc[3][3]; // columns of V
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            W[c][r] = Ain[r * 3 + c];
            Vc[c][r] = (r == c) ? 1.0 : 0.0;
        }
    const double eps = 1e-15;
    int sweep = 0;
    for (; sweep < 60; ++sweep) {
        bool rotated = false;
        static const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
        for (int pq = 0; pq < 3; ++pq) {
            int p = P[pq], q = Q[pq];
            double al = 0, be = 0, ga = 0;
            for (int r = 0; r < 3; ++r) {
                al += W[p][r] * W[p][r];
                be += W[q][r] * W[q][r];
                ga += W[p][r] * W[q][r];
            }
            if (ga == 0.0 || std::fabs(ga) <= eps * std::sqrt(al * be)) continue;
            rotated = true;
            double zeta = (be - al) / (2.0 * ga);
            double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
            double c = 1.0 / std::sqrt(1.0 + t * t);
            double s = c * t;
            for (int r = 0; r < 3; ++r) {
                double wp = W[p][r], wq = W[q][r];
                W[p][r] = c * wp - s * wq;
                W[q][r] = s * wp + c * wq;
                double vp = Vc[p][r], vq = Vc[q][r];
                Vc[p][r] = c * vp - s * vq;
                Vc[q][r] = s * vp + c * vq;
            }
        }
        if (!rotated) break;
    }
    double sig[3];
    for (int c = 0; c < 3; ++c)
        sig[c] = std::sqrt(W[c][0] * W[c][0] + W[c][1] * W[c][1] + W[c][2] * W[c][2]);
    int ord[3] = {0, 1, 2};
    std::sort(ord, ord + 3, [&](int a, int b) { return sig[a] > sig[b]; });
    V3 u[3];
    for (int k = 0; k < 3; ++k) {
        int c = ord[k];
        So[k] = sig[c];
        for (int r = 0; r < 3; ++r) Vo[r * 3 + k] = Vc[c][r];
    }
    const double smax = So[0];
    for (int k = 0; k < 3; ++k) {
        int c = ord[k];
        if (So[k] > 1e-12 * smax && So[k] > 0.0) {
            u[k] = scl(1.0 / So[k], v3(W[c][0], W[c][1], W[c][2]));
        } else if (k == 0) {
            u[0] = v3(1, 0, 0);
        } else if (k == 1) {
            // any unit vector orthogonal to u0
            V3 e = std::fabs(u[0][0]) < 0.9 ? v3(1, 0, 0) : v3(0, 1, 0);
            V3 t = sub(e, scl(dot(e, u[0]), u[0]));
            u[1] = scl(1.0 / norm(t), t);
        } else {
            u[2] = cross(u[0], u[1]);
        }
    }
    for (int k = 0; k < 3; ++k)
        for (int r = 0; r < 3; ++r) Uo[r * 3 + k] = u[k][r];
    return sweep;
}

// signum with sign(0) = 0, grid.cpp:15-20.
inline double signum(double x) {
    if (x > 0.0) return 1.0;
    if (x < 0.0) return -1.0;
    return 0.0;
}

// Cubic B-spline, grid.cpp:24-40.
double Nfun(double a, double b) {
    double d = std::fabs(a - b);
    if (d >= 0.0 && d < 1.0) return 0.5 * d * d * d - d * d + 2.0 / 3.0;
    if (d >= 1.0 && d < 2.0) return -1.0 / 6.0 * d * d * d + d * d - 2.0 * d + 4.0 / 3.0;
    return 0.0;
}

// Derivative w.r.t. the particle coordinate, grid.cpp:42-58.
double Ndash(double g, double p) {
    double d = std::fabs(g - p);
    if (d >= 0.0 && d < 1.0) return (1.5 * d * d - 2.0 * d) * signum(p - g);
    if (d >= 1.0 && d < 2.0) return (-0.5 * d * d + 2.0 * d - 2.0) * signum(p - g);
    return 0.0;
}

// GridIndex (grid.hpp:23-30) and a hash for it.  The hash only decides
// iteration order of the map, which never changes numerics: accumulation
// order is particle order, and per-node stages are independent.
struct Key {
    int i, j, k;
    bool operator==(const Key& o) const { return i == o.i && j == o.j && k == o.k; }
};
struct KeyHash {
    size_t operator()(const Key& q) const noexcept {
        uint64_t h = (uint64_t)(uint32_t)q.i * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)(uint32_t)q.j * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
        h ^= (uint64_t)(uint32_t)q.k * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};

// GridNode, grid.hpp:11-21 (density omitted: it never affects numerics).
// delta_v, delta_force and the collision flag serve the semi-implicit solve
// the reference declares (grid.hpp:18-20).
struct Node {
    double mass = 0.0;
    V3 momentum{{0, 0, 0}};
    V3 velocity{{0, 0, 0}};
    V3 force{{0, 0, 0}};
    V3 delta_v{{0, 0, 0}};
    V3 delta_force{{0, 0, 0}};
    bool constrained = false;  // a collider changed this node's velocity this step
};

using Grid = std::unordered_map<Key, Node, KeyHash>;

// StencilEntry / ParticleStencil, grid.hpp:56-70.
struct Entry {
    Key index;
    double weight;
    V3 gradient;
    Node* node;
};
using Stencil = std::array<Entry, 64>;

struct Particle {
    V3 x, v;
    double m, V0;
    M3 Fp, Fe;
    double mu0, lambda0;
};

Particle from_c(const orc_particle& p) {
    Particle q;
    q.x = v3(p.x[0], p.x[1], p.x[2]);
    q.v = v3(p.v[0], p.v[1], p.v[2]);
    q.m = p.m;
    q.V0 = p.V0;
    std::memcpy(q.Fp.a, p.Fp, sizeof q.Fp.a);
    std::memcpy(q.Fe.a, p.Fe, sizeof q.Fe.a);
    q.mu0 = p.mu0;
    q.lambda0 = p.lambda0;
    return q;
}
void to_c(const Particle& q, orc_particle& p) {
    for (int i = 0; i < 3; ++i) {
        p.x[i] = q.x[i];
        p.v[i] = q.v[i];
    }
    p.m = q.m;
    p.V0 = q.V0;
    std::memcpy(p.Fp, q.Fp.a, sizeof p.Fp);
    std::memcpy(p.Fe, q.Fe.a, sizeof p.Fe);
    p.mu0 = q.mu0;
    p.lambda0 = q.lambda0;
}

inline void base_of(const Particle& p, double dx, V3& gp, int base[3]) {
    gp = v3(p.x[0] / dx, p.x[1] / dx, p.x[2] / dx);
    for (int a = 0; a < 3; ++a) base[a] = (int)std::floor(gp[a]) - 1;
}

// build_stencil, grid.cpp:60-117.
void build_stencil(const Particle& p, double dx, Stencil& st) {
    V3 gp;
    int base[3];
    base_of(p, dx, gp, base);
    double w[3][4], d[3][4];
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < 4; ++i) {
            w[a][i] = Nfun(base[a] + i, gp[a]);
            d[a][i] = Ndash(base[a] + i, gp[a]) / dx;
        }
    int n = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k) {
                Entry& e = st[n++];
                e.index = Key{base[0] + i, base[1] + j, base[2] + k};
                e.weight = w[0][i] * w[1][j] * w[2][k];
                e.gradient = v3(d[0][i] * w[1][j] * w[2][k], w[0][i] * d[1][j] * w[2][k],
                                w[0][i] * w[1][j] * d[2][k]);
                e.node = nullptr;
            }
}

// P2G, grid.cpp:119-129.
void p2g(Grid& g, const Particle& p, Stencil& st) {
    const V3 mom = scl(p.m, p.v);
    for (Entry& e : st) {
        Node& nd = g[e.index];
        e.node = &nd;
        nd.mass += e.weight * p.m;
        nd.momentum = add(nd.momentum, scl(e.weight, mom));
    }
}

// calculate_volume, grid.cpp:138-160.
void calculate_volume(Grid& g, Particle& p, double dx) {
    V3 gp;
    int base[3];
    base_of(p, dx, gp, base);
    double density = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k) {
                Key idx{i + base[0], j + base[1], k + base[2]};
                const Node& nd = g[idx];
                double wp = Nfun(idx.i, gp[0]) * Nfun(idx.j, gp[1]) * Nfun(idx.k, gp[2]);
                density += (nd.mass * wp) / std::pow(dx, 3.0);
            }
    p.V0 = p.m / density;
}

// force_constitutive, grid.cpp:162-172.
M3 force_constitutive(const Particle& p, double mu0, double lambda0, double xi) {
    double h = std::exp(xi * (1.0 - det3(p.Fp.a)));
    double mu = mu0 * h, lambda = lambda0 * h;
    double J = det3(p.Fe.a);
    M3 U, V;
    double S[3];
    svd3(p.Fe.a, U.a, S, V.a);
    M3 R = m3_mul(U, m3_transpose(V));
    M3 FinvT = m3_transpose(inv3(p.Fe));
    M3 P;
    for (int i = 0; i < 9; ++i)
        P.a[i] = 2.0 * mu * (p.Fe.a[i] - R.a[i]) + lambda * (J - 1.0) * J * FinvT.a[i];
    return P;
}

// force_grid_accumulation, grid.cpp:174-182.
void force_grid_accumulation(const M3& P, const Particle& p, const Stencil& st) {
    M3 FT = m3_transpose(p.Fe);
    M3 Mf = m3_mul(P, FT);
    for (double& x : Mf.a) x *= -p.V0;
    for (const Entry& e : st) e.node->force = add(e.node->force, m3_apply(Mf, e.gradient));
}

// grid_velocity, grid.cpp:195-204.
void grid_velocity(Grid& g, const V3& gravity, double dt) {
    for (auto& kv : g) {
        Node& nd = kv.second;
        if (nd.mass > 0)
            nd.velocity = add(scl(1.0 / nd.mass, nd.momentum),
                              scl(dt, add(scl(1.0 / nd.mass, nd.force), gravity)));
    }
}

// Collider geometry.  Plane: collision.cpp:5-27 (phi = n.(x - p), constant
// normal).  Box and sphere are solid SDF bodies added for the workloads.
double coll_phi(const orc_collider& c, const V3& x, V3* nrm) {
    if (c.type == 0) {
        V3 n = v3(c.normal[0], c.normal[1], c.normal[2]);
        if (nrm) *nrm = n;
        return dot(n, sub(x, v3(c.point[0], c.point[1], c.point[2])));
    }
    if (c.type == 1) {
        V3 d = sub(x, v3(c.point[0], c.point[1], c.point[2]));
        double q[3], qmax = -1e300;
        int amax = 0;
        double outside = 0.0;
        for (int a = 0; a < 3; ++a) {
            q[a] = std::fabs(d[a]) - c.half_extent[a];
            if (q[a] > qmax) {
                qmax = q[a];
                amax = a;
            }
            if (q[a] > 0) outside += q[a] * q[a];
        }
        double phi = std::sqrt(outside) + std::min(qmax, 0.0);
        if (nrm) {
            V3 n = v3(0, 0, 0);
            if (qmax > 0) {
                for (int a = 0; a < 3; ++a)
                    n[a] = q[a] > 0 ? q[a] * (d[a] >= 0 ? 1.0 : -1.0) : 0.0;
                double l = norm(n);
                n = scl(1.0 / l, n);
            } else {
                n[amax] = d[amax] >= 0 ? 1.0 : -1.0;
            }
            *nrm = n;
        }
        return phi;
    }
    // sphere
    V3 d = sub(x, v3(c.point[0], c.point[1], c.point[2]));
    double l = norm(d);
    if (nrm) *nrm = l > 0 ? scl(1.0 / l, d) : v3(0, 1, 0);
    return l - c.radius;
}

// collision_helper, grid.cpp:206-231.
V3 collision_helper(const V3& x, const V3& v, const orc_collider& c) {
    V3 n;
    double phi = coll_phi(c, x, &n);
    if (phi <= 0) {
        V3 vc = v3(c.velocity[0], c.velocity[1], c.velocity[2]);
        V3 vrel = sub(v, vc);
        double vn = dot(vrel, n);
        if (vn < 0) {
            V3 vt = sub(vrel, scl(vn, n));
            V3 vd;
            if (norm(vt) <= -1 * c.friction * vn)
                vd = v3(0, 0, 0);
            else
                vd = add(vt, scl(c.friction * vn, scl(1.0 / norm(vt), vt)));
            return add(vd, vc);
        }
    }
    return v;
}

// grid_collisions, grid.cpp:233-243.  Also records which nodes a collider
// changed (the semi-implicit solve holds those fixed).
void grid_collisions(Grid& g, const orc_collider* cs, int nc, double dx) {
    for (auto& kv : g) {
        V3 xp = v3(dx * kv.first.i, dx * kv.first.j, dx * kv.first.k);
        const V3 before = kv.second.velocity;
        for (int b = 0; b < nc; ++b) kv.second.velocity = collision_helper(xp, kv.second.velocity, cs[b]);
        const V3& after = kv.second.velocity;
        kv.second.constrained = after[0] != before[0] || after[1] != before[1] || after[2] != before[2];
    }
}

// stress_update, grid.cpp:245-286.
void stress_update(const Grid& g, Particle& p, double dx, double dt, double tc, double ts) {
    V3 gp;
    int base[3];
    base_of(p, dx, gp, base);
    M3 L{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k) {
                Key idx{i + base[0], j + base[1], k + base[2]};
                auto it = g.find(idx);
                if (it == g.end()) continue;
                const Node& nd = it->second;
                double Nx = Nfun(idx.i, gp[0]), Ny = Nfun(idx.j, gp[1]), Nz = Nfun(idx.k, gp[2]);
                double Dx = Ndash(idx.i, gp[0]), Dy = Ndash(idx.j, gp[1]), Dz = Ndash(idx.k, gp[2]);
                V3 bg = v3(Dx * Ny * Nz * (1 / dx), Nx * Dy * Nz * (1 / dx), Nx * Ny * Dz * (1 / dx));
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) L.a[r * 3 + c] += nd.velocity[r] * bg[c];
            }
    M3 A = m3_identity();
    for (int i = 0; i < 9; ++i) A.a[i] += dt * L.a[i];
    M3 Fe_trial = m3_mul(A, p.Fe);
    M3 F_trial = m3_mul(Fe_trial, p.Fp);
    M3 U, V;
    double S[3];
    svd3(Fe_trial.a, U.a, S, V.a);
    for (int a = 0; a < 3; ++a) S[a] = std::clamp(S[a], 1 - tc, 1 + ts);
    M3 US = U;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) US.a[r * 3 + c] *= S[c];
    p.Fe = m3_mul(US, m3_transpose(V));
    p.Fp = m3_mul(inv3(p.Fe), F_trial);
}

// G2P, grid.cpp:288-325.
void g2p(const Grid& g, Particle& p, double dx, double alpha) {
    V3 vpic = v3(0, 0, 0), vtmp = v3(0, 0, 0);
    V3 gp;
    int base[3];
    base_of(p, dx, gp, base);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 4; ++k) {
                Key idx{i + base[0], j + base[1], k + base[2]};
                auto it = g.find(idx);
                if (it == g.end()) continue;
                const Node& nd = it->second;
                double wp = Nfun(idx.i, gp[0]) * Nfun(idx.j, gp[1]) * Nfun(idx.k, gp[2]);
                vpic = add(vpic, scl(wp, nd.velocity));
                if (nd.mass > 0) vtmp = add(vtmp, scl(wp, sub(nd.velocity, scl(1.0 / nd.mass, nd.momentum))));
            }
    V3 vflip = add(p.v, vtmp);
    p.v = add(scl(1 - alpha, vpic), scl(alpha, vflip));
}

// particle_collisions, grid.cpp:327-333.
void particle_collisions(Particle& p, const orc_collider* cs, int nc) {
    for (int b = 0; b < nc; ++b) p.v = collision_helper(p.x, p.v, cs[b]);
}

// position_update, grid.cpp:335-338.
void position_update(Particle& p, double dt) { p.x = add(p.x, scl(dt, p.v)); }

// Optional projection (not in the reference, which only advects): for the
// colliders with flags bit 0, in list order, a particle with phi(x) < 0 moves
// to x - phi n.  Velocities are left alone.
void position_projection(Particle& p, const orc_collider* cs, int nc) {
    for (int b = 0; b < nc; ++b) {
        if (!(cs[b].flags & 1)) continue;
        V3 n;
        const double phi = coll_phi(cs[b], p.x, &n);
        if (phi < 0) p.x = sub(p.x, scl(phi, n));
    }
}

struct GridBox {
    Grid g;
};

// ---------------------------------------------------------------------------
// Semi-implicit velocity update.  The reference declares it and never defines
// it (grid.hpp:77 constitutive_differential, 81 index_nodes, 89
// deformation_differential, 90 force_differential, 91 apply_A, 92
// Conjugate_residual; GridNode::delta_v / delta_force, grid.hpp:18-20).  This
// restates the method of Stomakhin et al. 2013 (section 4): after
// grid_velocity and grid_collisions, solve
//   (I + beta dt^2 M^-1 H) v = v*
// on the nodes with mass that no collider changed, where v* is the explicit
// grid velocity and H q = -df(q)/dt the Hessian of the elastic energy applied
// to a grid velocity field q:
//   dF_p = dt (sum_i q_i grad w_ip^T) F_Ep              deformation_differential
//   dP_p = d P / dF (F_Ep) : dF_p                       constitutive_differential
//   df_i = - sum_p V_p0 dP_p F_Ep^T grad w_ip            force_differential
//   A q = q - beta dt M^-1 df(q)                         apply_A
// M A = M + beta dt^2 H is symmetric, so A is self-adjoint in the mass inner
// product <a, b> = sum_i m_i a_i . b_i, and the conjugate-residual method in
// that product (Conjugate_residual) minimises |b - A x|_M over the Krylov
// space: the residual never grows.  The nodes are taken in index order
// (index_nodes), so every sum has a fixed order.  beta = 0 gives A = I and
// leaves v* unchanged.

// Corotated P(F) = 2 mu (F - R) + lambda (J - 1) J F^-T differentiated at F
// (force_constitutive, grid.cpp:162-172, with the same hardening):
//   dR = U W V^T, W_ij = (M_ij - M_ji) / (s_i + s_j), M = U^T dF V
//     (R^T dR is skew and R^T dR S + S R^T dR = R^T dF - dF^T R, S = R^T F)
//   d(J F^-T) = d cof(F) (cofactor matrix, bilinear), dJ = cof(F) : dF
M3 cofactor(const M3& F) {
    M3 C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            C.a[i * 3 + j] = F.a[i1 * 3 + j1] * F.a[i2 * 3 + j2] - F.a[i1 * 3 + j2] * F.a[i2 * 3 + j1];
        }
    return C;
}
M3 cofactor_differential(const M3& F, const M3& dF) {
    M3 C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            C.a[i * 3 + j] = dF.a[i1 * 3 + j1] * F.a[i2 * 3 + j2] + F.a[i1 * 3 + j1] * dF.a[i2 * 3 + j2] -
                             dF.a[i1 * 3 + j2] * F.a[i2 * 3 + j1] - F.a[i1 * 3 + j2] * dF.a[i2 * 3 + j1];
        }
    return C;
}

M3 constitutive_differential(const Particle& p, const M3& dF, double mu0, double lambda0, double xi) {
    const double h = std::exp(xi * (1.0 - det3(p.Fp.a)));
    const double mu = mu0 * h, lambda = lambda0 * h;
    const M3& F = p.Fe;
    const double J = det3(F.a);
    M3 U, V;
    double S[3];
    svd3(F.a, U.a, S, V.a);
    const M3 Mh = m3_mul(m3_transpose(U), m3_mul(dF, V));
    M3 W{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (i != j) W.a[i * 3 + j] = (Mh.a[i * 3 + j] - Mh.a[j * 3 + i]) / (S[i] + S[j]);
    const M3 dR = m3_mul(U, m3_mul(W, m3_transpose(V)));
    const M3 C = cofactor(F), dC = cofactor_differential(F, dF);
    double dJ = 0.0;
    for (int i = 0; i < 9; ++i) dJ += C.a[i] * dF.a[i];
    M3 dP;
    for (int i = 0; i < 9; ++i)
        dP.a[i] = 2.0 * mu * (dF.a[i] - dR.a[i]) + lambda * (dJ * C.a[i] + (J - 1.0) * dC.a[i]);
    return dP;
}

// deformation_differential: dF_p = dt (sum_i delta_v_i grad w_ip^T) F_Ep.
M3 deformation_differential(const Stencil& st, const Particle& p, double dt) {
    M3 L{};
    for (const Entry& e : st) {
        const V3& q = e.node->delta_v;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) L.a[r * 3 + c] += q[r] * e.gradient[c];
    }
    for (double& x : L.a) x *= dt;
    return m3_mul(L, p.Fe);
}

// force_differential: delta_force_i -= V_p0 dP_p F_Ep^T grad w_ip.
void force_differential(const Stencil& st, const Particle& p, double dt, double mu0, double lambda0, double xi) {
    const M3 dP = constitutive_differential(p, deformation_differential(st, p, dt), mu0, lambda0, xi);
    M3 G = m3_mul(dP, m3_transpose(p.Fe));
    for (double& x : G.a) x *= -p.V0;
    for (const Entry& e : st) e.node->delta_force = add(e.node->delta_force, m3_apply(G, e.gradient));
}

// index_nodes: the unknowns, nodes with mass that no collider changed, in
// index order.
std::vector<Node*> index_nodes(Grid& g) {
    std::vector<std::pair<Key, Node*>> v;
    for (auto& kv : g)
        if (kv.second.mass > 0.0 && !kv.second.constrained) v.push_back({kv.first, &kv.second});
    std::sort(v.begin(), v.end(), [](const auto& a, const auto& b) {
        if (a.first.i != b.first.i) return a.first.i < b.first.i;
        if (a.first.j != b.first.j) return a.first.j < b.first.j;
        return a.first.k < b.first.k;
    });
    std::vector<Node*> out;
    for (auto& e : v) out.push_back(e.second);
    return out;
}

using KrylovVector = std::vector<V3>;

double particle_mu0(const Particle& p, const orc_params& prm) { return p.mu0 >= 0 ? p.mu0 : prm.mu0; }
double particle_la0(const Particle& p, const orc_params& prm) { return p.lambda0 >= 0 ? p.lambda0 : prm.lambda0; }

// apply_A: y = q - beta dt M^-1 df(q) on the unknown nodes (q = 0 elsewhere).
KrylovVector apply_A(Grid& g, const std::vector<Particle>& snow, const std::vector<Stencil>& st,
                     const std::vector<Node*>& kn, const KrylovVector& q, const orc_params& prm) {
    for (auto& kv : g) {
        kv.second.delta_v = v3(0, 0, 0);
        kv.second.delta_force = v3(0, 0, 0);
    }
    for (size_t k = 0; k < kn.size(); ++k) kn[k]->delta_v = q[k];
    for (size_t i = 0; i < snow.size(); ++i)
        force_differential(st[i], snow[i], prm.dt, particle_mu0(snow[i], prm), particle_la0(snow[i], prm),
                           prm.hardening);
    KrylovVector y(kn.size());
    for (size_t k = 0; k < kn.size(); ++k)
        y[k] = sub(q[k], scl(prm.beta * prm.dt / kn[k]->mass, kn[k]->delta_force));
    return y;
}

double mdot(const std::vector<Node*>& kn, const KrylovVector& a, const KrylovVector& b) {
    double s = 0.0;
    for (size_t k = 0; k < kn.size(); ++k) s += kn[k]->mass * dot(a[k], b[k]);
    return s;
}

orc_cr_stats g_cr_stats{};

// Conjugate_residual: solves A v = v* for the unknown nodes' velocities.
void conjugate_residual(Grid& g, const std::vector<Particle>& snow, const std::vector<Stencil>& st,
                        const orc_params& prm) {
    std::vector<Node*> kn = index_nodes(g);
    orc_cr_stats stats{};
    stats.nodes = (int)kn.size();
    const size_t n = kn.size();
    KrylovVector x(n), b(n);
    for (size_t k = 0; k < n; ++k) x[k] = b[k] = kn[k]->velocity;
    KrylovVector Ax = apply_A(g, snow, st, kn, x, prm);
    KrylovVector r(n);
    for (size_t k = 0; k < n; ++k) r[k] = sub(b[k], Ax[k]);
    const double bb = mdot(kn, b, b), tol2 = prm.cr_tolerance * prm.cr_tolerance * bb;
    double rr = mdot(kn, r, r);
    stats.residual[0] = std::sqrt(rr);
    const int maxit = std::min(prm.cr_max_iterations, 256);
    if (rr > tol2) {
        KrylovVector p = r, Ar = apply_A(g, snow, st, kn, r, prm), Ap = Ar;
        double rho = mdot(kn, r, Ar);
        for (int it = 0; it < maxit; ++it) {
            const double den = mdot(kn, Ap, Ap);
            if (!(den > 0.0) || rho == 0.0) break;
            const double alpha = rho / den;
            for (size_t k = 0; k < n; ++k) {
                x[k] = add(x[k], scl(alpha, p[k]));
                r[k] = sub(r[k], scl(alpha, Ap[k]));
            }
            rr = mdot(kn, r, r);
            stats.iterations = it + 1;
            stats.residual[it + 1] = std::sqrt(rr);
            if (rr <= tol2) break;
            Ar = apply_A(g, snow, st, kn, r, prm);
            const double rho_new = mdot(kn, r, Ar);
            const double bet = rho_new / rho;
            rho = rho_new;
            for (size_t k = 0; k < n; ++k) {
                p[k] = add(r[k], scl(bet, p[k]));
                Ap[k] = add(Ar[k], scl(bet, Ap[k]));
            }
        }
    }
    for (size_t k = 0; k < n; ++k) kn[k]->velocity = x[k];
    g_cr_stats = stats;
}

double ms_since(std::chrono::steady_clock::time_point t0, std::chrono::steady_clock::time_point t1) {
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

// step, grid.cpp:359-450, on a particle vector and a caller-owned stencil
// vector (main.cpp:38).
void step_impl(std::vector<Particle>& snow, std::vector<Stencil>& stencils, const orc_collider* cs, int nc,
               const orc_params& prm, orc_timings* t, Grid* grid_out) {
    Grid grid;
    const double dx = prm.dx, dt = prm.dt;
    const V3 gravity = v3(prm.gravity[0], prm.gravity[1], prm.gravity[2]);
    const long n = (long)snow.size();
    using clk = std::chrono::steady_clock;
    orc_timings dummy{};
    if (!t) t = &dummy;

    auto t0 = clk::now();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) build_stencil(snow[i], dx, stencils[i]);
    auto t1 = clk::now();
    t->stencil += ms_since(t0, t1);

    t0 = clk::now();
    for (long i = 0; i < n; ++i) p2g(grid, snow[i], stencils[i]);
    t1 = clk::now();
    t->p2g += ms_since(t0, t1);

    std::vector<M3> stress(n);
    t0 = clk::now();
    for (long i = 0; i < n; ++i) {
        const Particle& p = snow[i];
        double mu0 = p.mu0 >= 0 ? p.mu0 : prm.mu0;
        double la0 = p.lambda0 >= 0 ? p.lambda0 : prm.lambda0;
        stress[i] = force_constitutive(p, mu0, la0, prm.hardening);
    }
    t1 = clk::now();
    t->forces_constitutive += ms_since(t0, t1);

    t0 = clk::now();
    for (long i = 0; i < n; ++i) force_grid_accumulation(stress[i], snow[i], stencils[i]);
    t1 = clk::now();
    t->forces_accumulative += ms_since(t0, t1);

    t0 = clk::now();
    grid_velocity(grid, gravity, dt);
    t1 = clk::now();
    t->grid_velocity += ms_since(t0, t1);

    t0 = clk::now();
    grid_collisions(grid, cs, nc, dx);
    t1 = clk::now();
    t->grid_collisions += ms_since(t0, t1);

    if (prm.beta > 0.0) conjugate_residual(grid, snow, stencils, prm);

    t0 = clk::now();
#pragma omp parallel for
    for (long i = 0; i < n; ++i) stress_update(grid, snow[i], dx, dt, prm.max_compression, prm.max_stretch);
    t1 = clk::now();
    t->stress_update += ms_since(t0, t1);

    t0 = clk::now();
#pragma omp parallel for
    for (long i = 0; i < n; ++i) g2p(grid, snow[i], dx, prm.alpha);
    t1 = clk::now();
    t->g2p += ms_since(t0, t1);

    t0 = clk::now();
    for (long i = 0; i < n; ++i) particle_collisions(snow[i], cs, nc);
    t1 = clk::now();
    t->particle_collisions += ms_since(t0, t1);

    t0 = clk::now();
    for (long i = 0; i < n; ++i) {
        position_update(snow[i], dt);
        position_projection(snow[i], cs, nc);
    }
    t1 = clk::now();
    t->position_update += ms_since(t0, t1);

    if (grid_out) *grid_out = std::move(grid);
}

}  // namespace

struct orc_grid {
    Grid g;
};

extern "C" {

double orc_N(double a, double b) { return Nfun(a, b); }
double orc_N_dash(double g, double p) { return Ndash(g, p); }
double orc_det3(const double A[9]) { return det3(A); }
int orc_inv3(const double A[9], double Ainv[9]) {
    M3 M;
    std::memcpy(M.a, A, sizeof M.a);
    if (det3(A) == 0.0) return -1;
    M3 r = inv3(M);
    std::memcpy(Ainv, r.a, sizeof r.a);
    return 0;
}
int orc_svd3(const double A[9], double U[9], double S[3], double V[9]) { return svd3(A, U, S, V); }

orc_collider orc_make_plane(const double point[3], const double normal[3], double friction) {
    orc_collider c{};
    c.type = 0;
    double l = std::sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
    for (int a = 0; a < 3; ++a) {
        c.point[a] = point[a];
        c.normal[a] = normal[a] / l;
    }
    c.friction = friction;
    return c;
}
orc_collider orc_make_box(const double centre[3], const double h[3], double friction) {
    orc_collider c{};
    c.type = 1;
    for (int a = 0; a < 3; ++a) {
        c.point[a] = centre[a];
        c.half_extent[a] = h[a];
    }
    c.friction = friction;
    return c;
}
orc_collider orc_make_sphere(const double centre[3], double radius, double friction) {
    orc_collider c{};
    c.type = 2;
    for (int a = 0; a < 3; ++a) c.point[a] = centre[a];
    c.radius = radius;
    c.friction = friction;
    return c;
}
double orc_collider_phi(const orc_collider* c, const double x[3]) {
    return coll_phi(*c, v3(x[0], x[1], x[2]), nullptr);
}
void orc_collision_helper(const double x[3], const double v[3], const orc_collider* c, double out[3]) {
    V3 r = collision_helper(v3(x[0], x[1], x[2]), v3(v[0], v[1], v[2]), *c);
    for (int a = 0; a < 3; ++a) out[a] = r[a];
}

orc_grid* orc_grid_new(void) { return new orc_grid(); }
void orc_grid_free(orc_grid* g) { delete g; }
size_t orc_grid_size(const orc_grid* g) { return g->g.size(); }
void orc_grid_set(orc_grid* g, int i, int j, int k, double mass, const double mom[3], const double vel[3],
                  const double f[3]) {
    Node& nd = g->g[Key{i, j, k}];
    nd.mass = mass;
    for (int a = 0; a < 3; ++a) {
        nd.momentum[a] = mom ? mom[a] : 0.0;
        nd.velocity[a] = vel ? vel[a] : 0.0;
        nd.force[a] = f ? f[a] : 0.0;
    }
}
int orc_grid_get(const orc_grid* g, int i, int j, int k, double* mass, double mom[3], double vel[3],
                 double f[3]) {
    auto it = g->g.find(Key{i, j, k});
    if (it == g->g.end()) return 0;
    const Node& nd = it->second;
    if (mass) *mass = nd.mass;
    for (int a = 0; a < 3; ++a) {
        if (mom) mom[a] = nd.momentum[a];
        if (vel) vel[a] = nd.velocity[a];
        if (f) f[a] = nd.force[a];
    }
    return 1;
}
size_t orc_grid_dump(const orc_grid* g, int* idx, double* mass, double* mom, double* vel, double* f) {
    size_t n = 0;
    for (const auto& kv : g->g) {
        if (idx) {
            idx[n * 3 + 0] = kv.first.i;
            idx[n * 3 + 1] = kv.first.j;
            idx[n * 3 + 2] = kv.first.k;
        }
        if (mass) mass[n] = kv.second.mass;
        for (int a = 0; a < 3; ++a) {
            if (mom) mom[n * 3 + a] = kv.second.momentum[a];
            if (vel) vel[n * 3 + a] = kv.second.velocity[a];
            if (f) f[n * 3 + a] = kv.second.force[a];
        }
        ++n;
    }
    return n;
}

void orc_build_stencil(const orc_particle* p, double dx, int idx[64 * 3], double w[64], double grad[64 * 3]) {
    Stencil st;
    build_stencil(from_c(*p), dx, st);
    for (int n = 0; n < 64; ++n) {
        idx[n * 3 + 0] = st[n].index.i;
        idx[n * 3 + 1] = st[n].index.j;
        idx[n * 3 + 2] = st[n].index.k;
        w[n] = st[n].weight;
        for (int a = 0; a < 3; ++a) grad[n * 3 + a] = st[n].gradient[a];
    }
}
// P2G(grid, particle, grid_spacing), grid.cpp:131-136.
void orc_p2g(orc_grid* g, const orc_particle* p, double dx) {
    Particle q = from_c(*p);
    Stencil st;
    build_stencil(q, dx, st);
    p2g(g->g, q, st);
}
void orc_calculate_volume(orc_grid* g, orc_particle* p, double dx) {
    Particle q = from_c(*p);
    calculate_volume(g->g, q, dx);
    to_c(q, *p);
}
void orc_force_constitutive(const orc_particle* p, double mu0, double lambda0, double xi, double P[9]) {
    M3 r = force_constitutive(from_c(*p), mu0, lambda0, xi);
    std::memcpy(P, r.a, sizeof r.a);
}
// force_calculation, grid.cpp:184-193.
void orc_force_calculation(orc_grid* g, const orc_particle* p, double dx, double mu0, double lambda0, double xi) {
    Particle q = from_c(*p);
    M3 P = force_constitutive(q, mu0, lambda0, xi);
    Stencil st;
    build_stencil(q, dx, st);
    for (Entry& e : st) e.node = &g->g[e.index];
    force_grid_accumulation(P, q, st);
}
void orc_grid_velocity(orc_grid* g, const double gravity[3], double dt) {
    grid_velocity(g->g, v3(gravity[0], gravity[1], gravity[2]), dt);
}
void orc_grid_collisions(orc_grid* g, const orc_collider* c, int nc, double dx) { grid_collisions(g->g, c, nc, dx); }
void orc_stress_update(const orc_grid* g, orc_particle* p, double dx, double dt, double tc, double ts) {
    Particle q = from_c(*p);
    stress_update(g->g, q, dx, dt, tc, ts);
    to_c(q, *p);
}
void orc_g2p(const orc_grid* g, orc_particle* p, double dx, double alpha) {
    Particle q = from_c(*p);
    g2p(g->g, q, dx, alpha);
    to_c(q, *p);
}
void orc_particle_collisions(orc_particle* p, const orc_collider* c, int nc) {
    Particle q = from_c(*p);
    particle_collisions(q, c, nc);
    to_c(q, *p);
}
void orc_position_update(orc_particle* p, double dt) {
    Particle q = from_c(*p);
    position_update(q, dt);
    to_c(q, *p);
}
void orc_position_projection(orc_particle* p, const orc_collider* c, int nc) {
    Particle q = from_c(*p);
    position_projection(q, c, nc);
    to_c(q, *p);
}

// setup, grid.cpp:340-357.
void orc_constitutive_differential(const orc_particle* p, const double dF[9], double mu0, double lambda0, double xi,
                                   double dP[9]) {
    M3 d;
    std::memcpy(d.a, dF, sizeof d.a);
    const M3 r = constitutive_differential(from_c(*p), d, mu0, lambda0, xi);
    std::memcpy(dP, r.a, sizeof r.a);
}

size_t orc_krylov_nodes(const orc_grid* g, int* idx) {
    std::vector<Key> keys;
    for (const auto& kv : g->g)
        if (kv.second.mass > 0.0 && !kv.second.constrained) keys.push_back(kv.first);
    std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& b) {
        if (a.i != b.i) return a.i < b.i;
        if (a.j != b.j) return a.j < b.j;
        return a.k < b.k;
    });
    if (idx)
        for (size_t k = 0; k < keys.size(); ++k) {
            idx[k * 3] = keys[k].i;
            idx[k * 3 + 1] = keys[k].j;
            idx[k * 3 + 2] = keys[k].k;
        }
    return keys.size();
}

void orc_apply_A(orc_grid* g, const orc_particle* ps, size_t n, const orc_params* prm, const double* q, double* y) {
    std::vector<Particle> snow(n);
    std::vector<Stencil> st(n);
    for (size_t i = 0; i < n; ++i) {
        snow[i] = from_c(ps[i]);
        build_stencil(snow[i], prm->dx, st[i]);
        for (Entry& e : st[i]) e.node = &g->g[e.index];
    }
    const std::vector<Node*> kn = index_nodes(g->g);
    KrylovVector qv(kn.size());
    for (size_t k = 0; k < kn.size(); ++k) qv[k] = v3(q[k * 3], q[k * 3 + 1], q[k * 3 + 2]);
    const KrylovVector r = apply_A(g->g, snow, st, kn, qv, *prm);
    for (size_t k = 0; k < kn.size(); ++k)
        for (int a = 0; a < 3; ++a) y[k * 3 + a] = r[k][a];
}

void orc_last_cr_stats(orc_cr_stats* out) { *out = g_cr_stats; }

void orc_setup(orc_particle* ps, size_t n, double dx) {
    std::vector<Particle> snow(n);
    for (size_t i = 0; i < n; ++i) snow[i] = from_c(ps[i]);
    Grid grid;
    std::vector<Stencil> st(n);
    for (size_t i = 0; i < n; ++i) build_stencil(snow[i], dx, st[i]);
    for (size_t i = 0; i < n; ++i) p2g(grid, snow[i], st[i]);
    for (size_t i = 0; i < n; ++i) calculate_volume(grid, snow[i], dx);
    for (size_t i = 0; i < n; ++i) to_c(snow[i], ps[i]);
}

void orc_step(orc_particle* ps, size_t n, const orc_collider* c, int nc, const orc_params* prm, orc_timings* t,
              int nthreads, orc_grid* grid_out) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    std::vector<Particle> snow(n);
    for (size_t i = 0; i < n; ++i) snow[i] = from_c(ps[i]);
    std::vector<Stencil> st(n);
    step_impl(snow, st, c, nc, *prm, t, grid_out ? &grid_out->g : nullptr);
    for (size_t i = 0; i < n; ++i) to_c(snow[i], ps[i]);
}

void orc_run(orc_particle* ps, size_t n, const orc_collider* c, int nc, const orc_params* prm, int nsteps,
             orc_timings* t, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    std::vector<Particle> snow(n);
    for (size_t i = 0; i < n; ++i) snow[i] = from_c(ps[i]);
    std::vector<Stencil> st(n);  // caller-owned scratch, main.cpp:38
    for (int s = 0; s < nsteps; ++s) step_impl(snow, st, c, nc, *prm, t, nullptr);
    for (size_t i = 0; i < n; ++i) to_c(snow[i], ps[i]);
}

}  // extern "C"
