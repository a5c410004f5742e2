// mpm_scenes.cpp -- seeded synthetic workloads.
//
// Lattices follow the seeder's convention (points at min + (i + 1/2) *
// spacing, blender/seed_export.py:106-131) with spacing dx/2, i.e. 8
// particles per cell.  Randomness is counter based (splitmix64 of seed and
// index), so the CPU oracle and the GPU receive bit-identical inputs.
//   kind 0  snow cube, side 0.4 m, 64,000 particles, dx 0.02, v (0,-2,0)
//   kind 1  two tori R 0.3, r 0.12, dx 0.01, jitter +-0.1 spacing, +-5 m/s
//   kind 2  slab 4 x 0.5 x 4 m, 8,000,000 particles, dx 0.02, jitter,
//           per-particle E ~ U[1e5, 2e5], 0.2 m above a box
//   kind 3  one scene of the batched distribution: N ~ logU[1e3, 2e5],
//           box/sphere/torus, E ~ U[5e4, 2e5], theta +-30 %, |v| ~ U[0, 8]
// Material defaults (all kinds): rho 400, E 1.4e5, nu 0.2, theta_c 2.5e-2,
// theta_s 7.5e-3, xi 10, alpha 0.95, g (0,-9.81,0).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/snowmpm.h"

namespace snowmpm {
void set_error(const std::string& s);
}

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
// uniform in [0,1) from (seed, stream, index)
inline double u01(uint64_t seed, uint64_t stream, uint64_t idx) {
    uint64_t h = splitmix64(seed ^ splitmix64(stream * 0xD1B54A32D192ED03ull + idx));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

struct Built {
    std::vector<double> x, v, m, E;
};

constexpr double kRho = 400.0, kE = 1.4e5, kNu = 0.2;

void lame(double E, double nu, double& mu, double& la) {
    mu = E / (2 * (1 + nu));
    la = E * nu / ((1 + nu) * (1 - 2 * nu));
}

void base_params(mpm_params& p, double dx, double dt) {
    std::memset(&p, 0, sizeof p);
    p.gravity[1] = -9.81;
    p.dt = dt;
    p.dx = dx;
    lame(kE, kNu, p.mu0, p.lambda0);
    p.hardening = 10.0;
    p.max_compression = 2.5e-2;
    p.max_stretch = 7.5e-3;
    p.alpha = 0.95;
    p.nu = kNu;
}

// Fill the lattice of spacing h over [lo, hi) with the points inside(x,y,z).
void lattice(Built& b, const double lo[3], const double hi[3], double h,
             const std::function<bool(double, double, double)>& inside, const double vel[3], double jitter,
             uint64_t seed, uint64_t stream) {
    const long nx = (long)std::floor((hi[0] - lo[0]) / h + 1e-9), ny = (long)std::floor((hi[1] - lo[1]) / h + 1e-9),
               nz = (long)std::floor((hi[2] - lo[2]) / h + 1e-9);
    const double m = kRho * h * h * h;
    for (long i = 0; i < nx; ++i)
        for (long j = 0; j < ny; ++j)
            for (long k = 0; k < nz; ++k) {
                double p[3] = {lo[0] + (i + 0.5) * h, lo[1] + (j + 0.5) * h, lo[2] + (k + 0.5) * h};
                if (!inside(p[0], p[1], p[2])) continue;
                const uint64_t idx = (uint64_t)b.m.size();
                for (int a = 0; a < 3; ++a) {
                    double jt = jitter > 0 ? (2.0 * u01(seed, stream * 4 + a, idx) - 1.0) * jitter * h : 0.0;
                    b.x.push_back(p[a] + jt);
                    b.v.push_back(vel[a]);
                }
                b.m.push_back(m);
            }
}

mpm_collider plane_y(double y, double mu) {
    double pt[3] = {0, y, 0}, n[3] = {0, 1, 0};
    return mpm_make_plane(pt, n, mu);
}

}  // namespace

extern "C" int mpm_scene_generate(int kind, uint64_t seed, int index, double scale, size_t* n, double** x,
                                  double** v, double** m, double** E, mpm_params* prm, mpm_collider* colliders,
                                  int* ncolliders) {
    if (!n || !x || !v || !m || !prm || !colliders || !ncolliders || !(scale > 0) || scale > 1.0) {
        snowmpm::set_error("mpm_scene_generate: bad arguments");
        return MPM_ERR_ARG;
    }
    Built b;
    bool perE = false;
    *ncolliders = 0;
    if (kind == 0) {
        base_params(*prm, 0.02, 1e-4);
        const double h = 0.01, half = 0.2 * scale;
        const double lo[3] = {-half, 0.02, -half}, hi[3] = {half, 0.02 + 2 * half, half};
        const double vel[3] = {0, -2, 0};
        lattice(b, lo, hi, h, [](double, double, double) { return true; }, vel, 0.0, seed, 0);
        colliders[0] = plane_y(0.0, 0.2);
        *ncolliders = 1;
    } else if (kind == 1) {
        base_params(*prm, 0.01, 5e-5);
        const double h = 0.005, R = 0.3 * scale, r = 0.12 * scale, gap = 0.1 * scale;
        for (int t = 0; t < 2; ++t) {
            const double cx = (t == 0 ? -1 : 1) * (R + r + 0.5 * gap);
            const double cy = r + 0.01 + (t == 0 ? 0.0 : 0.05 * scale);
            const double lo[3] = {cx - R - r, cy - r, -R - r}, hi[3] = {cx + R + r, cy + r, R + r};
            const double vel[3] = {t == 0 ? 5.0 : -5.0, 0, 0};
            lattice(
                b, lo, hi, h,
                [=](double px, double py, double pz) {
                    const double q = std::sqrt((px - cx) * (px - cx) + pz * pz) - R;
                    return q * q + (py - cy) * (py - cy) <= r * r;
                },
                vel, 0.1, seed, 1);
        }
        colliders[0] = plane_y(0.0, 0.3);
        *ncolliders = 1;
    } else if (kind == 2) {
        base_params(*prm, 0.02, 5e-5);
        const double h = 0.01, hx = 2.0 * scale;
        const double lo[3] = {-hx, 0.2, -hx}, hi[3] = {hx, 0.7, hx};
        const double vel[3] = {0, 0, 0};
        lattice(b, lo, hi, h, [](double, double, double) { return true; }, vel, 0.1, seed, 2);
        perE = true;
        b.E.resize(b.m.size());
        for (size_t i = 0; i < b.m.size(); ++i) b.E[i] = 1e5 + 1e5 * u01(seed, 9, i);
        const double c[3] = {0, -0.5, 0}, he[3] = {1.5 * scale, 0.5, 1.5 * scale};
        colliders[0] = mpm_make_box(c, he, 0.2);
        colliders[1] = plane_y(-1.0, 0.2);
        *ncolliders = 2;
    } else if (kind == 3) {
        base_params(*prm, 0.02, 5e-5);
        const uint64_t st = 1000 + (uint64_t)index * 16;
        const double h = 0.01;
        const double N = std::exp(std::log(1e3) + (std::log(2e5) - std::log(1e3)) * u01(seed, st, 0)) * scale;
        const int shape = (int)(3.0 * u01(seed, st, 1)) % 3;
        const double Ey = 5e4 + 1.5e5 * u01(seed, st, 2);
        lame(Ey, kNu, prm->mu0, prm->lambda0);
        prm->max_compression = 2.5e-2 * (0.7 + 0.6 * u01(seed, st, 3));
        prm->max_stretch = 7.5e-3 * (0.7 + 0.6 * u01(seed, st, 4));
        const double speed = 8.0 * u01(seed, st, 5);
        const double cz = 2.0 * u01(seed, st, 6) - 1.0, ph = 2.0 * M_PI * u01(seed, st, 7);
        const double sz = std::sqrt(std::max(0.0, 1.0 - cz * cz));
        const double vel[3] = {speed * sz * std::cos(ph), speed * sz * std::sin(ph), speed * cz};
        double ext;  // half height of the shape
        std::function<bool(double, double, double)> inside;
        double lo[3], hi[3];
        if (shape == 0) {
            const double L = h * std::cbrt(N);
            ext = 0.5 * L;
            inside = [](double, double, double) { return true; };
            for (int a = 0; a < 3; ++a) {
                lo[a] = -0.5 * L;
                hi[a] = 0.5 * L;
            }
        } else if (shape == 1) {
            const double r = h * std::cbrt(3.0 * N / (4.0 * M_PI));
            ext = r;
            inside = [=](double px, double py, double pz) { return px * px + py * py + pz * pz <= r * r; };
            for (int a = 0; a < 3; ++a) {
                lo[a] = -r;
                hi[a] = r;
            }
        } else {
            const double r = h * std::cbrt(N / (2.0 * M_PI * M_PI * 2.5)), R = 2.5 * r;
            ext = r;
            inside = [=](double px, double py, double pz) {
                const double q = std::sqrt(px * px + pz * pz) - R;
                return q * q + py * py <= r * r;
            };
            lo[0] = lo[2] = -R - r;
            hi[0] = hi[2] = R + r;
            lo[1] = -r;
            hi[1] = r;
        }
        const double lift = ext + 0.05;
        lo[1] += lift;
        hi[1] += lift;
        lattice(
            b, lo, hi, h, [&](double px, double py, double pz) { return inside(px, py - lift, pz); }, vel, 0.0, seed,
            st);
        colliders[0] = plane_y(0.0, 0.2);
        *ncolliders = 1;
    } else {
        snowmpm::set_error("mpm_scene_generate: unknown kind");
        return MPM_ERR_ARG;
    }
    const size_t k = b.m.size();
    *n = k;
    *x = (double*)std::malloc(sizeof(double) * 3 * (k ? k : 1));
    *v = (double*)std::malloc(sizeof(double) * 3 * (k ? k : 1));
    *m = (double*)std::malloc(sizeof(double) * (k ? k : 1));
    if (E) *E = nullptr;
    if (!*x || !*v || !*m) return MPM_ERR_OOM;
    std::memcpy(*x, b.x.data(), sizeof(double) * 3 * k);
    std::memcpy(*v, b.v.data(), sizeof(double) * 3 * k);
    std::memcpy(*m, b.m.data(), sizeof(double) * k);
    if (perE && E) {
        *E = (double*)std::malloc(sizeof(double) * (k ? k : 1));
        if (!*E) return MPM_ERR_OOM;
        std::memcpy(*E, b.E.data(), sizeof(double) * k);
    }
    return MPM_OK;
}
