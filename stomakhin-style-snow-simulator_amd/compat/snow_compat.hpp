// snow_compat.hpp -- C++20 drop-in for the reference solver's public API,
// running on the MI355X engine through the C ABI (include/snowmpm.h).
//
// Mirrors (paths relative to the reference repo):
//   include/snow/particle.hpp:5-28   Particle{x_p, v_p, m_p, V_p0, F_P, F_E}
//   include/snow/collision.hpp:6-14  CollisionBody, make_plane
//   include/snow/debug.hpp:3-15      StepTimings, print_timings
//   include/snow/grid.hpp:96-97      setup(), step()
//   include/snow/import.hpp:25       import_particles
//   include/snow/export.hpp:8        export_particles
// so a main.cpp-style driver compiles against it unchanged apart from the
// include.  Vector types are Eigen's when <Eigen/Dense> is available and a
// minimal value type with the same accessors otherwise.
//
// Semantics: step() is stateless like the reference: it takes the particle
// vector, advances it one step on the GPU and writes it back (x, v, F_E, F_P,
// V_p0).  Between calls the state stays resident on the device; it is only
// re-uploaded when the caller changed the vector.  `stencils` is accepted and
// ignored (the engine builds no per-particle stencils).  Failures throw
// std::runtime_error, as the reference's I/O does.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "snowmpm.h"

#if __has_include(<Eigen/Dense>)
#include <Eigen/Dense>
namespace snow_compat {
using Vec3 = Eigen::Vector3d;
using Mat3 = Eigen::Matrix3d;
inline Mat3 identity3() { return Mat3::Identity(); }
inline double m3(const Mat3& A, int r, int c) { return A(r, c); }
inline double& m3(Mat3& A, int r, int c) { return A(r, c); }
}  // namespace snow_compat
#else
namespace snow_compat {
struct Vec3 {
    double d[3]{0, 0, 0};
    Vec3() = default;
    Vec3(double a, double b, double c) : d{a, b, c} {}
    double& operator[](int i) { return d[i]; }
    double operator[](int i) const { return d[i]; }
    double& operator()(int i) { return d[i]; }
    double operator()(int i) const { return d[i]; }
    double x() const { return d[0]; }
    double y() const { return d[1]; }
    double z() const { return d[2]; }
    bool operator==(const Vec3& o) const { return d[0] == o.d[0] && d[1] == o.d[1] && d[2] == o.d[2]; }
};
struct Mat3 {
    double d[9]{1, 0, 0, 0, 1, 0, 0, 0, 1};
    double& operator()(int r, int c) { return d[r * 3 + c]; }
    double operator()(int r, int c) const { return d[r * 3 + c]; }
    bool operator==(const Mat3& o) const { return std::memcmp(d, o.d, sizeof d) == 0; }
};
inline Mat3 identity3() { return Mat3{}; }
inline double m3(const Mat3& A, int r, int c) { return A(r, c); }
inline double& m3(Mat3& A, int r, int c) { return A(r, c); }
}  // namespace snow_compat
#endif

namespace snow_compat {

// Particle, particle.hpp:5-28.
struct Particle {
    Vec3 x_p;
    Vec3 v_p;
    double m_p;
    double V_p0;
    Mat3 F_P;
    Mat3 F_E;
    Particle(const Vec3& x, const Vec3& v, double m, double V)
        : x_p(x), v_p(v), m_p(m), V_p0(V), F_P(identity3()), F_E(identity3()) {}
};

// CollisionBody, collision.hpp:6-12: the device needs a POD description, so
// the body carries one; phi/normal/velocity stay available on the host.
// Only bodies built by make_plane (or make_box / make_sphere below) have a
// description: step() throws for a body assembled from closures of its own
// (arbitrary SDF closures cannot run on the GPU), and for a body whose
// friction or velocity closure was changed afterwards; set_velocity() gives a
// body a constant velocity (the v_c of collision_helper, grid.cpp:206-231).
struct CollisionBody {
    std::function<double(const Vec3&)> phi;
    std::function<Vec3(const Vec3&)> normal;
    std::function<Vec3(const Vec3&)> velocity;
    double friction = 0.0;
    mpm_collider desc{};
    bool has_desc = false;
};

// make_plane, collision.hpp:14 / collision.cpp:5-27.
inline CollisionBody make_plane(const Vec3& point, const Vec3& normal, double friction) {
    const double p[3] = {point[0], point[1], point[2]}, n[3] = {normal[0], normal[1], normal[2]};
    mpm_collider d = mpm_make_plane(p, n, friction);
    const Vec3 un(d.normal[0], d.normal[1], d.normal[2]);
    CollisionBody b;
    b.phi = [point, un](const Vec3& x) {
        return un[0] * (x[0] - point[0]) + un[1] * (x[1] - point[1]) + un[2] * (x[2] - point[2]);
    };
    b.normal = [un](const Vec3&) { return un; };
    b.velocity = [](const Vec3&) { return Vec3(0, 0, 0); };
    b.friction = friction;
    b.desc = d;
    b.has_desc = true;
    return b;
}

namespace detail {
inline CollisionBody sdf_body(const mpm_collider& d) {
    CollisionBody b;
    b.desc = d;
    b.has_desc = true;
    b.friction = d.friction;
    b.phi = [d](const Vec3& x) {
        if (d.type == MPM_COLLIDER_SPHERE) {
            const double r0 = x[0] - d.point[0], r1 = x[1] - d.point[1], r2 = x[2] - d.point[2];
            return std::sqrt(r0 * r0 + r1 * r1 + r2 * r2) - d.radius;
        }
        double q[3], out = 0.0, qmax = -1e300;
        for (int a = 0; a < 3; ++a) {
            q[a] = std::fabs(x[a] - d.point[a]) - d.half_extent[a];
            qmax = std::max(qmax, q[a]);
            if (q[a] > 0) out += q[a] * q[a];
        }
        return std::sqrt(out) + std::min(qmax, 0.0);
    };
    b.velocity = [](const Vec3&) { return Vec3(0, 0, 0); };
    return b;
}
}  // namespace detail

// Solid box and sphere bodies (no counterpart in the reference, whose only
// factory is make_plane).
inline CollisionBody make_box(const Vec3& centre, const Vec3& half_extent, double friction) {
    const double c[3] = {centre[0], centre[1], centre[2]}, h[3] = {half_extent[0], half_extent[1], half_extent[2]};
    return detail::sdf_body(mpm_make_box(c, h, friction));
}
inline CollisionBody make_sphere(const Vec3& centre, double radius, double friction) {
    const double c[3] = {centre[0], centre[1], centre[2]};
    return detail::sdf_body(mpm_make_sphere(c, radius, friction));
}

// A constant body velocity: the velocity closure and the device description
// change together.
inline CollisionBody& set_velocity(CollisionBody& b, const Vec3& v) {
    for (int a = 0; a < 3; ++a) b.desc.velocity[a] = v[a];
    b.velocity = [v](const Vec3&) { return v; };
    return b;
}

// StepTimings, debug.hpp:3-15.
struct StepTimings {
    double stencil{0.0};
    double p2g{0.0};
    double forces_constitutive{0.0};
    double forces_accumulative{0.0};
    double grid_velocity{0.0};
    double grid_collisions{0.0};
    double stress_update{0.0};
    double g2p{0.0};
    double particle_collisions{0.0};
    double position_update{0.0};
};

// print_timings, debug.cpp:5-30 (same lines and units).
inline double print_timings(StepTimings t) {
    const double measured = t.stencil + t.p2g + t.forces_constitutive + t.forces_accumulative + t.grid_velocity +
                            t.grid_collisions + t.stress_update + t.g2p + t.particle_collisions + t.position_update;
    auto line = [&](const char* name, double v) {
        std::cout << name << ": " << v << "ms  " << 100.0 * v / measured << "% \n";
    };
    line("Stencil", t.stencil);
    line("P2G", t.p2g);
    line("Forces constitutive", t.forces_constitutive);
    line("Forces accumulative", t.forces_accumulative);
    line("Grid velocity", t.grid_velocity);
    line("Grid collisions", t.grid_collisions);
    line("Stress update", t.stress_update);
    line("G2P", t.g2p);
    line("Particle collisions", t.particle_collisions);
    line("Position update", t.position_update);
    return measured;
}

// The reference caches per-particle stencils in a caller-owned vector
// (main.cpp:38); the engine builds none, so the type is an empty placeholder.
struct ParticleStencil {};

namespace detail {

inline void check(int rc, const char* what) {
    if (rc != MPM_OK) throw std::runtime_error(std::string(what) + ": " + mpm_last_error());
}

// One resident engine per process, re-created when the particle count grows.
struct Resident {
    mpm_engine* eng = nullptr;
    int sid = -1;
    size_t cap = 0;
    std::vector<Particle> shadow;  // state as last written back to the caller
    std::vector<mpm_collider> coll;
    mpm_params prm{};
    ~Resident() {
        if (eng) mpm_destroy(eng);
    }
};
inline Resident& resident() {
    static Resident r;
    return r;
}

inline bool same(const std::vector<Particle>& a, const std::vector<Particle>& b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i) {
        const Particle &p = a[i], &q = b[i];
        if (!(p.x_p == q.x_p) || !(p.v_p == q.v_p) || p.m_p != q.m_p || p.V_p0 != q.V_p0 || !(p.F_P == q.F_P) ||
            !(p.F_E == q.F_E))
            return false;
    }
    return true;
}

inline bool same_params(const mpm_params& a, const mpm_params& b) { return std::memcmp(&a, &b, sizeof a) == 0; }

inline void upload(Resident& r, const std::vector<Particle>& snow, const mpm_params& prm,
                   const std::vector<mpm_collider>& coll, bool with_volume) {
    const size_t n = snow.size();
    if (!r.eng || r.cap < n) {
        if (r.eng) mpm_destroy(r.eng);
        r.eng = nullptr;
        mpm_config cfg{};
        cfg.particle_capacity = n + n / 4 + 1024;
        cfg.track_fp = 1;
        check(mpm_create(&cfg, &r.eng), "mpm_create");
        r.cap = cfg.particle_capacity;
        r.sid = -1;
    }
    if (r.sid >= 0) check(mpm_clear(r.eng), "mpm_clear");
    std::vector<double> x(3 * n), v(3 * n), m(n), V0(n), Fe(9 * n), Fp(9 * n);
    for (size_t i = 0; i < n; ++i) {
        const Particle& p = snow[i];
        for (int a = 0; a < 3; ++a) {
            x[i * 3 + a] = p.x_p[a];
            v[i * 3 + a] = p.v_p[a];
        }
        m[i] = p.m_p;
        V0[i] = p.V_p0;
        for (int rr = 0; rr < 3; ++rr)
            for (int c = 0; c < 3; ++c) {
                Fe[i * 9 + rr * 3 + c] = m3(p.F_E, rr, c);
                Fp[i * 9 + rr * 3 + c] = m3(p.F_P, rr, c);
            }
    }
    check(mpm_scene_add(r.eng, &prm, coll.data(), (int)coll.size(), n, x.data(), v.data(), m.data(),
                        with_volume ? V0.data() : nullptr, Fe.data(), Fp.data(), nullptr, &r.sid),
          "mpm_scene_add");
    r.prm = prm;
    r.coll = coll;
}

inline void write_back(Resident& r, std::vector<Particle>& snow) {
    const size_t n = snow.size();
    std::vector<double> x(3 * n), v(3 * n), Fe(9 * n), Fp(9 * n), V0(n);
    check(mpm_scene_download(r.eng, r.sid, x.data(), v.data(), Fe.data(), Fp.data(), nullptr, V0.data()),
          "mpm_scene_download");
    for (size_t i = 0; i < n; ++i) {
        Particle& p = snow[i];
        for (int a = 0; a < 3; ++a) {
            p.x_p[a] = x[i * 3 + a];
            p.v_p[a] = v[i * 3 + a];
        }
        p.V_p0 = V0[i];
        for (int rr = 0; rr < 3; ++rr)
            for (int c = 0; c < 3; ++c) {
                m3(p.F_E, rr, c) = Fe[i * 9 + rr * 3 + c];
                m3(p.F_P, rr, c) = Fp[i * 9 + rr * 3 + c];
            }
    }
    r.shadow = snow;
}

}  // namespace detail

// setup, grid.hpp:96 / grid.cpp:340-357: V_p0 from the P2G mass.
inline void setup(std::vector<Particle>& snow, double grid_spacing) {
    if (snow.empty()) return;
    auto& r = detail::resident();
    mpm_params prm{};
    prm.dx = grid_spacing;
    prm.dt = 1.0;  // not used by setup
    detail::upload(r, snow, prm, {}, false);
    detail::check(mpm_setup(r.eng), "mpm_setup");
    detail::write_back(r, snow);
}

namespace detail {
// The device description of each body, after checking that it still
// describes the body (see CollisionBody).
inline std::vector<mpm_collider> descriptions(const std::vector<CollisionBody>& colliders) {
    std::vector<mpm_collider> coll;
    for (const auto& b : colliders) {
        if (!b.has_desc)
            throw std::runtime_error("step: a CollisionBody not built by make_plane/make_box/make_sphere "
                                     "(custom phi/normal closures cannot run on the GPU)");
        if (b.friction != b.desc.friction)
            throw std::runtime_error("step: CollisionBody::friction changed after the body was made");
        if (b.velocity) {
            const Vec3 probe = b.velocity(Vec3(0, 0, 0));
            if (probe[0] != b.desc.velocity[0] || probe[1] != b.desc.velocity[1] || probe[2] != b.desc.velocity[2])
                throw std::runtime_error("step: CollisionBody::velocity changed outside set_velocity()");
        }
        coll.push_back(b.desc);
    }
    return coll;
}
}  // namespace detail

// nsub calls of step() (grid.hpp:97 / grid.cpp:359-450) with the particle
// vector written back once at the end: the frame loop of a driver that only
// reads the particles between frames (src/main.cpp:46-57) needs no download
// per substep.  Same results as nsub calls of step().
inline void step_substeps(std::vector<Particle>& snow, const std::vector<CollisionBody>& colliders,
                          std::vector<ParticleStencil>& /*stencils*/, Vec3 gravity, double time_step,
                          double grid_spacing, double mu_0, double lambda_0, double hardening_coefficient,
                          double max_compression, double max_stretch, double alpha, StepTimings& timings, int nsub) {
    if (snow.empty() || nsub <= 0) return;
    auto& r = detail::resident();
    mpm_params prm{};
    for (int a = 0; a < 3; ++a) prm.gravity[a] = gravity[a];
    prm.dt = time_step;
    prm.dx = grid_spacing;
    prm.mu0 = mu_0;
    prm.lambda0 = lambda_0;
    prm.hardening = hardening_coefficient;
    prm.max_compression = max_compression;
    prm.max_stretch = max_stretch;
    prm.alpha = alpha;
    prm.nu = 0.2;
    const std::vector<mpm_collider> coll = detail::descriptions(colliders);
    const bool resident_ok = r.eng && r.sid >= 0 && detail::same_params(prm, r.prm) && coll.size() == r.coll.size() &&
                             std::memcmp(coll.data(), r.coll.data(), coll.size() * sizeof(mpm_collider)) == 0 &&
                             detail::same(snow, r.shadow);
    if (!resident_ok) detail::upload(r, snow, prm, coll, true);
    mpm_timings t{};
    detail::check(mpm_step(r.eng, nsub, &t), "mpm_step");
    timings.stencil += t.stencil;
    timings.p2g += t.p2g;
    timings.forces_constitutive += t.forces_constitutive;
    timings.forces_accumulative += t.forces_accumulative;
    timings.grid_velocity += t.grid_velocity;
    timings.grid_collisions += t.grid_collisions;
    timings.stress_update += t.stress_update;
    timings.g2p += t.g2p;
    timings.particle_collisions += t.particle_collisions;
    timings.position_update += t.position_update;
    detail::write_back(r, snow);
}

// step, grid.hpp:97 / grid.cpp:359-450.
inline void step(std::vector<Particle>& snow, const std::vector<CollisionBody>& colliders,
                 std::vector<ParticleStencil>& stencils, Vec3 gravity, double time_step, double grid_spacing,
                 double mu_0, double lambda_0, double hardening_coefficient, double max_compression,
                 double max_stretch, double alpha, StepTimings& timings) {
    step_substeps(snow, colliders, stencils, gravity, time_step, grid_spacing, mu_0, lambda_0, hardening_coefficient,
                  max_compression, max_stretch, alpha, timings, 1);
}

// import_particles, import.hpp:25 / import.cpp:131-137.
inline std::vector<Particle> import_particles(const std::string& groups, const std::string& seeds) {
    size_t n = 0;
    double *x = nullptr, *v = nullptr, *m = nullptr;
    const int rc = mpm_import_particles(groups.c_str(), seeds.c_str(), &n, &x, &v, &m);
    if (rc != MPM_OK) throw std::runtime_error(mpm_last_error());
    std::vector<Particle> out;
    out.reserve(n);
    for (size_t i = 0; i < n; ++i)
        out.emplace_back(Vec3(x[i * 3], x[i * 3 + 1], x[i * 3 + 2]), Vec3(v[i * 3], v[i * 3 + 1], v[i * 3 + 2]), m[i],
                         0.0);
    mpm_free(x);
    mpm_free(v);
    mpm_free(m);
    return out;
}

// export_particles, export.hpp:8 / export.cpp:10-22 (into output/).
inline void export_particles(const std::vector<Particle>& particles, int frame_no, const char* dir = "output") {
    std::vector<double> x(3 * particles.size());
    for (size_t i = 0; i < particles.size(); ++i)
        for (int a = 0; a < 3; ++a) x[i * 3 + a] = particles[i].x_p[a];
    detail::check(mpm_export_frame(dir, frame_no, particles.size(), x.data()), "export_particles");
}

}  // namespace snow_compat
