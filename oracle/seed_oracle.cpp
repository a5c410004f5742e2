// seed_oracle.cpp -- CPU restatement of the reference's mesh seeding.
//
// TEST INFRASTRUCTURE ONLY (part of liboracle.so, see snow_oracle.h): the
// checker for the GPU seeding entry point, never linked into the product.
//
// Restates blender/seed_export.py:43-131 without Blender:
//   world_bounds       (43-62)   axis-aligned bounds of the mesh, float32
//   ray_inside         (64-86)   hits counted along a ray by re-casting from
//                                just past each hit (256 casts at most,
//                                epsilon 1e-5), inside when the count is odd
//   point_inside_mesh  (88-104)  majority vote (>= 2) of three fixed rays
//   seed_mesh          (106-131) lattice at min + spacing/2 + k spacing per
//                                axis (Python double accumulation), x outer,
//                                z inner, points stored as float32
// Blender's own pieces are restated from their published definitions:
// mathutils vectors are float32 (positions, directions, the stored settings),
// Vector.normalized() scales by 1/sqrt(dot) in float32, and Object.ray_cast
// returns the nearest double-sided hit at distance >= 0 (below
// 1.70141e38, the default distance) at location origin + t dir.  The
// ray/triangle test is the watertight test of Woop, Benthin and Wald,
// "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013 (shear to the
// ray's dominant axis, edge functions U, V, W redone in double when one is
// zero, no back-face culling), with t = T/det correctly rounded.  Whether
// Blender's BVH ray caster agrees bit for bit on points within rounding of
// the surface cannot be checked here (bpy is not available): parity with
// Blender is unpinned for such points; the tests pin this restatement by
// closed forms (boxes, convex polyhedra, a box with a cavity) and the lattice
// by the add-on's own output shipped as the reference's input/particles.csv.
//
// Compiled without floating-point contraction (-ffp-contract=off): every
// product and sum is rounded on its own, as the GPU kernel does.
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "snow_oracle.h"

namespace {

constexpr float kDistMax = 1.70141183e38f;  // Object.ray_cast default distance
constexpr float kEps = 1e-5f;               // ray_inside's epsilon (seed_export.py:68)
constexpr int kMaxCasts = 256;              // seed_export.py:70

struct SeedRay {
    float d[3];
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

// RAY_DIRECTIONS (seed_export.py:7-11), normalized in float32.
SeedRay make_ray(float a, float b, float c) {
    SeedRay r;
    const float v[3] = {a, b, c};
    const float dot = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const float inv = 1.0f / std::sqrt(dot);
    for (int i = 0; i < 3; ++i) r.d[i] = v[i] * inv;
    int kz = 0;
    for (int i = 1; i < 3; ++i)
        if (std::fabs(r.d[i]) > std::fabs(r.d[kz])) kz = i;
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (r.d[kz] < 0.0f) {
        const int t = kx;
        kx = ky;
        ky = t;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.Sx = r.d[kx] / r.d[kz];
    r.Sy = r.d[ky] / r.d[kz];
    r.Sz = 1.0f / r.d[kz];
    return r;
}

const SeedRay& ray(int k) {
    static const SeedRay rays[3] = {make_ray(1.0f, 0.371f, 0.529f), make_ray(-0.427f, 1.0f, 0.233f),
                                    make_ray(0.191f, -0.463f, 1.0f)};
    return rays[k];
}

// Watertight test of the ray (o, r) against triangle (A, B, C); on a hit at
// distance t >= 0 returns true.
bool hit_triangle(const SeedRay& r, const float o[3], const float* A, const float* B, const float* C, float& t) {
    float a[3], b[3], c[3];
    for (int i = 0; i < 3; ++i) {
        a[i] = A[i] - o[i];
        b[i] = B[i] - o[i];
        c[i] = C[i] - o[i];
    }
    const float ax = a[r.kx] - r.Sx * a[r.kz], ay = a[r.ky] - r.Sy * a[r.kz];
    const float bx = b[r.kx] - r.Sx * b[r.kz], by = b[r.ky] - r.Sy * b[r.kz];
    const float cx = c[r.kx] - r.Sx * c[r.kz], cy = c[r.ky] - r.Sy * c[r.kz];
    float U = cx * by - cy * bx;
    float V = ax * cy - ay * cx;
    float W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)cx * (double)by - (double)cy * (double)bx);
        V = (float)((double)ax * (double)cy - (double)ay * (double)cx);
        W = (float)((double)bx * (double)ay - (double)by * (double)ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * a[r.kz], Bz = r.Sz * b[r.kz], Cz = r.Sz * c[r.kz];
    const float T = U * Az + V * Bz + W * Cz;
    if ((det > 0.0f && T < 0.0f) || (det < 0.0f && T > 0.0f)) return false;
    t = (float)((double)T / (double)det);
    return true;
}

// ray_inside (seed_export.py:64-86): casts until nothing is hit.
int count_hits(const float* tri, size_t ntris, const float p[3], const SeedRay& r) {
    float o[3] = {p[0], p[1], p[2]};
    int hits = 0;
    for (int cast = 0; cast < kMaxCasts; ++cast) {
        float best = kDistMax;
        bool any = false;
        for (size_t k = 0; k < ntris; ++k) {
            float t;
            if (hit_triangle(r, o, tri + 9 * k, tri + 9 * k + 3, tri + 9 * k + 6, t) && t < best) {
                best = t;
                any = true;
            }
        }
        if (!any) break;
        ++hits;
        for (int i = 0; i < 3; ++i) {
            const float loc = o[i] + r.d[i] * best;
            o[i] = loc + r.d[i] * kEps;
        }
    }
    return hits;
}

std::vector<float> triangles(const double* verts, size_t nverts, const int* tris, size_t ntris, bool& ok) {
    std::vector<float> t(9 * ntris);
    ok = true;
    for (size_t k = 0; k < ntris; ++k)
        for (int v = 0; v < 3; ++v) {
            const int id = tris[3 * k + v];
            if (id < 0 || (size_t)id >= nverts) {
                ok = false;
                return t;
            }
            for (int i = 0; i < 3; ++i) t[9 * k + 3 * v + i] = (float)verts[3 * (size_t)id + i];
        }
    return t;
}

}  // namespace

extern "C" {

size_t orc_seed_axis(float lo, float hi, double spacing, float* out, size_t cap) {
    const double s = (double)(float)spacing;  // the add-on's spacing is a float32 property
    size_t n = 0;
    for (double x = (double)lo + s * 0.5; x < (double)hi; x += s) {
        if (out && n < cap) out[n] = (float)x;
        ++n;
    }
    return n;
}

void orc_seed_direction(int k, float d[3]) {
    for (int i = 0; i < 3; ++i) d[i] = ray(k).d[i];
}

int orc_seed_ray_hits(const double* verts, size_t nverts, const int* tris, size_t ntris, const double p[3], int k) {
    bool ok;
    const std::vector<float> t = triangles(verts, nverts, tris, ntris, ok);
    if (!ok || k < 0 || k > 2) return -1;
    const float q[3] = {(float)p[0], (float)p[1], (float)p[2]};
    return count_hits(t.data(), ntris, q, ray(k));
}

long orc_seed_mesh(const double* verts, size_t nverts, const int* tris, size_t ntris, double spacing, double* out,
                   size_t cap) {
    bool ok;
    const std::vector<float> t = triangles(verts, nverts, tris, ntris, ok);
    if (!ok || !(spacing > 0.0)) return -1;
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
    if (nverts > 0)
        for (int i = 0; i < 3; ++i) {
            lo[i] = hi[i] = (float)verts[i];
            for (size_t v = 1; v < nverts; ++v) {
                const float c = (float)verts[3 * v + i];
                lo[i] = std::fmin(lo[i], c);
                hi[i] = std::fmax(hi[i], c);
            }
        }
    std::vector<float> ax[3];
    for (int i = 0; i < 3; ++i) {
        ax[i].resize(orc_seed_axis(lo[i], hi[i], spacing, nullptr, 0));
        orc_seed_axis(lo[i], hi[i], spacing, ax[i].data(), ax[i].size());
    }
    long n = 0;
    for (float x : ax[0])
        for (float y : ax[1])
            for (float z : ax[2]) {
                const float p[3] = {x, y, z};
                int votes = 0;
                for (int k = 0; k < 3; ++k) votes += count_hits(t.data(), ntris, p, ray(k)) & 1;
                if (votes >= 2) {
                    if (out && (size_t)n < cap) {
                        out[3 * n] = x;
                        out[3 * n + 1] = y;
                        out[3 * n + 2] = z;
                    }
                    ++n;
                }
            }
    return n;
}

}  // extern "C"
