// snow_sim.cpp -- the reference driver (src/main.cpp:12-59) on the MI355X
// engine: the same call order (import -> setup -> export frame 0 -> frames of
// substeps -> print_timings -> export), written against snow_compat.hpp.
// Usage: snow_sim [groups.csv particles.csv [frames [substeps [outdir]]]]
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "snow_compat.hpp"

using namespace snow_compat;

int main(int argc, char** argv) {
    const std::string groups = argc > 1 ? argv[1] : "input/groups.csv";
    const std::string seeds = argc > 2 ? argv[2] : "input/particles.csv";
    const int frame_count = argc > 3 ? std::atoi(argv[3]) : 50;
    const int substeps_per_frame = argc > 4 ? std::atoi(argv[4]) : 80;
    const char* outdir = argc > 5 ? argv[5] : "output";

    // scene constants of main.cpp:14-29
    constexpr double grid_spacing = 0.05;
    constexpr double time_step = 0.0005;
    const double Youngs_modulus = 2.5e5, Poisson_ratio = 0.2;
    constexpr double hardening = 10.0, max_compression = 0.015, max_stretch = 0.0075, flip_ratio = 0.95;
    const double mu_0 = Youngs_modulus / (2 * (1 + Poisson_ratio));
    const double lambda_0 = (Youngs_modulus * Poisson_ratio) / ((1 + Poisson_ratio) * (1 - 2 * Poisson_ratio));

    try {
        std::vector<Particle> particles = import_particles(groups, seeds);
        std::cout << particles.size() << " particles\n";
        std::vector<CollisionBody> colliders;
        colliders.push_back(make_plane(Vec3(0, 0, 0), Vec3(0, 1, 0), 1));
        const Vec3 gravity(0, -10, 0);
        std::vector<ParticleStencil> stencils(particles.size());
        setup(particles, grid_spacing);
        export_particles(particles, 0, outdir);
        for (int frame = 1; frame <= frame_count; ++frame) {
            StepTimings timings;
            const auto t0 = std::chrono::steady_clock::now();
            for (int s = 0; s < substeps_per_frame; ++s)
                step(particles, colliders, stencils, gravity, time_step, grid_spacing, mu_0, lambda_0, hardening,
                     max_compression, max_stretch, flip_ratio, timings);
            const auto t1 = std::chrono::steady_clock::now();
            const double frame_time = std::chrono::duration<double, std::milli>(t1 - t0).count();
            std::cout << "Frame " << frame << ": " << frame_time << "ms\n";
            const double measured = print_timings(timings);
            std::cout << "Unacounted time: " << frame_time - measured << "ms \n\n";
            export_particles(particles, frame, outdir);
        }
    } catch (const std::exception& e) {
        std::cerr << "snow_sim: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
