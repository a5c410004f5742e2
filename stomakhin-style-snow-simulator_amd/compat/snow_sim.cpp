// snow_sim.cpp -- command-line driver of the MI355X engine through the C++
// drop-in API (snow_compat.hpp): reads a particle/group input in the
// reference's CSV format, computes V_p0, runs frames of substeps and writes
// one frame CSV per frame.  Every parameter comes from flags or a scene file;
// the defaults are the scene the reference's driver hard-codes
// (src/main.cpp:14-37: dx 0.05, dt 5e-4, E 2.5e5, nu 0.2, xi 10,
// theta_c 0.015, theta_s 0.0075, alpha 0.95, g = (0, -10, 0), ground plane
// y = 0 with friction 1), and the API calls keep the order that reference
// driver uses (import, setup, frame 0, substeps, frame).
//
//   snow_sim [--scene FILE] [--groups F] [--particles F] [--frames N]
//            [--substeps N] [--out DIR] [--dx H] [--dt T] [--E E] [--nu NU]
//            [--xi XI] [--theta-c C] [--theta-s S] [--alpha A] [--gravity X,Y,Z]
//            [--plane PX,PY,PZ,NX,NY,NZ,MU[,VX,VY,VZ]] ... [--quiet]
// A scene file holds the same settings as "key value" lines ('#' comments);
// flags given after --scene override it.  Any --plane replaces the default
// ground plane; several planes are applied in the order given; VX,VY,VZ give
// a plane a constant velocity (a conveyor surface).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "snow_compat.hpp"

namespace {

struct Scene {
    std::string groups = "input/groups.csv", particles = "input/particles.csv", out = "output";
    int frames = 50, substeps = 80;
    double dx = 0.05, dt = 5e-4, E = 2.5e5, nu = 0.2, xi = 10.0, theta_c = 0.015, theta_s = 0.0075, alpha = 0.95;
    double gravity[3] = {0.0, -10.0, 0.0};
    std::vector<std::vector<double>> planes;  // px py pz nx ny nz mu [vx vy vz]
    bool quiet = false;
};

std::vector<double> numbers(const std::string& s, size_t want, size_t or_want, const std::string& key) {
    std::vector<double> v;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ',')) v.push_back(std::stod(tok));
    if (v.size() != want && v.size() != or_want)
        throw std::runtime_error(key + " takes " + std::to_string(want) + " comma-separated numbers" +
                                 (or_want != want ? " (or " + std::to_string(or_want) + ")" : std::string()));
    return v;
}

void set(Scene& sc, const std::string& key, const std::string& val) {
    if (key == "groups") sc.groups = val;
    else if (key == "particles") sc.particles = val;
    else if (key == "out") sc.out = val;
    else if (key == "frames") sc.frames = std::stoi(val);
    else if (key == "substeps") sc.substeps = std::stoi(val);
    else if (key == "dx") sc.dx = std::stod(val);
    else if (key == "dt") sc.dt = std::stod(val);
    else if (key == "E") sc.E = std::stod(val);
    else if (key == "nu") sc.nu = std::stod(val);
    else if (key == "xi") sc.xi = std::stod(val);
    else if (key == "theta-c") sc.theta_c = std::stod(val);
    else if (key == "theta-s") sc.theta_s = std::stod(val);
    else if (key == "alpha") sc.alpha = std::stod(val);
    else if (key == "gravity") {
        const auto g = numbers(val, 3, 3, key);
        for (int a = 0; a < 3; ++a) sc.gravity[a] = g[a];
    } else if (key == "plane") sc.planes.push_back(numbers(val, 7, 10, key));
    else throw std::runtime_error("unknown setting '" + key + "'");
}

void read_scene_file(Scene& sc, const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open scene file " + path);
    std::string line;
    while (std::getline(f, line)) {
        const auto hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::stringstream ss(line);
        std::string key, val;
        if (!(ss >> key)) continue;
        if (!(ss >> val)) throw std::runtime_error("scene file: '" + key + "' has no value");
        set(sc, key, val);
    }
}

Scene parse(int argc, char** argv) {
    Scene sc;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--quiet") {
            sc.quiet = true;
            continue;
        }
        if (a.rfind("--", 0) != 0 || i + 1 >= argc) throw std::runtime_error("bad argument '" + a + "'");
        const std::string key = a.substr(2), val = argv[++i];
        if (key == "scene")
            read_scene_file(sc, val);
        else
            set(sc, key, val);
    }
    if (sc.planes.empty()) sc.planes.push_back({0, 0, 0, 0, 1, 0, 1.0});
    return sc;
}

}  // namespace

int main(int argc, char** argv) {
    using namespace snow_compat;
    try {
        const Scene sc = parse(argc, argv);
        // Lame parameters from Young's modulus and Poisson's ratio
        const double mu0 = sc.E / (2.0 * (1.0 + sc.nu));
        const double lambda0 = sc.E * sc.nu / ((1.0 + sc.nu) * (1.0 - 2.0 * sc.nu));
        std::vector<Particle> snow = import_particles(sc.groups, sc.particles);
        std::cout << "loaded " << snow.size() << " particles from " << sc.particles << "\n";
        std::vector<CollisionBody> bodies;
        for (const auto& p : sc.planes) {
            bodies.push_back(make_plane(Vec3(p[0], p[1], p[2]), Vec3(p[3], p[4], p[5]), p[6]));
            if (p.size() == 10) set_velocity(bodies.back(), Vec3(p[7], p[8], p[9]));
        }
        const Vec3 g(sc.gravity[0], sc.gravity[1], sc.gravity[2]);
        std::vector<ParticleStencil> scratch(snow.size());
        setup(snow, sc.dx);
        export_particles(snow, 0, sc.out.c_str());
        double total_ms = 0.0;
        for (int frame = 1; frame <= sc.frames; ++frame) {
            StepTimings stages;
            const auto t0 = std::chrono::steady_clock::now();
            step_substeps(snow, bodies, scratch, g, sc.dt, sc.dx, mu0, lambda0, sc.xi, sc.theta_c, sc.theta_s,
                          sc.alpha, stages, sc.substeps);
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            total_ms += ms;
            export_particles(snow, frame, sc.out.c_str());
            if (!sc.quiet) {
                std::printf("frame %d: %d substeps in %.3f ms (GPU stages: bin+sort %.3f, P2G %.3f, grid %.3f, "
                            "G2P %.3f ms)\n",
                            frame, sc.substeps, ms, stages.stencil, stages.p2g, stages.grid_velocity, stages.g2p);
            }
        }
        const double psteps = (double)snow.size() * sc.frames * sc.substeps;
        std::printf("%d frames, %.3f s, %.3g particle-steps/s including frame downloads\n", sc.frames,
                    total_ms * 1e-3, total_ms > 0 ? psteps / (total_ms * 1e-3) : 0.0);
    } catch (const std::exception& e) {
        std::cerr << "snow_sim: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
