// mpm_io.cpp -- the reference's on-disk formats, host side.
//
//  groups.csv     group_id,vx,vy,vz,mass        (src/import.cpp:37-76)
//  particles.csv  group_id,x,y,z                (src/import.cpp:78-108)
//  frame_%06d.csv header "x,y,z", rows x,-z,y   (src/export.cpp:10-22)
//
// Error behaviour follows import.cpp: the header line is skipped unchecked,
// empty lines are skipped, a trailing '\r' is stripped from each field, and a
// missing or empty field, a duplicate group id or a particle naming an
// undefined group is an error carrying the same message text.  The exporter
// writes with the same default stream formatting (6 significant digits) but,
// unlike the reference, creates the directory and reports failure.
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/snowmpm.h"

namespace snowmpm {
void set_error(const std::string& s);
}

namespace {

struct Group {
    double v[3];
    double m;
};

// read_field, import.cpp:11-35
std::string read_field(std::stringstream& row, const std::string& column) {
    std::string field;
    if (!std::getline(row, field, ',')) throw std::runtime_error("Missing CSV field: " + column);
    if (!field.empty() && field.back() == '\r') field.pop_back();
    if (field.empty()) throw std::runtime_error("Empty CSV field: " + column);
    return field;
}

// import_groups, import.cpp:37-76
std::map<int, Group> import_groups(const std::string& filename) {
    std::map<int, Group> groups;
    std::ifstream file{filename};
    if (!file.is_open()) throw std::runtime_error("File not open: " + filename);
    std::string line;
    if (!std::getline(file, line)) throw std::runtime_error("Group data file is empty: " + filename);
    while (std::getline(file, line)) {
        if (line.empty()) continue;
        std::stringstream row(line);
        const int id = std::stoi(read_field(row, "group_ID"));
        Group g;
        g.v[0] = std::stod(read_field(row, "v_x"));
        g.v[1] = std::stod(read_field(row, "v_y"));
        g.v[2] = std::stod(read_field(row, "v_z"));
        g.m = std::stod(read_field(row, "m"));
        if (!groups.emplace(id, g).second) throw std::runtime_error("Duplicate group id: " + std::to_string(id));
    }
    return groups;
}

struct Seed {
    int group;
    double x[3];
};

// import_particle_seeds, import.cpp:78-108
std::vector<Seed> import_seeds(const std::string& filename) {
    std::vector<Seed> out;
    std::ifstream file{filename};
    if (!file.is_open()) throw std::runtime_error("File not open: " + filename);
    std::string line;
    if (!std::getline(file, line)) throw std::runtime_error("Particle data file is empty: " + filename);
    while (std::getline(file, line)) {
        if (line.empty()) continue;
        std::stringstream row(line);
        Seed s;
        s.group = std::stoi(read_field(row, "group_ID"));
        s.x[0] = std::stod(read_field(row, "x"));
        s.x[1] = std::stod(read_field(row, "y"));
        s.x[2] = std::stod(read_field(row, "z"));
        out.push_back(s);
    }
    return out;
}

}  // namespace

extern "C" {

// import_particles + assemble_particles, import.cpp:110-137 (V_p0 = 0, F = I).
int mpm_import_particles(const char* groups_csv, const char* particles_csv, size_t* n, double** x, double** v,
                         double** m) {
    if (!groups_csv || !particles_csv || !n || !x || !v || !m) {
        snowmpm::set_error("mpm_import_particles: null argument");
        return MPM_ERR_ARG;
    }
    *x = *v = *m = nullptr;
    *n = 0;
    try {
        const auto groups = import_groups(groups_csv);
        const auto seeds = import_seeds(particles_csv);
        const size_t k = seeds.size();
        double* X = (double*)std::malloc(sizeof(double) * 3 * (k ? k : 1));
        double* V = (double*)std::malloc(sizeof(double) * 3 * (k ? k : 1));
        double* M = (double*)std::malloc(sizeof(double) * (k ? k : 1));
        if (!X || !V || !M) {
            std::free(X);
            std::free(V);
            std::free(M);
            snowmpm::set_error("out of host memory");
            return MPM_ERR_IO;
        }
        for (size_t i = 0; i < k; ++i) {
            auto it = groups.find(seeds[i].group);
            if (it == groups.end()) {
                std::free(X);
                std::free(V);
                std::free(M);
                throw std::runtime_error("Particle references undefined group id: " + std::to_string(seeds[i].group));
            }
            for (int a = 0; a < 3; ++a) {
                X[i * 3 + a] = seeds[i].x[a];
                V[i * 3 + a] = it->second.v[a];
            }
            M[i] = it->second.m;
        }
        *x = X;
        *v = V;
        *m = M;
        *n = k;
        return MPM_OK;
    } catch (const std::exception& e) {
        snowmpm::set_error(e.what());
        return MPM_ERR_IO;
    }
}

// export_particles, export.cpp:10-22 (solver axes -> Blender axes: x, -z, y).
int mpm_export_frame(const char* directory, int frame_no, size_t n, const double* x) {
    if (!directory || (n && !x)) {
        snowmpm::set_error("mpm_export_frame: null argument");
        return MPM_ERR_ARG;
    }
    if (::mkdir(directory, 0777) != 0 && errno != EEXIST) {
        snowmpm::set_error(std::string("cannot create directory ") + directory + ": " + std::strerror(errno));
        return MPM_ERR_IO;
    }
    std::ostringstream name;
    name << directory << "/frame_" << std::setw(6) << std::setfill('0') << frame_no << ".csv";
    std::ofstream file{name.str()};
    if (!file.is_open()) {
        snowmpm::set_error("cannot open " + name.str());
        return MPM_ERR_IO;
    }
    file << "x,y,z\n";
    for (size_t i = 0; i < n; ++i) file << x[i * 3 + 0] << "," << x[i * 3 + 2] * -1 << "," << x[i * 3 + 1] << "\n";
    file.flush();
    if (!file) {
        snowmpm::set_error("write failed: " + name.str());
        return MPM_ERR_IO;
    }
    return MPM_OK;
}

}  // extern "C"
