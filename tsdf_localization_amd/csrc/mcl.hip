// .mcl snapshot files (MCLFile, src/util/mcl_file.cpp:14-113): a text dump
// of one sensor update's inputs -- scan points with their ring numbers,
// particles with weights, the scanner->base transform and a pose -- written
// with the default stream formatting (6 significant digits) and read back
// with stream extraction.  Host only.
#include "tsl_internal.h"

#include <fstream>
#include <sstream>
#include <string>
#include <vector>

using namespace tsl;

extern "C" {

int tsl_mcl_write(const char* path, const tsl_point* points, const int32_t* rings, size_t num_points,
                  const tsl_particle* particles, size_t num_particles, const float tf[16], const float pose[7])
{
    if (!path || !tf || !pose || (num_points && (!points || !rings)) || (num_particles && !particles))
        return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    std::ofstream file(path);
    if (!file) return set_error(TSL_ERR_INVALID_ARGUMENT, "Error while opening file for writing");
    std::stringstream data;  // mcl_file.cpp:25-55
    data << num_points << "\n\n";
    for (size_t i = 0; i < num_points; ++i) data << points[i].x << ' ' << points[i].y << ' ' << points[i].z << '\n';
    for (size_t i = 0; i < num_points; ++i) data << rings[i] << '\n';
    data << num_particles << '\n';
    for (size_t i = 0; i < num_particles; ++i) {
        const tsl_particle& p = particles[i];
        data << p.first[0] << ' ' << p.first[1] << ' ' << p.first[2] << ' ' << p.first[3] << ' ' << p.first[4] << ' '
             << p.first[5] << "  " << p.second << '\n';
    }
    data << '\n';
    for (int k = 0; k < 16; ++k) data << tf[k] << ' ';
    data << '\n';
    data << pose[0] << " " << pose[1] << " " << pose[2] << " " << pose[3] << " " << pose[4] << " " << pose[5] << " "
         << pose[6];
    file << data.str();
    file.flush();
    if (!file) return set_error(TSL_ERR_INVALID_ARGUMENT, "Error while writing into file");
    return TSL_OK;
}

int tsl_mcl_read(const char* path, tsl_point* points, int32_t* rings, size_t points_capacity, tsl_particle* particles,
                 size_t particles_capacity, float tf[16], float pose[7], size_t* num_points, size_t* num_particles)
{
    if (!path || !num_points || !num_particles) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    std::ifstream file(path);
    if (!file) return set_error(TSL_ERR_INVALID_ARGUMENT, "Error while opening file for reading");
    // mcl_file.cpp:77-111
    size_t np = 0, nq = 0;
    file >> np;
    std::vector<tsl_point> pts(np);
    std::vector<int32_t> rg(np);
    for (size_t i = 0; i < np; ++i) file >> pts[i].x >> pts[i].y >> pts[i].z;
    for (size_t i = 0; i < np; ++i) file >> rg[i];
    file >> nq;
    std::vector<tsl_particle> ps(nq);
    for (size_t i = 0; i < nq; ++i)
        file >> ps[i].first[0] >> ps[i].first[1] >> ps[i].first[2] >> ps[i].first[3] >> ps[i].first[4] >>
            ps[i].first[5] >> ps[i].second;
    float t[16], q[7];
    for (int k = 0; k < 16; ++k) file >> t[k];
    for (int k = 0; k < 7; ++k) file >> q[k];
    if (!file) return set_error(TSL_ERR_INVALID_ARGUMENT, "Error: Could not read mcl data from file");
    *num_points = np;
    *num_particles = nq;
    if (np > points_capacity || nq > particles_capacity)
        return set_error(TSL_ERR_CAPACITY, "file holds %zu points and %zu particles", np, nq);
    if (np && (!points || !rings)) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL point buffers");
    if (nq && !particles) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL particle buffer");
    for (size_t i = 0; i < np; ++i) {
        points[i] = pts[i];
        rings[i] = rg[i];
    }
    for (size_t i = 0; i < nq; ++i) particles[i] = ps[i];
    if (tf)
        for (int k = 0; k < 16; ++k) tf[k] = t[k];
    if (pose)
        for (int k = 0; k < 7; ++k) pose[k] = q[k];
    return TSL_OK;
}

}  // extern "C"
