// Seeded synthetic inputs for the sensor update: analytic scenes (solid
// axis-aligned boxes and vertical cylinders), TSDF maps in the reference's
// CudaSubVoxelMap layout (1 m blocks in linear block order, voxels x-fastest
// inside a block; cuda_sub_voxel_map.tcc:171-231), LiDAR scans ray-cast
// against the same primitives, and particle sets.
//
// Host-only C++/OpenMP.  Determinism: a counter-based generator (splitmix64)
// seeded per item, so results do not depend on the thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/tsl.h"

#define SYNTH_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

struct Box {
    double lo[3], hi[3];
};
struct Cyl {
    double cx, cy, z0, z1, r;
};

struct Scene {
    std::vector<Box> boxes;
    std::vector<Cyl> cyls;
};

inline uint64_t splitmix(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(splitmix(seed ^ 0x5eed5eed5eed5eedull)) {}
    uint64_t next()
    {
        s += 0x9e3779b97f4a7c15ull;
        uint64_t x = s;
        x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
        x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
        return x ^ (x >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }            // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal()
    {
        double u1 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

double sd_box(const Box& b, double x, double y, double z)
{
    const double c[3] = {(b.lo[0] + b.hi[0]) * 0.5, (b.lo[1] + b.hi[1]) * 0.5, (b.lo[2] + b.hi[2]) * 0.5};
    const double h[3] = {(b.hi[0] - b.lo[0]) * 0.5, (b.hi[1] - b.lo[1]) * 0.5, (b.hi[2] - b.lo[2]) * 0.5};
    const double q[3] = {std::fabs(x - c[0]) - h[0], std::fabs(y - c[1]) - h[1], std::fabs(z - c[2]) - h[2]};
    const double ox = std::max(q[0], 0.0), oy = std::max(q[1], 0.0), oz = std::max(q[2], 0.0);
    return std::sqrt(ox * ox + oy * oy + oz * oz) + std::min(std::max(q[0], std::max(q[1], q[2])), 0.0);
}

double sd_cyl(const Cyl& c, double x, double y, double z)
{
    const double dr = std::hypot(x - c.cx, y - c.cy) - c.r;
    const double zc = (c.z0 + c.z1) * 0.5, hz = (c.z1 - c.z0) * 0.5;
    const double dz = std::fabs(z - zc) - hz;
    const double ox = std::max(dr, 0.0), oz = std::max(dz, 0.0);
    return std::sqrt(ox * ox + oz * oz) + std::min(std::max(dr, dz), 0.0);
}

// distance from a point to an axis-aligned region, used to cull primitives
double aabb_gap(const double lo[3], const double hi[3], const double qlo[3], const double qhi[3])
{
    double d2 = 0.0;
    for (int a = 0; a < 3; ++a) {
        double g = 0.0;
        if (qhi[a] < lo[a]) g = lo[a] - qhi[a];
        else if (qlo[a] > hi[a]) g = qlo[a] - hi[a];
        d2 += g * g;
    }
    return std::sqrt(d2);
}

struct Cand {
    std::vector<const Box*> b;
    std::vector<const Cyl*> c;
    double sdf(double x, double y, double z) const
    {
        double d = 1e30;
        for (const Box* p : b) d = std::min(d, sd_box(*p, x, y, z));
        for (const Cyl* p : c) d = std::min(d, sd_cyl(*p, x, y, z));
        return d;
    }
};

Cand candidates(const Scene& s, const double qlo[3], const double qhi[3], double margin)
{
    Cand out;
    for (const Box& b : s.boxes)
        if (aabb_gap(b.lo, b.hi, qlo, qhi) < margin) out.b.push_back(&b);
    for (const Cyl& c : s.cyls) {
        const double lo[3] = {c.cx - c.r, c.cy - c.r, c.z0}, hi[3] = {c.cx + c.r, c.cy + c.r, c.z1};
        if (aabb_gap(lo, hi, qlo, qhi) < margin) out.c.push_back(&c);
    }
    return out;
}

double scene_sdf(const Scene& s, double x, double y, double z)
{
    double d = 1e30;
    for (const Box& b : s.boxes) d = std::min(d, sd_box(b, x, y, z));
    for (const Cyl& c : s.cyls) d = std::min(d, sd_cyl(c, x, y, z));
    return d;
}

// ray / box slab test: entry distance in [tmin, tmax] or +inf
double ray_box(const Box& b, const double o[3], const double d[3], double tmin, double tmax)
{
    double t0 = tmin, t1 = tmax;
    for (int a = 0; a < 3; ++a) {
        if (std::fabs(d[a]) < 1e-15) {
            if (o[a] < b.lo[a] || o[a] > b.hi[a]) return INFINITY;
            continue;
        }
        double ta = (b.lo[a] - o[a]) / d[a], tb = (b.hi[a] - o[a]) / d[a];
        if (ta > tb) std::swap(ta, tb);
        t0 = std::max(t0, ta);
        t1 = std::min(t1, tb);
        if (t0 > t1) return INFINITY;
    }
    return t0;
}

double ray_cyl(const Cyl& c, const double o[3], const double d[3], double tmin, double tmax)
{
    double best = INFINITY;
    const double ox = o[0] - c.cx, oy = o[1] - c.cy;
    const double a = d[0] * d[0] + d[1] * d[1];
    if (a > 1e-15) {
        const double b = 2.0 * (ox * d[0] + oy * d[1]);
        const double cc = ox * ox + oy * oy - c.r * c.r;
        const double disc = b * b - 4.0 * a * cc;
        if (disc >= 0.0) {
            const double sq = std::sqrt(disc);
            for (double t : {(-b - sq) / (2.0 * a), (-b + sq) / (2.0 * a)}) {
                if (t < tmin || t > tmax) continue;
                const double z = o[2] + t * d[2];
                if (z >= c.z0 && z <= c.z1) {
                    best = std::min(best, t);
                    break;
                }
            }
        }
    }
    if (std::fabs(d[2]) > 1e-15) {
        for (double zc : {c.z0, c.z1}) {
            const double t = (zc - o[2]) / d[2];
            if (t < tmin || t > tmax) continue;
            const double x = ox + t * d[0], y = oy + t * d[1];
            if (x * x + y * y <= c.r * c.r) best = std::min(best, t);
        }
    }
    return best;
}

// rotation R = Rz(yaw) Ry(pitch) Rx(roll), row-major
void rpy_matrix(double roll, double pitch, double yaw, double R[9])
{
    const double sa = std::sin(roll), ca = std::cos(roll), sb = std::sin(pitch), cb = std::cos(pitch),
                 sg = std::sin(yaw), cg = std::cos(yaw);
    R[0] = cb * cg; R[1] = sa * sb * cg - ca * sg; R[2] = ca * sb * cg + sa * sg;
    R[3] = cb * sg; R[4] = sa * sb * sg + ca * cg; R[5] = ca * sb * sg - sa * cg;
    R[6] = -sb;     R[7] = sa * cb;                R[8] = ca * cb;
}

void add_random_objects(Scene& s, Rng& rng, int n, double x0, double y0, double x1, double y1, double z_floor,
                        double z_top, double clear_x, double clear_y, double clear_r)
{
    int placed = 0, tries = 0;
    while (placed < n && tries < 100 * n + 100) {
        ++tries;
        const bool cyl = rng.uniform() < 0.4;
        const double cx = rng.uniform(x0 + 0.6, x1 - 0.6), cy = rng.uniform(y0 + 0.6, y1 - 0.6);
        if (std::hypot(cx - clear_x, cy - clear_y) < clear_r) continue;
        if (cyl) {
            const double r = rng.uniform(0.1, 0.5);
            const double h = rng.uniform(0.5, std::min(2.6, z_top - z_floor));
            s.cyls.push_back(Cyl{cx, cy, z_floor - 0.1, z_floor + h, r});
        } else {
            const double hx = rng.uniform(0.15, 0.75), hy = rng.uniform(0.15, 0.75);
            const double h = rng.uniform(0.5, std::min(2.0, z_top - z_floor));
            s.boxes.push_back(Box{{cx - hx, cy - hy, z_floor - 0.1}, {cx + hx, cy + hy, z_floor + h}});
        }
        ++placed;
    }
}

}  // namespace

struct synth_scene {
    Scene s;
};

SYNTH_EXPORT synth_scene* synth_scene_new(void) { return new synth_scene(); }
SYNTH_EXPORT void synth_scene_free(synth_scene* s) { delete s; }
SYNTH_EXPORT void synth_free(void* p) { std::free(p); }

SYNTH_EXPORT int synth_scene_add_box(synth_scene* sc, double x0, double y0, double z0, double x1, double y1, double z1)
{
    sc->s.boxes.push_back(Box{{x0, y0, z0}, {x1, y1, z1}});
    return (int)(sc->s.boxes.size() + sc->s.cyls.size());
}

SYNTH_EXPORT int synth_scene_add_cylinder(synth_scene* sc, double cx, double cy, double z0, double z1, double r)
{
    sc->s.cyls.push_back(Cyl{cx, cy, z0, z1, r});
    return (int)(sc->s.boxes.size() + sc->s.cyls.size());
}

SYNTH_EXPORT int synth_scene_count(const synth_scene* sc) { return (int)(sc->s.boxes.size() + sc->s.cyls.size()); }

// One room spanning [0, sx] x [0, sy] x [0, sz]: floor top at 0.3, ceiling
// underside at sz - 0.3, walls with inner faces 0.4 m from the bounds, and
// n_objects random boxes/cylinders kept clear_r away from (clear_x, clear_y).
SYNTH_EXPORT void synth_scene_build_room(synth_scene* sc, uint64_t seed, double sx, double sy, double sz,
                                         int n_objects, double clear_x, double clear_y, double clear_r)
{
    Scene& s = sc->s;
    const double t = 1.0;
    s.boxes.push_back(Box{{-t, -t, -t}, {sx + t, sy + t, 0.3}});                // floor
    s.boxes.push_back(Box{{-t, -t, sz - 0.3}, {sx + t, sy + t, sz + t}});       // ceiling
    s.boxes.push_back(Box{{-t, -t, -t}, {0.4, sy + t, sz + t}});                // walls
    s.boxes.push_back(Box{{sx - 0.4, -t, -t}, {sx + t, sy + t, sz + t}});
    s.boxes.push_back(Box{{-t, -t, -t}, {sx + t, 0.4, sz + t}});
    s.boxes.push_back(Box{{-t, sy - 0.4, -t}, {sx + t, sy + t, sz + t}});
    Rng rng(seed);
    add_random_objects(s, rng, n_objects, 0.4, 0.4, sx - 0.4, sy - 0.4, 0.3, sz - 0.3, clear_x, clear_y, clear_r);
}

// A one-storey building over [0, sx] x [0, sy]: floor slab (top at 0.3),
// outer walls, interior walls on a room grid with door gaps, a ceiling slab
// (underside at wall_h) over a random subset of rooms, random furniture.
SYNTH_EXPORT void synth_scene_build_building(synth_scene* sc, uint64_t seed, double sx, double sy, double wall_h,
                                             double room, int objects_per_room, double ceiling_fraction)
{
    Scene& s = sc->s;
    Rng rng(seed);
    const double t = 1.0, w = 0.1;  // outer slab overhang, half wall thickness
    s.boxes.push_back(Box{{-t, -t, -t}, {sx + t, sy + t, 0.3}});
    s.boxes.push_back(Box{{-t, -t, -t}, {0.4, sy + t, wall_h}});
    s.boxes.push_back(Box{{sx - 0.4, -t, -t}, {sx + t, sy + t, wall_h}});
    s.boxes.push_back(Box{{-t, -t, -t}, {sx + t, 0.4, wall_h}});
    s.boxes.push_back(Box{{-t, sy - 0.4, -t}, {sx + t, sy + t, wall_h}});
    const int nx = std::max(1, (int)std::floor((sx - 0.8) / room + 0.5));
    const int ny = std::max(1, (int)std::floor((sy - 0.8) / room + 0.5));
    const double rx = (sx - 0.8) / nx, ry = (sy - 0.8) / ny;
    // interior walls along x (constant x) with one door per room span
    for (int i = 1; i < nx; ++i) {
        const double x = 0.4 + i * rx;
        for (int j = 0; j < ny; ++j) {
            const double y0 = 0.4 + j * ry, y1 = y0 + ry;
            const double door = rng.uniform(y0 + 0.8, y1 - 1.8);
            s.boxes.push_back(Box{{x - w, y0, 0.0}, {x + w, door, wall_h}});
            s.boxes.push_back(Box{{x - w, door + 1.0, 0.0}, {x + w, y1, wall_h}});
        }
    }
    for (int j = 1; j < ny; ++j) {
        const double y = 0.4 + j * ry;
        for (int i = 0; i < nx; ++i) {
            const double x0 = 0.4 + i * rx, x1 = x0 + rx;
            const double door = rng.uniform(x0 + 0.8, x1 - 1.8);
            s.boxes.push_back(Box{{x0, y - w, 0.0}, {door, y + w, wall_h}});
            s.boxes.push_back(Box{{door + 1.0, y - w, 0.0}, {x1, y + w, wall_h}});
        }
    }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            const double x0 = 0.4 + i * rx, y0 = 0.4 + j * ry;
            if (rng.uniform() < ceiling_fraction)
                s.boxes.push_back(Box{{x0 - w, y0 - w, wall_h}, {x0 + rx + w, y0 + ry + w, wall_h + 0.2}});
            add_random_objects(s, rng, objects_per_room, x0 + w, y0 + w, x0 + rx - w, y0 + ry - w, 0.3,
                               wall_h, x0 + rx * 0.5, y0 + ry * 0.5, 1.2);
        }
}

SYNTH_EXPORT double synth_sdf(const synth_scene* sc, double x, double y, double z)
{
    return scene_sdf(sc->s, x, y, z);
}

SYNTH_EXPORT void synth_sdf_batch(const synth_scene* sc, const double* xyz, size_t n, double* out)
{
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; ++i) out[i] = scene_sdf(sc->s, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

// TSDF map in the reference layout.  coef must come from the
// CudaSubVoxelMap constructor arithmetic (tsl_map_coef_init); on return
// coef->data_size is set, grid_occ (grid_occ_size entries) holds voxel
// offsets in linear block order (-1 = unallocated), and *tsdf_out is a
// malloc'ed int16 array (data_size entries) of signed distances in mm at
// voxel centres, INT16_MIN where |d| >= truncation (such voxels read the
// map's init value, like cells createTSDFMap skips).  dense != 0 allocates
// every block; otherwise only blocks that hold a voxel with |d| < truncation.
SYNTH_EXPORT int synth_map_build(const synth_scene* sc, tsl_map_coef* coef, int truncation_mm, int dense,
                                 int32_t* grid_occ, int16_t** tsdf_out)
{
    const Scene& s = sc->s;
    const uint64_t ux = coef->up_dim_x, uy = coef->up_dim_y, uz = coef->up_dim_z, sd = coef->sub_dim;
    const double res = (double)coef->resolution;
    const double trunc = truncation_mm * 0.001;
    const uint64_t nblocks = ux * uy * uz;
    const uint64_t vox = sd * sd * sd;
    std::vector<uint8_t> alloc(nblocks, 0);
    auto block_lo = [&](uint64_t b, double lo[3], double hi[3]) {
        const uint64_t bx = b % ux, by = (b / ux) % uy, bz = b / (ux * uy);
        lo[0] = coef->min_x + (double)bx; lo[1] = coef->min_y + (double)by; lo[2] = coef->min_z + (double)bz;
        hi[0] = lo[0] + 1.0; hi[1] = lo[1] + 1.0; hi[2] = lo[2] + 1.0;
    };
    #pragma omp parallel for schedule(dynamic, 4)
    for (long b = 0; b < (long)nblocks; ++b) {
        if (dense) {
            alloc[b] = 1;
            continue;
        }
        double lo[3], hi[3];
        block_lo((uint64_t)b, lo, hi);
        const Cand cand = candidates(s, lo, hi, trunc + 1e-6);
        if (cand.b.empty() && cand.c.empty()) continue;
        bool any = false;
        for (uint64_t v = 0; v < vox && !any; ++v) {
            const uint64_t vx = v % sd, vy = (v / sd) % sd, vz = v / (sd * sd);
            const double x = lo[0] + (vx + 0.5) * res, y = lo[1] + (vy + 0.5) * res, z = lo[2] + (vz + 0.5) * res;
            const double d = cand.sdf(x, y, z);
            if (std::fabs(d) < trunc && std::lround(d * 1000.0) > -truncation_mm &&
                std::lround(d * 1000.0) < truncation_mm)
                any = true;
        }
        alloc[b] = any ? 1 : 0;
    }
    int64_t offset = 0;
    for (uint64_t b = 0; b < nblocks; ++b) {
        if (alloc[b]) {
            if (offset > 0x7fffffff - (int64_t)vox) return -1;
            grid_occ[b] = (int32_t)offset;
            offset += (int64_t)vox;
        } else {
            grid_occ[b] = -1;
        }
    }
    coef->data_size = (uint64_t)offset;
    int16_t* tsdf = (int16_t*)std::malloc(sizeof(int16_t) * (offset ? offset : 1));
    if (!tsdf) return -2;
    #pragma omp parallel for schedule(dynamic, 4)
    for (long b = 0; b < (long)nblocks; ++b) {
        if (grid_occ[b] < 0) continue;
        double lo[3], hi[3];
        block_lo((uint64_t)b, lo, hi);
        const Cand cand = candidates(s, lo, hi, trunc + 1e-6);
        int16_t* out = tsdf + grid_occ[b];
        for (uint64_t v = 0; v < vox; ++v) {
            const uint64_t vx = v % sd, vy = (v / sd) % sd, vz = v / (sd * sd);
            const double x = lo[0] + (vx + 0.5) * res, y = lo[1] + (vy + 0.5) * res, z = lo[2] + (vz + 0.5) * res;
            const double d = (cand.b.empty() && cand.c.empty()) ? 1e30 : cand.sdf(x, y, z);
            const long mm = std::fabs(d) < 30.0 ? std::lround(d * 1000.0) : 30000;
            out[v] = (mm > -truncation_mm && mm < truncation_mm) ? (int16_t)mm : (int16_t)INT16_MIN;
        }
    }
    *tsdf_out = tsdf;
    return 0;
}

// Ray-cast LiDAR scan.  sensor_pose = (x, y, z, roll, pitch, yaw) of the
// scanner in the map frame.  beams rows evenly spaced in elevation over
// [el_min, el_max] degrees, azimuths columns over 360 degrees.  Returns hits
// (range in (0, max_range]) in the scanner frame with Gaussian range noise,
// ring (beam) index, in ring-major order.  points_out needs 3*beams*azimuths
// floats, rings_out beams*azimuths ints.
SYNTH_EXPORT size_t synth_raycast(const synth_scene* sc, const double pose[6], int beams, int azimuths,
                                  double el_min, double el_max, double max_range, double noise_sigma, uint64_t seed,
                                  float* points_out, int32_t* rings_out)
{
    const Scene& s = sc->s;
    double R[9];
    rpy_matrix(pose[3], pose[4], pose[5], R);
    const double o[3] = {pose[0], pose[1], pose[2]};
    const long total = (long)beams * azimuths;
    std::vector<double> range(total, INFINITY);
    std::vector<double> dirs(3 * total);
    #pragma omp parallel for schedule(dynamic, 64)
    for (long r = 0; r < total; ++r) {
        const int b = (int)(r / azimuths), a = (int)(r % azimuths);
        const double el = (beams > 1 ? el_min + (el_max - el_min) * b / (beams - 1) : el_min) * M_PI / 180.0;
        const double az = 2.0 * M_PI * a / azimuths;
        const double ds[3] = {std::cos(el) * std::cos(az), std::cos(el) * std::sin(az), std::sin(el)};
        const double dw[3] = {R[0] * ds[0] + R[1] * ds[1] + R[2] * ds[2], R[3] * ds[0] + R[4] * ds[1] + R[5] * ds[2],
                              R[6] * ds[0] + R[7] * ds[1] + R[8] * ds[2]};
        double best = INFINITY;
        for (const Box& bx : s.boxes) best = std::min(best, ray_box(bx, o, dw, 1e-6, max_range));
        for (const Cyl& c : s.cyls) best = std::min(best, ray_cyl(c, o, dw, 1e-6, max_range));
        if (best <= max_range) {
            Rng rng(seed * 0x100000001b3ull + (uint64_t)r);
            best += noise_sigma * rng.normal();
        }
        range[r] = best;
        dirs[3 * r] = ds[0]; dirs[3 * r + 1] = ds[1]; dirs[3 * r + 2] = ds[2];
    }
    size_t n = 0;
    for (long r = 0; r < total; ++r) {
        if (!(range[r] > 0.0 && range[r] <= max_range)) continue;
        points_out[3 * n] = (float)(dirs[3 * r] * range[r]);
        points_out[3 * n + 1] = (float)(dirs[3 * r + 1] * range[r]);
        points_out[3 * n + 2] = (float)(dirs[3 * r + 2] * range[r]);
        rings_out[n] = (int32_t)(r / azimuths);
        ++n;
    }
    return n;
}

// n particles (x, y, z, roll, pitch, yaw, weight = 1/n) drawn from
// independent normals around mean with the given sigmas.
SYNTH_EXPORT void synth_particles_normal(uint64_t seed, size_t n, const double mean[6], const double sigma[6],
                                         float* out)
{
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; ++i) {
        Rng rng(seed * 0x9e3779b97f4a7c15ull + (uint64_t)i);
        for (int k = 0; k < 6; ++k) out[7 * i + k] = (float)(mean[k] + sigma[k] * rng.normal());
        out[7 * i + 6] = (float)(1.0 / (double)n);
    }
}

// n particles uniform over free space: position uniform in [lo, hi] with
// scene distance >= clearance, yaw uniform in [-pi, pi), roll and pitch
// uniform in [-rp, rp] radians.  Returns the number written (n unless the
// free space is too small to hit within 1000 tries per particle).
SYNTH_EXPORT size_t synth_particles_free(const synth_scene* sc, uint64_t seed, size_t n, const double lo[3],
                                         const double hi[3], double clearance, double rp, float* out)
{
    const Scene& s = sc->s;
    size_t bad = 0;
    #pragma omp parallel for schedule(dynamic, 256) reduction(+ : bad)
    for (long i = 0; i < (long)n; ++i) {
        Rng rng(seed * 0xd1342543de82ef95ull + (uint64_t)i);
        bool ok = false;
        double p[3] = {0, 0, 0};
        for (int t = 0; t < 1000 && !ok; ++t) {
            for (int a = 0; a < 3; ++a) p[a] = rng.uniform(lo[a], hi[a]);
            ok = scene_sdf(s, p[0], p[1], p[2]) >= clearance;
        }
        if (!ok) ++bad;
        out[7 * i + 0] = (float)p[0];
        out[7 * i + 1] = (float)p[1];
        out[7 * i + 2] = (float)p[2];
        out[7 * i + 3] = (float)rng.uniform(-rp, rp);
        out[7 * i + 4] = (float)rng.uniform(-rp, rp);
        out[7 * i + 5] = (float)rng.uniform(-M_PI, M_PI);
        out[7 * i + 6] = (float)(1.0 / (double)n);
    }
    return n - bad;
}

// Voxel-grid filter of a scan: keep the first point (in input order) of
// every (cell) at the given cell size, dropping points closer than
// min_range; output keeps input order.  Returns the number kept.
SYNTH_EXPORT size_t synth_voxel_filter(const float* pts, const int32_t* rings, size_t n, double cell, double min_range,
                                       float* out_pts, int32_t* out_rings)
{
    std::vector<std::pair<uint64_t, uint32_t>> keys;
    keys.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        const double x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
        if (std::sqrt(x * x + y * y + z * z) < min_range) continue;
        const int64_t cx = (int64_t)std::floor(x / cell) + (1 << 20), cy = (int64_t)std::floor(y / cell) + (1 << 20),
                      cz = (int64_t)std::floor(z / cell) + (1 << 20);
        const uint64_t key = ((uint64_t)cx << 42) ^ ((uint64_t)cy << 21) ^ (uint64_t)cz;
        keys.emplace_back(key, (uint32_t)i);
    }
    std::stable_sort(keys.begin(), keys.end(),
                     [](const std::pair<uint64_t, uint32_t>& a, const std::pair<uint64_t, uint32_t>& b) {
                         return a.first < b.first;
                     });
    std::vector<uint32_t> keep;
    for (size_t k = 0; k < keys.size(); ++k)
        if (k == 0 || keys[k].first != keys[k - 1].first) keep.push_back(keys[k].second);
    std::sort(keep.begin(), keep.end());
    for (size_t k = 0; k < keep.size(); ++k) {
        const uint32_t i = keep[k];
        out_pts[3 * k] = pts[3 * i];
        out_pts[3 * k + 1] = pts[3 * i + 1];
        out_pts[3 * k + 2] = pts[3 * i + 2];
        if (out_rings) out_rings[k] = rings ? rings[i] : 0;
    }
    return keep.size();
}
