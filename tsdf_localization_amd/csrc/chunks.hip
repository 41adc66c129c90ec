// TSDF map chunks -> the cell list of CudaSubVoxelMap::setData, as
// createTSDFMap does after reading the /map group of a map file
// (include/tsdf_localization/map/map_util.h:17-154).  The file read itself
// (HighFive over libhdf5) is the caller's: it hands over each chunk's
// position and its CHUNK_SIZE^3 raw TSDFValue words.  Host only.
#include "tsl_internal.h"

#include <cmath>
#include <cstdlib>

using namespace tsl;

namespace {

constexpr int kChunkSize = 64;           // CHUNK_SIZE = 1 << CHUNK_SHIFT, grid_map.h:18-19
constexpr int kMapResolutionMm = 64;     // MAP_RESOLUTION = 1 << MAP_SHIFT, grid_map.h:21-22
constexpr float kTruncationMm = 600.0f;  // truncation, grid_map.h:24
constexpr size_t kChunkVoxels = (size_t)kChunkSize * kChunkSize * kChunkSize;

// static_cast<FLOAT_T>(x) * MAP_RESOLUTION * 0.001 (map_util.h:114-116): the
// product with the int constant is float, the one with 0.001 double, and the
// result lands in the float cell tuple / CudaPoint.
inline float cell_coord(int x)
{
    return (float)((double)((float)x * (float)kMapResolutionMm) * 0.001);
}

// TSDFValue (util/tsdf.h:11-45): a union of the 32-bit raw word with
// {int16 value; int16 weight}, value first, so on a little-endian host the
// value is the low half and the weight the high half.
inline int16_t tsdf_value(uint32_t raw) { return (int16_t)(uint16_t)(raw & 0xffffu); }
inline int16_t tsdf_weight(uint32_t raw) { return (int16_t)(uint16_t)(raw >> 16); }

}  // namespace

extern "C" {

int tsl_map_chunks_to_cells(const int32_t* chunk_pos, const uint32_t* chunk_raw, size_t num_chunks, float sigma,
                            tsl_map_coef* coef, float* cell_xyz, float* cell_value, size_t cell_capacity,
                            size_t* num_cells, float* free_xyz, size_t free_capacity, size_t* num_free)
{
    if (!coef || !num_cells || !num_free || (num_chunks && (!chunk_pos || !chunk_raw)))
        return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    // Bounding box over the chunk positions, map_util.h:23-64.  min and max
    // start at 0, so the box always holds the origin.
    float mn[3] = {0.0f, 0.0f, 0.0f}, mx[3] = {0.0f, 0.0f, 0.0f};
    for (size_t c = 0; c < num_chunks; ++c) {
        for (int a = 0; a < 3; ++a) {
            const float v = (float)chunk_pos[3 * c + a];
            if (v < mn[a]) mn[a] = v;
            if (v > mx[a]) mx[a] = v;
        }
    }
    for (int a = 0; a < 3; ++a) {
        mn[a] = (float)((double)(mn[a] * (float)kChunkSize * (float)kMapResolutionMm) * 0.001);
        mx[a] = (float)((double)((mx[a] * (float)kChunkSize + (float)kChunkSize) * (float)kMapResolutionMm) * 0.001);
    }
    // MapT(min, max, MAP_RESOLUTION * 0.001, init), map_util.h:66-73: the
    // double resolution narrows to the map's float index type.
    TSL_TRY(tsl_map_coef_init(mn[0], mn[1], mn[2], mx[0], mx[1], mx[2], (float)(kMapResolutionMm * 0.001),
                              tsl_likelihood_init(sigma), coef));

    // Cells, map_util.h:78-148: chunks in the order given, then i (x), j (y),
    // k (z) with the chunk stored x-major.  weight != 0 and |tsdf| < truncation
    // gives a likelihood cell; weight != 0 and |tsdf| >= truncation a free-space
    // point; weight == 0 neither.
    size_t nc = 0, nf = 0;
    for (size_t c = 0; c < num_chunks; ++c) {
        const uint32_t* raw = chunk_raw + c * kChunkVoxels;
        const int bx = kChunkSize * chunk_pos[3 * c], by = kChunkSize * chunk_pos[3 * c + 1],
                  bz = kChunkSize * chunk_pos[3 * c + 2];
        for (int i = 0; i < kChunkSize; ++i) {
            for (int j = 0; j < kChunkSize; ++j) {
                for (int k = 0; k < kChunkSize; ++k) {
                    const uint32_t r = raw[(size_t)kChunkSize * kChunkSize * i + (size_t)kChunkSize * j + k];
                    const int16_t w = tsdf_weight(r);
                    if (w == 0) continue;
                    const int16_t v = tsdf_value(r);
                    const bool valid = std::fabs((float)v) < kTruncationMm;
                    float* dst;
                    if (valid) {
                        dst = (cell_xyz && nc < cell_capacity) ? cell_xyz + 3 * nc : nullptr;
                        if (dst && cell_value) cell_value[nc] = tsl_likelihood_value(v, sigma);
                        ++nc;
                    } else {
                        dst = (free_xyz && nf < free_capacity) ? free_xyz + 3 * nf : nullptr;
                        ++nf;
                    }
                    if (dst) {
                        dst[0] = cell_coord(bx + i);
                        dst[1] = cell_coord(by + j);
                        dst[2] = cell_coord(bz + k);
                    }
                }
            }
        }
    }
    *num_cells = nc;
    *num_free = nf;
    if (nc > cell_capacity || nf > free_capacity || (nc && (!cell_xyz || !cell_value)) || (nf && !free_xyz))
        return set_error(TSL_ERR_CAPACITY, "chunks hold %zu cells and %zu free points", nc, nf);
    return TSL_OK;
}

}  // extern "C"
