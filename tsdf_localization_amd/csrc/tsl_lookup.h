// Device-side voxel lookup with the reference's exact index semantics.
//
// Reference: CudaSubVoxelMap::getIndex / getEntry,
// include/tsdf_localization/cuda/cuda_sub_voxel_map.tcc:51-158 (the CPU
// evaluator's lookup).  The returned index is the reference's linear voxel
// index (block offset + sx + sy*sub_dim + sz*sub_dim^2); data_size means
// "rejected, read init_value".
#pragma once

#include "tsl_internal.h"

namespace tsl {

// General path: every float->size_t conversion with x86-64 semantics, size_t
// arithmetic modulo 2^64, IEEE float division.  Taken for offsets that are
// negative, >= 2^23 or NaN (points outside the map on the low side, or
// garbage); bit-exact with the reference for all inputs.
__device__ __attribute__((noinline)) uint64_t lookup_index_general(const DevMap& m, float ox,
                                                                     float oy, float oz)
{
    // :53-62 integer-metre bounds check
    uint64_t gx = cvt_f32_u64(ox), gy = cvt_f32_u64(oy), gz = cvt_f32_u64(oz);
    uint64_t qx = cvt_f32_u64(cvt_u64_f32(gx) / m.res);
    uint64_t qy = cvt_f32_u64(cvt_u64_f32(gy) / m.res);
    uint64_t qz = cvt_f32_u64(cvt_u64_f32(gz) / m.res);
    if (qx >= m.dim_x || qy >= m.dim_y || qz >= m.dim_z) return m.data_size;
    // :67-84: x_offset / 1.0f == x_offset, so the block coordinate equals gx
    float spx = ox - cvt_u64_f32(gx) * 1.0f;
    float spy = oy - cvt_u64_f32(gy) * 1.0f;
    float spz = oz - cvt_u64_f32(gz) * 1.0f;
    uint64_t up = gx + gy * m.up_dim_x + gz * m.up_dim_2;
    if (up >= m.grid_occ_size) return m.data_size;
    int32_t s = m.grid_occ[up];
    if (s < 0) return m.data_size;
    uint64_t sx = cvt_f32_u64(spx / m.res);
    uint64_t sy = cvt_f32_u64(spy / m.res);
    uint64_t sz = cvt_f32_u64(spz / m.res);
    return (uint64_t)(int64_t)s + sx + sy * m.sub_dim + sz * m.sub_dim_2;
}

// Voxel index of a map-frame point.  Offsets in [0, 2^23) use 32-bit
// integer conversions (exact there) and an integer form of the grid check
// (u < gthr, derived on the host from the same float division); the rest
// goes through lookup_index_general.
__device__ __forceinline__ uint64_t lookup_index(const DevMap& m, float tx, float ty, float tz)
{
    const float ox = tx - m.min_x;
    const float oy = ty - m.min_y;
    const float oz = tz - m.min_z;
    const bool fast = ox >= 0.0f && oy >= 0.0f && oz >= 0.0f && ox < 8388608.0f &&
                      oy < 8388608.0f && oz < 8388608.0f;
    if (__builtin_expect(fast, 1)) {
        const uint32_t ux = (uint32_t)ox, uy = (uint32_t)oy, uz = (uint32_t)oz;
        if (ux >= m.gthr_x || uy >= m.gthr_y || uz >= m.gthr_z) return m.data_size;
        const uint64_t up = (uint64_t)ux + (uint64_t)uy * m.up_dim_x + (uint64_t)uz * m.up_dim_2;
        if (up >= m.grid_occ_size) return m.data_size;
        const int32_t s = m.grid_occ[up];
        if (s < 0) return m.data_size;
        const float spx = ox - (float)ux;
        const float spy = oy - (float)uy;
        const float spz = oz - (float)uz;
        const uint32_t sx = (uint32_t)(spx / m.res);
        const uint32_t sy = (uint32_t)(spy / m.res);
        const uint32_t sz = (uint32_t)(spz / m.res);
        return (uint64_t)(int64_t)s + sx + (uint64_t)sy * m.sub_dim + (uint64_t)sz * m.sub_dim_2;
    }
    return lookup_index_general(m, ox, oy, oz);
}

}  // namespace tsl
