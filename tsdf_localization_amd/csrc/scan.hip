// Scan preparation on the device: spatial ordering of the points and the
// per-point score constant.
//
// The evaluation kernel gives each wavefront 64 consecutive points.  Points
// kept in the reference's ring order spread over metres of wall, so one
// wavefront's voxel gathers touch ~60 distinct cache lines; sorted by the
// Morton code of their scanner-frame position (0.25 m cells) they cluster,
// and every particle's rigid transform keeps them clustered.  The order of
// the points only changes the float summation order of the weights (the
// voxel index of every (particle, point) pair is unchanged); perm[j] keeps
// the caller's index of sorted point j for the index trace.
#include <hipcub/hipcub.hpp>

#include "tsl_internal.h"

namespace tsl {

namespace {

__device__ __forceinline__ uint64_t spread3(uint64_t v)
{
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__device__ __forceinline__ uint64_t cell_coord(float x)
{
    // 0.25 m cells, offset so that +-262 km maps into 21 bits; NaN -> 0
    float c = floorf(x * 4.0f) + 1048576.0f;
    c = c > 0.0f ? c : 0.0f;
    c = c < 2097151.0f ? c : 2097151.0f;
    return (uint64_t)c;
}

__global__ void k_morton(const tsl_point* __restrict__ in, uint32_t n, uint64_t* __restrict__ keys,
                         uint32_t* __restrict__ vals)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const tsl_point p = in[i];
    keys[i] = spread3(cell_coord(p.x)) | (spread3(cell_coord(p.y)) << 1) | (spread3(cell_coord(p.z)) << 2);
    vals[i] = i;
}

// Sorted point j = input point perm[j], with its constant score term
// (tsdf_evaluator.cpp:56-65): |p|^2 evaluated as the reference does.
__global__ void k_prep_sorted(const tsl_point* __restrict__ in, const uint32_t* __restrict__ perm, uint32_t n,
                              float range_term, float a_max, float mr2, float4* __restrict__ out)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const tsl_point p = in[perm[j]];
    const float d2 = p.x * p.x + p.y * p.y + p.z * p.z;
    out[j] = make_float4(p.x, p.y, p.z, d2 < mr2 ? range_term : a_max);
}

// Evaluation order of the particles: Morton code of (x, y) in 0.5 m cells
// and yaw in 0.1 rad cells.  Particles close in position and heading
// transform the scan onto nearly the same voxels, so evaluating them
// together on one XCD turns most of their voxel gathers into L2 hits.
// Slots past the device count get the largest key and sort last.
__device__ __forceinline__ uint64_t quant(float v, float inv_cell)
{
    float c = floorf(v * inv_cell) + 1048576.0f;
    c = c > 0.0f ? c : 0.0f;  // NaN -> 0
    c = c < 2097151.0f ? c : 2097151.0f;
    return (uint64_t)c;
}

__global__ void k_particle_keys(const tsl_particle* __restrict__ in, const uint32_t* __restrict__ count,
                                uint32_t n_bound, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_bound) return;
    uint64_t key = ~0ull;
    if (i < *count) {
        const tsl_particle p = in[i];
        key = spread3(quant(p.first[0], 2.0f)) | (spread3(quant(p.first[1], 2.0f)) << 1) |
              (spread3(quant(p.first[5], 10.0f)) << 2);
    }
    keys[i] = key;
    vals[i] = i;
}

}  // namespace

int launch_particle_order(tsl_filter* f, hipStream_t stream)
{
    const uint32_t n = (uint32_t)f->count_bound;
    if (n == 0) return TSL_OK;
    TSL_TRY(f->pkeys.reserve(sizeof(uint64_t) * 2 * n));
    TSL_TRY(f->pvals.reserve(sizeof(uint32_t) * n));
    TSL_TRY(f->porder.reserve(sizeof(uint32_t) * n));
    TSL_TRY(f->pinv.reserve(sizeof(uint32_t) * n));
    uint64_t* keys = f->pkeys.as<uint64_t>();
    size_t temp = 0;
    TSL_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, keys, keys + n, f->pvals.as<uint32_t>(),
                                                     f->porder.as<uint32_t>(), (int)n, 0, 64, stream));
    TSL_TRY(f->ptemp.reserve(temp));
    hipLaunchKernelGGL(k_particle_keys, dim3((n + 255) / 256), dim3(256), 0, stream, f->particles.as<tsl_particle>(),
                       &f->state.as<FilterState>()->count, n, keys, f->pvals.as<uint32_t>());
    TSL_HIP_CHECK(hipGetLastError());
    TSL_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(f->ptemp.ptr, temp, keys, keys + n, f->pvals.as<uint32_t>(),
                                                     f->porder.as<uint32_t>(), (int)n, 0, 64, stream));
    return TSL_OK;
}

int launch_prep_points(tsl_filter* f, hipStream_t stream)
{
    const uint32_t n = (uint32_t)f->num_points;
    if (n == 0) return TSL_OK;
    TSL_TRY(f->sort_keys.reserve(sizeof(uint64_t) * 2 * n));
    TSL_TRY(f->sort_vals.reserve(sizeof(uint32_t) * n));
    TSL_TRY(f->perm.reserve(sizeof(uint32_t) * n));
    uint64_t* keys = f->sort_keys.as<uint64_t>();
    uint32_t* vals = f->sort_vals.as<uint32_t>();
    size_t temp = 0;
    TSL_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, keys, keys + n, vals, f->perm.as<uint32_t>(),
                                                     (int)n, 0, 63, stream));
    TSL_TRY(f->sort_temp.reserve(temp));
    const unsigned g = (n + 255) / 256;
    hipLaunchKernelGGL(k_morton, dim3(g), dim3(256), 0, stream, f->points_in.as<tsl_point>(), n, keys, vals);
    TSL_HIP_CHECK(hipGetLastError());
    TSL_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(f->sort_temp.ptr, temp, keys, keys + n, vals,
                                                     f->perm.as<uint32_t>(), (int)n, 0, 63, stream));
    const tsl_evaluator* ev = f->ev;
    hipLaunchKernelGGL(k_prep_sorted, dim3(g), dim3(256), 0, stream, f->points_in.as<tsl_point>(),
                       f->perm.as<uint32_t>(), n, ev->range_term, ev->a_max, ev->max_range_squared,
                       f->points.as<float4>());
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

}  // namespace tsl
