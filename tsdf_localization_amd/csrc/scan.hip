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

#include <algorithm>

namespace tsl {

namespace {

// particle-order buckets: kXyBits bits for each of x and y, kYawBits for
// the heading
#ifndef TSL_ORDER_XY_BITS
#define TSL_ORDER_XY_BITS 6
#endif
#ifndef TSL_ORDER_YAW_BITS
#define TSL_ORDER_YAW_BITS 5
#endif
constexpr uint32_t kXyBits = TSL_ORDER_XY_BITS;
constexpr uint32_t kYawBits = TSL_ORDER_YAW_BITS;
constexpr uint32_t kBuckets = 1u << (2 * kXyBits + kYawBits);

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

// Evaluation order of the particles: a counting sort into Morton-ordered
// buckets of (x, y) cells over the map and yaw sectors.  Particles close in
// position and heading transform the scan onto nearly the same voxels, so
// evaluating them together on one XCD turns most of their voxel gathers
// into L2 hits; the thousands of particles an XCD has in flight at once
// span many buckets, so the order inside a bucket (arrival order of the
// scatter's atomics) does not matter, and no particle's weight depends on
// its slot.
struct BucketGrid {
    float min_x, min_y, inv_x, inv_y, inv_yaw;
};

__device__ __forceinline__ uint32_t bucket_coord(float v, float inv_cell, uint32_t cells)
{
    float c = floorf(v * inv_cell);
    c = c > 0.0f ? c : 0.0f;  // NaN -> 0
    const float top = (float)(cells - 1);
    c = c < top ? c : top;
    return (uint32_t)c;
}

// Morton key: from the lowest bit up, bit k of x, of y (k < kXyBits) and of
// the heading sector (k < kYawBits)
__device__ __forceinline__ uint32_t bucket_of(const tsl_particle& p, const BucketGrid& g)
{
    const uint32_t x = bucket_coord(p.first[0] - g.min_x, g.inv_x, 1u << kXyBits);
    const uint32_t y = bucket_coord(p.first[1] - g.min_y, g.inv_y, 1u << kXyBits);
    const uint32_t a = bucket_coord(p.first[5] + 3.14159265f, g.inv_yaw, 1u << kYawBits);
    uint32_t key = 0, pos = 0;
#pragma unroll
    for (uint32_t k = 0; k < (kXyBits > kYawBits ? kXyBits : kYawBits); ++k) {
        if (k < kXyBits) {
            key |= ((x >> k) & 1u) << pos++;
            key |= ((y >> k) & 1u) << pos++;
        }
        if (k < kYawBits) key |= ((a >> k) & 1u) << pos++;
    }
    return key;
}

__global__ void k_bucket_zero(uint32_t* __restrict__ hist)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kBuckets) hist[i] = 0u;
}

// hist[key] += the number of active lanes with that key, one atomic per
// distinct key of the wavefront (the particles of a tracking filter share a
// few buckets, and per-lane atomics on one counter serialise); returns the
// counter's old value plus the lane's rank among the lanes with its key.
// The lanes are grouped by key first (ballots only), then every group's
// first lane adds its group's size in one atomic instruction for all
// groups, so the wavefront waits for one atomic round trip however many
// keys it holds.  Every lane of the wavefront calls it.
__device__ __forceinline__ uint32_t wave_bucket_add(uint32_t* __restrict__ hist, uint32_t key, bool active)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    uint64_t pending = __ballot(active);
    uint32_t leader_of = lane, rank = 0, group = 0;
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;  // wavefront-uniform
        const uint32_t lk = __builtin_amdgcn_readlane(key, leader);
        const uint64_t same = __ballot(active && key == lk);
        if (active && key == lk) {
            leader_of = (uint32_t)leader;
            rank = (uint32_t)__popcll(same & below);
            group = (uint32_t)__popcll(same);
        }
        pending &= ~same;
    }
    uint32_t base = 0;
    if (active && leader_of == lane) base = atomicAdd(&hist[key], group);
    base = __shfl(base, (int)leader_of, 64);
    return base + rank;
}

__global__ void k_bucket_count(const tsl_particle* __restrict__ in, const uint32_t* __restrict__ count,
                               BucketGrid g, uint32_t* __restrict__ hist, uint32_t* __restrict__ bucket)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < *count;
    uint32_t k = 0;
    if (active) {
        k = bucket_of(in[i], g);
        bucket[i] = k;
    }
    (void)wave_bucket_add(hist, k, active);
}

__global__ void k_bucket_scatter(const uint32_t* __restrict__ count, const uint32_t* __restrict__ bucket,
                                 uint32_t* __restrict__ hist, uint32_t* __restrict__ order)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < *count;
    const uint32_t k = active ? bucket[i] : 0u;
    const uint32_t pos = wave_bucket_add(hist, k, active);
    if (active) order[pos] = i;
}

// ---- scan reduction (TSDFEvaluator::evaluateParticles,
// src/evaluation/tsdf_evaluator.cpp:304-376) ---------------------------------

// keep[j]: the reference skips points closer than 1 m (sqrtf, :316-322);
// the cloud reduction of CudaEvaluator keeps every point
// (src/cuda/cuda_evaluator.cu:84-91)
__global__ void k_red_keep(const tsl_point* __restrict__ raw, uint32_t n, int mode, uint32_t* __restrict__ keep)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (mode == kReduceCloud) {
        keep[j] = 1u;
        return;
    }
    const tsl_point p = raw[j];
    const float dist = sqrtf(p.x * p.x + p.y * p.y + p.z * p.z);
    keep[j] = (dist < 1.0f) ? 0u : 1u;
}

// Kept point j becomes entry k = kidx[j] (its index in the reference); it
// reads ring[k] because the reference's ring iterator only advances on kept
// points (:319-331).  Entries past the kept count are padded with ring 255.
// ring_out is the dedup part of the second sort key, ring_ord the order key.
// Cloud mode (cuda_evaluator.cu:78-116): every point kept, rings are int32
// (the reference reads the field as int) used for the order only, the
// dedup key is the centre alone, computed in double with the literal
// 0.064 m cell (:95-97), and the centre itself is the reduced point.
__global__ void k_red_build(const tsl_point* __restrict__ raw, const void* __restrict__ ring_any, uint32_t n,
                            int mode, const uint32_t* __restrict__ keep, const uint32_t* __restrict__ kidx,
                            float cell, float half, uint64_t* __restrict__ key_a, uint32_t* __restrict__ ring_out,
                            uint32_t* __restrict__ ring_ord, uint32_t* __restrict__ cx_out,
                            tsl_point* __restrict__ orig, uint32_t* __restrict__ vals, uint32_t* __restrict__ bad)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    vals[j] = j;
    const uint32_t kept_total = kidx[n - 1] + keep[n - 1];
    if (j >= kept_total) {  // padding entry (filled by the thread of the same slot)
        key_a[j] = ~0ull;
        ring_out[j] = 255u;
        ring_ord[j] = 0xffffffffu;
        cx_out[j] = 0xffffffffu;
    }
    if (!keep[j]) return;
    const uint32_t k = kidx[j];
    const tsl_point p = raw[j];
    if (mode == kReduceCloud) {
        const int r = static_cast<const int32_t*>(ring_any)[k];
        const double c = 0.064, h = 0.032;
        const float cx = (float)(floor((double)p.x / c) * c + h);
        const float cy = (float)(floor((double)p.y / c) * c + h);
        const float cz = (float)(floor((double)p.z / c) * c + h);
        key_a[k] = ((uint64_t)__float_as_uint(cy) << 32) | __float_as_uint(cz);
        ring_out[k] = 0u;
        ring_ord[k] = (uint32_t)r ^ 0x80000000u;  // signed order
        cx_out[k] = __float_as_uint(cx);
        orig[k] = tsl_point{cx, cy, cz};
        return;
    }
    const int r = static_cast<const int16_t*>(ring_any)[k];
    if (r < 0 || r >= 64) atomicOr(bad, 1u);
    const float cx = floorf(p.x / cell) * cell + half;
    const float cy = floorf(p.y / cell) * cell + half;
    const float cz = floorf(p.z / cell) * cell + half;
    key_a[k] = ((uint64_t)__float_as_uint(cy) << 32) | __float_as_uint(cz);
    ring_out[k] = (uint32_t)(r & 0xff);
    ring_ord[k] = (uint32_t)(r & 0xff);
    cx_out[k] = __float_as_uint(cx);
    orig[k] = p;
}

// key of the second (stable) pass: ring, then the x centre
__global__ void k_red_key_b(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ ring,
                            const uint32_t* __restrict__ cx, uint32_t n, uint64_t* __restrict__ key_b)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t e = perm[t];
    key_b[t] = ((uint64_t)ring[e] << 32) | cx[e];
}

// first entry of each (ring, centre) group wins; NaN centres never match;
// winners are ordered by (ring_ord, entry)
__global__ void k_red_mark(const uint32_t* __restrict__ perm, const uint64_t* __restrict__ key_a,
                           const uint64_t* __restrict__ key_b_sorted, const uint32_t* __restrict__ ring,
                           const uint32_t* __restrict__ cx, const uint32_t* __restrict__ keep,
                           const uint32_t* __restrict__ kidx, uint32_t n, uint64_t* __restrict__ out_key,
                           uint32_t* __restrict__ win)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t kept_total = kidx[n - 1] + keep[n - 1];
    const uint32_t e = perm[t];
    bool w = false;
    if (t < kept_total) {
        const uint64_t ka = key_a[e];
        const float fx = __uint_as_float(cx[e]);
        const float fy = __uint_as_float((uint32_t)(ka >> 32));
        const float fz = __uint_as_float((uint32_t)ka);
        const bool nan = (fx != fx) | (fy != fy) | (fz != fz);
        w = t == 0 || nan || key_b_sorted[t] != key_b_sorted[t - 1] || ka != key_a[perm[t - 1]];
    }
    win[t] = w ? 1u : 0u;
    out_key[t] = w ? (((uint64_t)ring[e] << 32) | e) : ~0ull;  // ring = ring_ord here
}

__global__ void k_red_order(const uint64_t* __restrict__ sorted_key, uint32_t n, uint32_t* __restrict__ order)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    order[r] = (uint32_t)sorted_key[r];  // entry index of the r-th winner (padding past the count)
}

__global__ void k_red_gather(const uint32_t* __restrict__ order, const tsl_point* __restrict__ orig,
                             const uint32_t* __restrict__ count, tsl_point* __restrict__ out)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= *count) return;
    out[r] = orig[order[r]];
}

}  // namespace

// Reduce a raw scan on the device: writes the reduced points to out (room
// for n) and their number to *count_dev, and 1 to *bad_dev if a kept point's
// ring is outside [0, 64).  Scratch buffers belong to the caller's filter.
int launch_reduce_scan(tsl_filter* f, int mode, const tsl_point* raw, const void* ring, uint32_t n, float cell,
                       tsl_point* out, uint32_t* count_dev, uint32_t* bad_dev, hipStream_t stream)
{
    if (n == 0) return TSL_OK;
    const float half = cell / 2;
    DevBuf& sc = f->red_scratch;
    // layout: keep, kidx, ring, ring_ord, cx, vals, perm1, perm2, win, order (uint32 x n each),
    //         key_a, key_b, key_b_sorted, out_key, out_key_sorted (uint64 x n each), orig (tsl_point x n)
    const size_t u32 = sizeof(uint32_t) * (size_t)n, u64 = sizeof(uint64_t) * (size_t)n;
    size_t t_scan = 0, t_sort = 0, t_sum = 0, t_keys = 0;
    TSL_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t_scan, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)n,
                                                   stream));
    TSL_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, t_keys, (uint64_t*)nullptr, (uint64_t*)nullptr, (int)n, 0,
                                                    64, stream));
    TSL_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, t_sort, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                                     (uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 64, stream));
    TSL_HIP_CHECK(hipcub::DeviceReduce::Sum(nullptr, t_sum, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, stream));
    const size_t temp = std::max(std::max(t_scan, t_keys), std::max(t_sort, t_sum));
    const size_t total = 10 * u32 + 5 * u64 + sizeof(tsl_point) * (size_t)n + temp + 17 * 256;
    TSL_TRY(sc.reserve(total));
    char* base = sc.as<char>();
    auto take = [&](size_t bytes) {
        char* p = base;
        base += (bytes + 255) / 256 * 256;
        return p;
    };
    uint32_t* keep = (uint32_t*)take(u32);
    uint32_t* kidx = (uint32_t*)take(u32);
    uint32_t* ring_e = (uint32_t*)take(u32);
    uint32_t* ring_ord = (uint32_t*)take(u32);
    uint32_t* cx = (uint32_t*)take(u32);
    uint32_t* vals = (uint32_t*)take(u32);
    uint32_t* perm1 = (uint32_t*)take(u32);
    uint32_t* perm2 = (uint32_t*)take(u32);
    uint32_t* win = (uint32_t*)take(u32);
    uint32_t* order = (uint32_t*)take(u32);
    uint64_t* key_a = (uint64_t*)take(u64);
    uint64_t* key_a_sorted = (uint64_t*)take(u64);
    uint64_t* key_b = (uint64_t*)take(u64);
    uint64_t* key_b_sorted = (uint64_t*)take(u64);
    uint64_t* out_key = (uint64_t*)take(u64);
    tsl_point* orig = (tsl_point*)take(sizeof(tsl_point) * (size_t)n);
    void* tmp = take(temp);
    if ((size_t)(base - sc.as<char>()) > sc.bytes) return set_error(TSL_ERR_HIP, "scan reduction scratch layout");
    const unsigned g = (n + 255) / 256;
    hipLaunchKernelGGL(k_red_keep, dim3(g), dim3(256), 0, stream, raw, n, mode, keep);
    TSL_HIP_CHECK(hipGetLastError());
    size_t tb = temp;
    TSL_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tb, keep, kidx, (int)n, stream));
    hipLaunchKernelGGL(k_red_build, dim3(g), dim3(256), 0, stream, raw, ring, n, mode, keep, kidx, cell, half, key_a,
                       ring_e, ring_ord, cx, orig, vals, bad_dev);
    TSL_HIP_CHECK(hipGetLastError());
    // pass A: (y centre, z centre), stable; pass B: (ring, x centre), stable
    tb = temp;
    TSL_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, key_a, key_a_sorted, vals, perm1, (int)n, 0, 64, stream));
    hipLaunchKernelGGL(k_red_key_b, dim3(g), dim3(256), 0, stream, perm1, ring_e, cx, n, key_b);
    TSL_HIP_CHECK(hipGetLastError());
    tb = temp;
    TSL_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, key_b, key_b_sorted, perm1, perm2, (int)n, 0, 64, stream));
    hipLaunchKernelGGL(k_red_mark, dim3(g), dim3(256), 0, stream, perm2, key_a, key_b_sorted, ring_ord, cx, keep, kidx,
                       n, out_key, win);
    TSL_HIP_CHECK(hipGetLastError());
    tb = temp;
    TSL_HIP_CHECK(hipcub::DeviceReduce::Sum(tmp, tb, win, count_dev, (int)n, stream));
    // winners by (ring, index): the entry position is the reference's index
    tb = temp;
    TSL_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(tmp, tb, out_key, key_b, (int)n, 0, 64, stream));
    hipLaunchKernelGGL(k_red_order, dim3(g), dim3(256), 0, stream, key_b, n, order);
    TSL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_red_gather, dim3(g), dim3(256), 0, stream, order, orig, count_dev, out);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

// Layout of f->pkeys: bucket counts, bucket offsets (their exclusive scan),
// the bucket of each particle, the scan's temporary storage; each part
// 256-byte aligned.
static size_t round256(size_t b) { return (b + 255) / 256 * 256; }

static int bucket_scan_temp_bytes(size_t* bytes)
{
    static size_t t = 0;
    static bool known = false;
    if (!known) {
        TSL_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                       (int)kBuckets, (hipStream_t)0));
        known = true;
    }
    *bytes = t;
    return TSL_OK;
}

int reserve_particle_order(tsl_filter* f)
{
    const size_t n = f->count_bound;
    size_t temp = 0;
    TSL_TRY(bucket_scan_temp_bytes(&temp));
    TSL_TRY(f->pkeys.reserve(2 * round256(sizeof(uint32_t) * kBuckets) + round256(sizeof(uint32_t) * n) +
                             round256(temp)));
    TSL_TRY(f->porder.reserve(sizeof(uint32_t) * n));
    return f->pinv.reserve(sizeof(uint32_t) * n);
}

int launch_particle_order(tsl_filter* f, hipStream_t stream)
{
    const uint32_t n = (uint32_t)f->count_bound;
    if (n == 0) return TSL_OK;
    TSL_TRY(reserve_particle_order(f));
    size_t temp = 0;
    TSL_TRY(bucket_scan_temp_bytes(&temp));
    char* at = f->pkeys.as<char>();
    uint32_t* hist = reinterpret_cast<uint32_t*>(at);
    at += round256(sizeof(uint32_t) * kBuckets);
    uint32_t* offsets = reinterpret_cast<uint32_t*>(at);
    at += round256(sizeof(uint32_t) * kBuckets);
    uint32_t* bucket = reinterpret_cast<uint32_t*>(at);
    at += round256(sizeof(uint32_t) * n);
    void* scan_temp = at;
    const uint32_t* count = &f->state.as<FilterState>()->count;
    // x and y cells spanning the map's extent (at least 0.5 m), yaw sectors
    // spanning [-pi, pi)
    const tsl_map_coef& c = f->ev->map->coef;
    const float cells = (float)(1u << kXyBits);
    BucketGrid g;
    g.min_x = c.min_x;
    g.min_y = c.min_y;
    g.inv_x = 1.0f / std::max(0.5f, (c.max_x - c.min_x) / cells);
    g.inv_y = 1.0f / std::max(0.5f, (c.max_y - c.min_y) / cells);
    g.inv_yaw = (float)(1u << kYawBits) / 6.2831853f;
    hipLaunchKernelGGL(k_bucket_zero, dim3(kBuckets / 256), dim3(256), 0, stream, hist);
    TSL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_bucket_count, dim3((n + 255) / 256), dim3(256), 0, stream, f->particles.as<tsl_particle>(),
                       count, g, hist, bucket);
    TSL_HIP_CHECK(hipGetLastError());
    size_t tb = temp;
    TSL_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(scan_temp, tb, hist, offsets, (int)kBuckets, stream));
    hipLaunchKernelGGL(k_bucket_scatter, dim3((n + 255) / 256), dim3(256), 0, stream, count, bucket, offsets,
                       f->porder.as<uint32_t>());
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

// Reserve every buffer the scan preparation of up to n points uses, so that
// later preparations allocate nothing (and never synchronise).
int filter_reserve_scan(tsl_filter* f, size_t n)
{
    if (n == 0) n = 1;
    TSL_TRY(filter_reserve_points(f, n));
    TSL_TRY(f->sort_keys.reserve(sizeof(uint64_t) * 2 * n));
    TSL_TRY(f->sort_vals.reserve(sizeof(uint32_t) * n));
    TSL_TRY(f->perm.reserve(sizeof(uint32_t) * n));
    size_t temp = 0;
    uint64_t* keys = f->sort_keys.as<uint64_t>();
    TSL_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, keys, keys + n, f->sort_vals.as<uint32_t>(),
                                                     f->perm.as<uint32_t>(), (int)n, 0, 63, (hipStream_t)0));
    TSL_TRY(f->sort_temp.reserve(temp));
    return TSL_OK;
}

size_t prep_points_temp_bytes(uint32_t n)
{
    size_t temp = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                             (uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 63, (hipStream_t)0);
    return temp;
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

namespace {

// Sorted position i of scan blockIdx.y: the point and its score constant,
// as k_prep_sorted writes them for one filter.
__global__ void k_prep_sorted_seg(const tsl_point* __restrict__ in, const uint32_t* __restrict__ off,
                                  const uint32_t* __restrict__ sorted_idx, float range_term, float a_max, float mr2,
                                  const ScanDest* __restrict__ dst)
{
    const uint32_t r = blockIdx.y;
    const uint32_t b = off[r], n = off[r + 1] - b;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t src = sorted_idx[b + i];
    const tsl_point p = in[src];
    const float d2 = p.x * p.x + p.y * p.y + p.z * p.z;
    dst[r].points[i] = make_float4(p.x, p.y, p.z, d2 < mr2 ? range_term : a_max);
    dst[r].perm[i] = src - b;
}

}  // namespace

size_t prep_points_batch_temp(uint32_t total, uint32_t n)
{
    size_t t1 = 0, t2 = 0;
    if (total == 0 || n == 0) return 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, t1, (uint64_t*)nullptr, (uint64_t*)nullptr, (uint32_t*)nullptr,
                                           (uint32_t*)nullptr, (int)total, 0, 63, (hipStream_t)0) != hipSuccess ||
        hipcub::DeviceRadixSort::SortPairs(nullptr, t2, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (uint32_t*)nullptr, (int)total, 0, 32, (hipStream_t)0) != hipSuccess)
        return 0;
    return std::max(t1, t2);
}

// Sorted position t: the scan (segment of off) that point order[t] belongs
// to (the last r with off[r] <= order[t]; empty scans own no point).
__global__ void k_seg_keys(const uint32_t* __restrict__ off, uint32_t n, const uint32_t* __restrict__ order,
                           uint32_t total, uint32_t* __restrict__ seg)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint32_t idx = order[t];
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= idx) lo = mid; else hi = mid;
    }
    seg[t] = lo;
}

// Every scan of a tick prepared at once: one device-wide stable sort of all
// points by Morton key, then a stable sort by scan.  Within a scan the order
// is by key, ties in point order: the order the per-scan sort gives.
int launch_prep_points_batch(const tsl_evaluator* ev, const tsl_point* pts, const uint32_t* off, const ScanDest* dst,
                             uint32_t n, uint32_t total, uint32_t max_count, uint64_t* keys, uint32_t* vals,
                             void* temp, size_t temp_bytes, hipStream_t stream)
{
    if (n == 0 || total == 0) return TSL_OK;
    const unsigned g = (total + 255) / 256;
    hipLaunchKernelGGL(k_morton, dim3(g), dim3(256), 0, stream, pts, total, keys, vals);
    TSL_HIP_CHECK(hipGetLastError());
    TSL_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys + total, vals, vals + total,
                                                     (int)total, 0, 63, stream));
    // the keys' first half is free again: scan numbers, in and out
    uint32_t* seg = reinterpret_cast<uint32_t*>(keys);
    hipLaunchKernelGGL(k_seg_keys, dim3(g), dim3(256), 0, stream, off, n, vals + total, total, seg);
    TSL_HIP_CHECK(hipGetLastError());
    int bits = 1;
    while ((1u << bits) < n) ++bits;
    TSL_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, seg, seg + total, vals + total, vals,
                                                     (int)total, 0, bits, stream));
    hipLaunchKernelGGL(k_prep_sorted_seg, dim3((max_count + 255) / 256, n), dim3(256), 0, stream, pts, off, vals,
                       ev->range_term, ev->a_max, ev->max_range_squared, dst);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

}  // namespace tsl
