// Sensor-update kernels for gfx950.
//
// Reference path (CPU semantics, src/evaluation/tsdf_evaluator.cpp:78-221):
//   for every particle: pose matrix L (a1), for every point: t = L*p (a2),
//   v = map.getEntry(t) (a3), term = a_hit*v + (|p|^2 < mr^2 ? a_range/mr :
//   a_max) (a4), w += term; then sum, w /= sum, weighted pose (a6).
//
// Work decomposition: a wavefront owns K particles and walks a chunk of the
// scan with its 64 lanes on consecutive points, so the K pose matrices are
// wavefront-uniform (scalar registers) and neighbouring lanes transform
// neighbouring scan points, which land in neighbouring voxels: the block
// table and voxel gathers of a wavefront hit few cache lines.  The voxel
// value palette sits in LDS.  Per-particle sums are reduced across the
// wavefront with a fixed butterfly and across point chunks in a fixed
// order, so results are deterministic.
#include "tsl_lookup.h"

namespace tsl {

namespace {

constexpr int kBlock = 256;           // threads per workgroup (4 wavefronts)
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide double sum with a fixed combination order; result valid in
// thread 0.
__device__ double block_sum_d(double v, double* sh)
{
    v = wave_sum_d(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += sh[w];
    }
    __syncthreads();
    return r;
}

struct Tf12 {
    float m[16];
};

__global__ void k_pose(const tsl_particle* __restrict__ particles, const uint32_t* __restrict__ count,
                       uint32_t n_static, Tf12 tf, float* __restrict__ L)
{
    const uint32_t n = count ? *count : n_static;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float pose[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) pose[k] = particles[i].first[k];
    float out[12];
    pose_matrix(pose, tf.m, out);
#pragma unroll
    for (int k = 0; k < 12; ++k) L[(size_t)i * 12 + k] = out[k];
}

// Scan points -> (x, y, z, c) with c the point's constant score term
// (tsdf_evaluator.cpp:56-65): |p|^2 is evaluated as the reference does.
__global__ void k_prep_points(const tsl_point* __restrict__ in, uint32_t n, float range_term,
                              float a_max, float mr2, float4* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = in[i].x, y = in[i].y, z = in[i].z;
    const float d2 = x * x + y * y + z * z;
    out[i] = make_float4(x, y, z, d2 < mr2 ? range_term : a_max);
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_eval(DevMap map, const float* __restrict__ L,
                                                 const uint32_t* __restrict__ count,
                                                 const float4* __restrict__ pts, uint32_t N,
                                                 uint32_t chunk_len, float a_hit,
                                                 float* __restrict__ partial, uint32_t stride)
{
    extern __shared__ float lut[];
    for (uint32_t i = threadIdx.x; i < map.palette_size; i += blockDim.x) lut[i] = map.palette[i];
    __syncthreads();

    const uint32_t P = *count;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t p0 = (blockIdx.x * kWaves + wave) * K;
    if (p0 >= P) return;
    const uint32_t j0 = blockIdx.y * chunk_len;
    const uint32_t j1 = min(N, j0 + chunk_len);

    float m[K][12];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t p = min(p0 + k, P - 1);
#pragma unroll
        for (int e = 0; e < 12; ++e) m[k][e] = L[(size_t)p * 12 + e];
    }
    float acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0f;

    for (uint32_t j = j0 + lane; j < j1; j += 64) {
        const float4 q = pts[j];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const float tx = m[k][0] * q.x + m[k][1] * q.y + m[k][2] * q.z + m[k][3];
            const float ty = m[k][4] * q.x + m[k][5] * q.y + m[k][6] * q.z + m[k][7];
            const float tz = m[k][8] * q.x + m[k][9] * q.y + m[k][10] * q.z + m[k][11];
            const uint64_t idx = lookup_index(map, tx, ty, tz);
            const float v = idx < map.data_size ? lut[map.codes[idx]] : map.init;
            acc[k] += a_hit * v + q.w;
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = wave_sum(acc[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (p0 + k < P) partial[(size_t)blockIdx.y * stride + p0 + k] = acc[k];
    }
}

// raw[p] = sum over chunks (ascending) of partial[c][p], in double, stored as
// float; block partial sums of raw in double; the last block to finish adds
// the block partials in index order into stats[0].
__global__ __launch_bounds__(kBlock) void k_reduce(const float* __restrict__ partial, uint32_t chunks,
                                                   uint32_t stride, float* __restrict__ raw,
                                                   double* __restrict__ blocks, FilterState* st)
{
    __shared__ double sh[kWaves];
    __shared__ bool last;
    const uint32_t P = st->count;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double w = 0.0;
    if (i < P) {
        double acc = 0.0;
        for (uint32_t c = 0; c < chunks; ++c) acc += (double)partial[(size_t)c * stride + i];
        const float wf = (float)acc;
        raw[i] = wf;
        w = (double)wf;
    }
    const double b = block_sum_d(w, sh);
    if (threadIdx.x == 0) {
        blocks[blockIdx.x] = b;
        __threadfence();
        const uint32_t done = atomicAdd(&st->counter0, 1u);
        last = (done == gridDim.x - 1);
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        __threadfence();
        double total = 0.0;
        for (uint32_t k = 0; k < gridDim.x; ++k) total += __builtin_nontemporal_load(&blocks[k]);
        st->stats[0] = total;
        st->status = ((float)total == 0.0f) ? 1u : 0u;
        st->counter0 = 0;
    }
}

// w_i = raw_i / (float)sum written to particles[i].second (the reference's
// particles[index].second /= weight_sum, float division), plus the nine
// weighted pose sums (tsdf_evaluator.cpp:198-217) in double.
__global__ __launch_bounds__(kBlock) void k_normalize(const float* __restrict__ raw,
                                                      tsl_particle* __restrict__ particles,
                                                      double* __restrict__ blocks, FilterState* st)
{
    __shared__ double sh[kWaves];
    __shared__ bool last;
    const uint32_t P = st->count;
    const float sum = (float)st->stats[0];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double q[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) q[k] = 0.0;
    if (i < P && sum != 0.0f) {
        const float w = raw[i] / sum;
        particles[i].second = w;
        const double wd = (double)w;
        q[0] = (double)particles[i].first[0] * wd;
        q[1] = (double)particles[i].first[1] * wd;
        q[2] = (double)particles[i].first[2] * wd;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float s, c;
            sincosf_exact(particles[i].first[3 + a], &s, &c);
            q[3 + 2 * a] = (double)s * wd;
            q[4 + 2 * a] = (double)c * wd;
        }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const double b = block_sum_d(q[k], sh);
        if (threadIdx.x == 0) blocks[(size_t)k * gridDim.x + blockIdx.x] = b;
    }
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t done = atomicAdd(&st->counter1, 1u);
        last = (done == gridDim.x - 1);
    }
    __syncthreads();
    if (last) {
        __threadfence();
        if (threadIdx.x < 9) {
            double t = 0.0;
            for (uint32_t k = 0; k < gridDim.x; ++k)
                t += __builtin_nontemporal_load(&blocks[(size_t)threadIdx.x * gridDim.x + k]);
            st->stats[1 + threadIdx.x] = t;
        }
        __syncthreads();
        if (threadIdx.x == 0) st->counter1 = 0;
    }
}

__global__ void k_trace(DevMap map, const float* __restrict__ L, uint32_t P,
                        const float4* __restrict__ pts, uint32_t N, uint64_t* __restrict__ out)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)P * N) return;
    const uint32_t p = (uint32_t)(t / N), j = (uint32_t)(t % N);
    const float* m = L + (size_t)p * 12;
    const float4 q = pts[j];
    const float tx = m[0] * q.x + m[1] * q.y + m[2] * q.z + m[3];
    const float ty = m[4] * q.x + m[5] * q.y + m[6] * q.z + m[7];
    const float tz = m[8] * q.x + m[9] * q.y + m[10] * q.z + m[11];
    out[t] = lookup_index(map, tx, ty, tz);
}

// Copy the device count into the filter state (count comes from the host).
__global__ void k_set_count(FilterState* st, uint32_t n)
{
    st->count = n;
}

struct Tile {
    int K;
    uint32_t chunk_len;
    uint32_t chunks;
};

// Tile shape: enough (particle-group, point-chunk) wavefronts to fill 256
// CUs several times over, chunks no shorter than 512 points.
Tile choose_tile(size_t P, size_t N, int per_point)
{
    Tile t;
    t.K = 4;
    const size_t groups = (P + t.K - 1) / t.K;
    const size_t target_waves = 256 * 4 * 8;
    size_t chunks = 1;
    if (per_point || groups < target_waves) {
        chunks = (target_waves + groups - 1) / groups;
        const size_t max_chunks = (N + 511) / 512;
        if (chunks > max_chunks) chunks = max_chunks;
        if (chunks < 1) chunks = 1;
    }
    size_t len = (N + chunks - 1) / chunks;
    len = (len + 63) / 64 * 64;
    if (len == 0) len = 64;
    t.chunk_len = (uint32_t)len;
    t.chunks = (uint32_t)((N + len - 1) / len);
    if (t.chunks == 0) t.chunks = 1;
    return t;
}

}  // namespace

int launch_pose_matrices(const tsl_particle* d_particles, uint32_t n, const float tf[16], float* d_out,
                         hipStream_t stream)
{
    if (n == 0) return TSL_OK;
    Tf12 t;
    for (int k = 0; k < 16; ++k) t.m[k] = tf[k];
    hipLaunchKernelGGL(k_pose, dim3((n + 255) / 256), dim3(256), 0, stream, d_particles, nullptr, n, t,
                       d_out);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

int launch_prep_points(tsl_filter* f, hipStream_t stream)
{
    const uint32_t n = (uint32_t)f->num_points;
    if (n == 0) return TSL_OK;
    const tsl_evaluator* ev = f->ev;
    hipLaunchKernelGGL(k_prep_points, dim3((n + 255) / 256), dim3(256), 0, stream,
                       f->points_in.as<tsl_point>(), n, ev->range_term, ev->a_max,
                       ev->max_range_squared, f->points.as<float4>());
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

static int launch_eval_core(tsl_filter* f, hipStream_t stream, Tile& tile)
{
    const size_t P = f->count_bound;
    const size_t N = f->num_points;
    FilterState* st = f->state.as<FilterState>();
    Tf12 t;
    for (int k = 0; k < 16; ++k) t.m[k] = f->tf[k];
    hipLaunchKernelGGL(k_pose, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, stream,
                       f->particles.as<tsl_particle>(), &st->count, (uint32_t)P, t, f->L.as<float>());
    TSL_HIP_CHECK(hipGetLastError());
    tile = choose_tile(P, N, f->ev->per_point);
    TSL_TRY(f->partial.reserve(sizeof(float) * (size_t)tile.chunks * f->capacity));
    const tsl_map* map = f->ev->map;
    const size_t groups = (P + tile.K - 1) / tile.K;
    dim3 grid((unsigned)((groups + kWaves - 1) / kWaves), tile.chunks);
    const size_t lds = sizeof(float) * map->dev.palette_size;
    hipLaunchKernelGGL(k_eval<4>, grid, dim3(kBlock), lds, stream, map->dev, f->L.as<float>(),
                       &st->count, f->points.as<float4>(), (uint32_t)N, tile.chunk_len, f->ev->a_hit,
                       f->partial.as<float>(), (uint32_t)f->capacity);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

int launch_update(tsl_filter* f, hipStream_t stream)
{
    const size_t P = f->count_bound;
    if (P == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "filter holds no particles");
    Tile tile;
    TSL_TRY(launch_eval_core(f, stream, tile));
    FilterState* st = f->state.as<FilterState>();
    const unsigned nb = (unsigned)((P + kBlock - 1) / kBlock);
    TSL_TRY(f->blocks.reserve(sizeof(double) * 9 * (size_t)nb));
    hipLaunchKernelGGL(k_reduce, dim3(nb), dim3(kBlock), 0, stream, f->partial.as<float>(), tile.chunks,
                       (uint32_t)f->capacity, f->raw.as<float>(), f->blocks.as<double>(), st);
    TSL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_normalize, dim3(nb), dim3(kBlock), 0, stream, f->raw.as<float>(),
                       f->particles.as<tsl_particle>(), f->blocks.as<double>(), st);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

int launch_raw(tsl_filter* f, hipStream_t stream, float* raw_out_device)
{
    const size_t P = f->count_bound;
    if (P == 0) return TSL_OK;
    Tile tile;
    TSL_TRY(launch_eval_core(f, stream, tile));
    FilterState* st = f->state.as<FilterState>();
    const unsigned nb = (unsigned)((P + kBlock - 1) / kBlock);
    TSL_TRY(f->blocks.reserve(sizeof(double) * 9 * (size_t)nb));
    hipLaunchKernelGGL(k_reduce, dim3(nb), dim3(kBlock), 0, stream, f->partial.as<float>(), tile.chunks,
                       (uint32_t)f->capacity, raw_out_device, f->blocks.as<double>(), st);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

int launch_trace(tsl_filter* f, hipStream_t stream, uint64_t* out_device)
{
    const size_t P = f->count_bound;
    const size_t N = f->num_points;
    if (P == 0 || N == 0) return TSL_OK;
    TSL_TRY(launch_pose_matrices(f->particles.as<tsl_particle>(), (uint32_t)P, f->tf, f->L.as<float>(),
                                 stream));
    const uint64_t total = (uint64_t)P * N;
    hipLaunchKernelGGL(k_trace, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream,
                       f->ev->map->dev, f->L.as<float>(), (uint32_t)P, f->points.as<float4>(),
                       (uint32_t)N, out_device);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

int launch_set_count(tsl_filter* f, uint32_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, stream, f->state.as<FilterState>(), n);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

}  // namespace tsl
