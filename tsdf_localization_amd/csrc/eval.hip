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

// One wavefront: K particles (pose matrices wavefront-uniform) x the points
// [j0, j1) of one chunk, 64 consecutive points per iteration.  Each
// iteration picks, by wavefront-uniform ballot, the cheapest lookup tier
// every lane qualifies for (fast, else wrap, else per-lane wrap/general).
// kMode 1 writes the voxel index of every pair, kMode 2 the voxel value,
// instead of accumulating the score, with the same tiling, tier selection
// and addressing as the production kernel (kMode 0); trace_out is indexed
// by the caller's point order (perm).
template <int K, int kMode, bool kU24, bool kBrick>
__global__ __launch_bounds__(kBlock) void k_eval(DevMap map, const float* __restrict__ L,
                                                 const uint32_t* __restrict__ count,
                                                 const float4* __restrict__ pts, uint32_t N,
                                                 uint32_t chunk_len, float a_hit,
                                                 float* __restrict__ partial, uint32_t stride,
                                                 void* __restrict__ trace_out, const uint32_t* __restrict__ perm)
{
    constexpr bool kTrace = kMode != 0;
    extern __shared__ uint4 smem_raw[];
    uint64_t* neg = reinterpret_cast<uint64_t*>(smem_raw);
    SubEntry* tab = reinterpret_cast<SubEntry*>(neg + 8);
    float* lut = reinterpret_cast<float*>(tab + map.sub_entries);
    if (threadIdx.x < 8) neg[threadIdx.x] = map.neg_tab[threadIdx.x];
    for (uint32_t i = threadIdx.x; i < map.sub_entries; i += blockDim.x) tab[i] = map.sub_tab[i];
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

    for (uint32_t base = j0; base < j1; base += 64) {
        const uint32_t j = base + lane;
        const bool act = j < j1;
        const float4 q = pts[act ? j : j0];
        float ox[K], oy[K], oz[K];
        bool fast = true;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            // tsdf_evaluator.cpp:44-46: ((L0*x + L1*y) + L2*z) + L3, no FMA
            const float tx = m[k][0] * q.x + m[k][1] * q.y + m[k][2] * q.z + m[k][3];
            const float ty = m[k][4] * q.x + m[k][5] * q.y + m[k][6] * q.z + m[k][7];
            const float tz = m[k][8] * q.x + m[k][9] * q.y + m[k][10] * q.z + m[k][11];
            ox[k] = tx - map.min_x;
            oy[k] = ty - map.min_y;
            oz[k] = tz - map.min_z;
            fast = fast & fast_ok(ox[k], oy[k], oz[k], map.range_bits);
        }
        float v[K];
        uint64_t full[K];
        if (__builtin_amdgcn_ballot_w64(!fast) == 0) {
            int32_t s[K];
            uint32_t small[K], pbrick[K];
            bool carry[K];
#pragma unroll
            for (int k = 0; k < K; ++k)
                fast_block<kU24, kBrick>(map, tab, ox[k], oy[k], oz[k], s[k], small[k], pbrick[k], carry[k]);
            uint32_t code[K];
            bool valid[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint32_t sum = (uint32_t)s[k] + small[k];
                valid[k] = (s[k] >= 0) & (sum < map.ds);
                if (kTrace) full[k] = s[k] >= 0 ? (uint64_t)(uint32_t)s[k] + small[k] : map.data_size;
                uint32_t addr = (uint32_t)s[k] + pbrick[k];
                if (kBrick && __builtin_expect(carry[k], 0)) addr = phys_from_linear(map, sum);
                code[k] = 0;
                if (valid[k]) code[k] = map.codes[addr];
            }
#pragma unroll
            for (int k = 0; k < K; ++k) v[k] = valid[k] ? lut[code[k]] : map.init;
        } else {
            bool wrap = true;
#pragma unroll
            for (int k = 0; k < K; ++k) wrap = wrap & wrap_ok(ox[k], oy[k], oz[k], map.range_bits);
            if (__builtin_amdgcn_ballot_w64(!wrap) == 0) {
#pragma unroll
                for (int k = 0; k < K; ++k) full[k] = wrap_index<kU24>(map, tab, neg, ox[k], oy[k], oz[k]);
            } else {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if (wrap_ok(ox[k], oy[k], oz[k], map.range_bits))
                        full[k] = wrap_index<kU24>(map, tab, neg, ox[k], oy[k], oz[k]);
                    else
                        full[k] = lookup_index_general(*map.full, ox[k], oy[k], oz[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                v[k] = map.init;
                if (full[k] < map.data_size) {
                    const uint32_t x = (uint32_t)full[k];
                    v[k] = lut[map.codes[kBrick ? phys_from_linear(map, x) : x]];
                }
            }
        }
        if (kMode == 1) {
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (act && p0 + k < P) static_cast<uint64_t*>(trace_out)[(size_t)(p0 + k) * N + perm[j]] = full[k];
        } else if (kMode == 2) {
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (act && p0 + k < P) static_cast<float*>(trace_out)[(size_t)(p0 + k) * N + perm[j]] = v[k];
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (act) acc[k] += a_hit * v[k] + q.w;  // tsdf_evaluator.cpp:56-67
        }
    }
    if (kTrace) return;
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
    if (last) {
        __threadfence();
        // fixed order: thread t adds partials t, t+256, ...; then the block tree
        double t = 0.0;
        for (uint32_t k = threadIdx.x; k < gridDim.x; k += blockDim.x) t += *(volatile const double*)&blocks[k];
        const double total = block_sum_d(t, sh);
        if (threadIdx.x == 0) {
            st->stats[0] = total;
            st->status = ((float)total == 0.0f) ? 1u : 0u;
            st->counter0 = 0;
        }
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
        for (int q = 0; q < 9; ++q) {
            double t = 0.0;
            for (uint32_t k = threadIdx.x; k < gridDim.x; k += blockDim.x)
                t += *(volatile const double*)&blocks[(size_t)q * gridDim.x + k];
            const double total = block_sum_d(t, sh);
            if (threadIdx.x == 0) st->stats[1 + q] = total;
        }
        if (threadIdx.x == 0) st->counter1 = 0;
    }
}

// Copy the device count into the filter state (count comes from the host).
__global__ void k_set_count(FilterState* st, uint32_t n)
{
    st->count = n;
}

// Restore the resident set from the snapshot and reset the count.
__global__ void k_restore(const tsl_particle* __restrict__ src, tsl_particle* __restrict__ dst, uint32_t n,
                          FilterState* st)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
    if (i == 0) st->count = n;
}

// Tile shape: K particles per wavefront (template instantiations 1, 2)
// and the scan chunk a wavefront walks.  Automatic choice: K = 1 (fewest
// scalar registers, full occupancy on gfx950) and enough point chunks to
// give the 256 CUs at least four full rounds of wavefronts.
Tile choose_tile(const tsl_filter* f, size_t P, size_t N)
{
    Tile t;
    t.K = f->tile_k ? f->tile_k : 1;
    const size_t groups = (P + t.K - 1) / t.K;
    size_t len;
    if (f->tile_chunk) {
        len = f->tile_chunk;
    } else {
        const size_t target_waves = 256 * 4 * 8 * 4;
        size_t chunks = 1;
        if (f->ev->per_point || groups < target_waves) {
            chunks = (target_waves + groups - 1) / groups;
            const size_t max_chunks = (N + 255) / 256;
            if (chunks > max_chunks) chunks = max_chunks;
            if (chunks < 1) chunks = 1;
        }
        len = (N + chunks - 1) / chunks;
    }
    len = (len + 63) / 64 * 64;
    if (len == 0) len = 64;
    t.chunk_len = (uint32_t)len;
    t.chunks = (uint32_t)((N + len - 1) / len);
    if (t.chunks == 0) t.chunks = 1;
    return t;
}

template <int kMode>
int launch_eval_kernel(const Tile& tile, const tsl_filter* f, size_t P, size_t N, hipStream_t stream,
                       float* partial, void* trace_out)
{
    const tsl_map* map = f->ev->map;
    FilterState* st = f->state.as<FilterState>();
    const size_t groups = (P + tile.K - 1) / tile.K;
    dim3 grid((unsigned)((groups + kWaves - 1) / kWaves), tile.chunks);
    const size_t lds = 8 * sizeof(uint64_t) + sizeof(SubEntry) * map->dev.sub_entries +
                       sizeof(float) * map->dev.palette_size;
    const bool u24 = map->dev.up_dim_x < (1u << 23) && map->dev.up_dim_2 < (1u << 23) &&
                     map->dev.sub_dim_2 < (1 << 23);
    const bool brick = map->dev.brick != 0;
#define TSL_LAUNCH_EVAL(KK, U24, BR)                                                                            \
    hipLaunchKernelGGL((k_eval<KK, kMode, U24, BR>), grid, dim3(kBlock), lds, stream, map->dev, f->L.as<float>(), \
                       &st->count, f->points.as<float4>(), (uint32_t)N, tile.chunk_len, f->ev->a_hit, partial,     \
                       (uint32_t)f->capacity, trace_out, f->perm.as<uint32_t>())
    const int sel = tile.K * 4 + (u24 ? 2 : 0) + (brick ? 1 : 0);
    switch (sel) {
    case 4: TSL_LAUNCH_EVAL(1, false, false); break;
    case 5: TSL_LAUNCH_EVAL(1, false, true); break;
    case 6: TSL_LAUNCH_EVAL(1, true, false); break;
    case 7: TSL_LAUNCH_EVAL(1, true, true); break;
    case 8: TSL_LAUNCH_EVAL(2, false, false); break;
    case 9: TSL_LAUNCH_EVAL(2, false, true); break;
    case 10: TSL_LAUNCH_EVAL(2, true, false); break;
    case 11: TSL_LAUNCH_EVAL(2, true, true); break;
    default: return set_error(TSL_ERR_INVALID_ARGUMENT, "tile K must be 1 or 2 (got %d)", tile.K);
    }
#undef TSL_LAUNCH_EVAL
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
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
    tile = choose_tile(f, P, N);
    TSL_TRY(f->partial.reserve(sizeof(float) * (size_t)tile.chunks * f->capacity));
    if (f->timing) {
        if (f->timing_pending) {
            // fold the previous bracket in before reusing the events
            float ms = 0.0f;
            TSL_HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(f->ev_stop)));
            TSL_HIP_CHECK(hipEventElapsedTime(&ms, static_cast<hipEvent_t>(f->ev_start),
                                              static_cast<hipEvent_t>(f->ev_stop)));
            f->eval_ms += ms;
            f->timing_pending = false;
        }
        TSL_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(f->ev_start), stream));
    }
    TSL_TRY(launch_eval_kernel<0>(tile, f, P, N, stream, f->partial.as<float>(), nullptr));
    if (f->timing) {
        TSL_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(f->ev_stop), stream));
        f->timing_pending = true;
        f->eval_launches += 1;
        f->eval_pairs += (uint64_t)P * N;
    }
    return TSL_OK;
}

// Time every candidate tile shape on the current particles and scan; keep
// the fastest.  Candidates: K in {1, 2} x {automatic chunks, 1/2/4/8
// chunks of the scan}.
int autotune_tile(tsl_filter* f, hipStream_t stream, int reps, int* best_k, uint32_t* best_chunk, float* best_ms)
{
    const size_t P = f->count_bound;
    const size_t N = f->num_points;
    if (P == 0 || N == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "autotune needs particles and a scan");
    const int saved_k = f->tile_k;
    const uint32_t saved_chunk = f->tile_chunk;
    TSL_TRY(launch_pose_matrices(f->particles.as<tsl_particle>(), (uint32_t)P, f->tf, f->L.as<float>(), stream));
    hipEvent_t a, b;
    TSL_HIP_CHECK(hipEventCreate(&a));
    TSL_HIP_CHECK(hipEventCreate(&b));
    float best = 1e30f;
    int bk = 2;
    uint32_t bc = 0;
    const int ks[2] = {1, 2};
    const uint32_t div[5] = {0, 1, 2, 4, 8};
    int rc = TSL_OK;
    for (int ki = 0; ki < 2 && rc == TSL_OK; ++ki) {
        for (int ci = 0; ci < 5 && rc == TSL_OK; ++ci) {
            f->tile_k = ks[ki];
            f->tile_chunk = div[ci] ? (uint32_t)((N + div[ci] - 1) / div[ci]) : 0u;
            const Tile tile = choose_tile(f, P, N);
            rc = f->partial.reserve(sizeof(float) * (size_t)tile.chunks * f->capacity);
            if (rc != TSL_OK) break;
            rc = launch_eval_kernel<0>(tile, f, P, N, stream, f->partial.as<float>(), nullptr);  // warm-up
            if (rc != TSL_OK) break;
            (void)hipEventRecord(a, stream);
            for (int r = 0; r < reps && rc == TSL_OK; ++r)
                rc = launch_eval_kernel<0>(tile, f, P, N, stream, f->partial.as<float>(), nullptr);
            (void)hipEventRecord(b, stream);
            (void)hipEventSynchronize(b);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, a, b);
            ms /= (float)(reps > 0 ? reps : 1);
            if (ms < best) {
                best = ms;
                bk = ks[ki];
                bc = f->tile_chunk;
            }
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc != TSL_OK) {
        f->tile_k = saved_k;
        f->tile_chunk = saved_chunk;
        return rc;
    }
    f->tile_k = bk;
    f->tile_chunk = bc;
    if (best_k) *best_k = bk;
    if (best_chunk) *best_chunk = bc;
    if (best_ms) *best_ms = best;
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

int launch_trace(tsl_filter* f, hipStream_t stream, void* out_device, int values)
{
    const size_t P = f->count_bound;
    const size_t N = f->num_points;
    if (P == 0 || N == 0) return TSL_OK;
    TSL_TRY(launch_pose_matrices(f->particles.as<tsl_particle>(), (uint32_t)P, f->tf, f->L.as<float>(),
                                 stream));
    const Tile tile = choose_tile(f, P, N);
    if (values) return launch_eval_kernel<2>(tile, f, P, N, stream, nullptr, out_device);
    return launch_eval_kernel<1>(tile, f, P, N, stream, nullptr, out_device);
}

int launch_restore(tsl_filter* f, hipStream_t stream)
{
    const uint32_t n = (uint32_t)f->snapshot_count;
    hipLaunchKernelGGL(k_restore, dim3((n + 255) / 256 ? (n + 255) / 256 : 1), dim3(256), 0, stream,
                       f->snapshot.as<tsl_particle>(), f->particles.as<tsl_particle>(), n,
                       f->state.as<FilterState>());
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
