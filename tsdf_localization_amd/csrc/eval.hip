// Sensor-update kernels for gfx950: pose matrices, the tile choice and
// dispatch of the evaluation kernels (eval_kernels.h, instantiated per tile
// shape in eval_tiles.hip), reduction, normalisation and pose sums, and the
// batched form's plan and launches.
//
// Reference path (CPU semantics, src/evaluation/tsdf_evaluator.cpp:78-221):
//   for every particle: pose matrix L (a1), for every point: t = L*p (a2),
//   v = map.getEntry(t) (a3), term = a_hit*v + (|p|^2 < mr^2 ? a_range/mr :
//   a_max) (a4), w += term; then sum, w /= sum, weighted pose (a6).
//
// Per-particle sums are reduced across the wavefront with a fixed butterfly
// and across point chunks in a fixed order, so results are deterministic.
#include "eval_kernels.h"

namespace tsl {

namespace {

struct Tf12 {
    float m[16];
};

__global__ void k_pose(const tsl_particle* __restrict__ particles, const uint32_t* __restrict__ count,
                       uint32_t n_static, Tf12 tf, int trig, float* __restrict__ L)
{
    const uint32_t n = count ? *count : n_static;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float pose[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) pose[k] = particles[i].first[k];
    float out[12];
    pose_matrix(pose, tf.m, out, trig);
#pragma unroll
    for (int k = 0; k < 12; ++k) L[(size_t)i * 12 + k] = out[k];
}

// Exact lookup of one (particle, point) pair outside the fast tier: wrap
// tier if every axis offset is within R, else the general tier.
__global__ void k_pose_ordered(const tsl_particle* __restrict__ particles, const uint32_t* __restrict__ count,
                               const uint32_t* __restrict__ order, Tf12 tf, int trig, float* __restrict__ L,
                               uint32_t* __restrict__ inv)
{
    const uint32_t n = *count;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = order[i];
    float pose[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) pose[k] = particles[o].first[k];
    float out[12];
    pose_matrix(pose, tf.m, out, trig);
#pragma unroll
    for (int k = 0; k < 12; ++k) L[(size_t)i * 12 + k] = out[k];
    inv[o] = i;
}

// ---------------------------------------------------------------------------
// Member-indexed forms (blockIdx.y = member): a single filter is a batch of
// one.  Sums are formed exactly as in the kernels above, so a filter gives
// the same bits alone and inside a batch: extra blocks of a wider grid add
// zeros after the filter's own blocks.
// ---------------------------------------------------------------------------

__global__ void k_pose_m(const Member* __restrict__ members)
{
    const Member& m = members[blockIdx.y];
    if (m.inv) return;  // pose matrices already written in the member's spatial order
    const uint32_t n = m.st->count;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float pose[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) pose[k] = m.particles[i].first[k];
    float tf[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) tf[k] = m.tf[k];
    float out[12];
    pose_matrix(pose, tf, out, m.trig);
#pragma unroll
    for (int k = 0; k < 12; ++k) m.L[(size_t)i * 12 + k] = out[k];
}

// Reduction partials of a member: blocks[0, nb) the weight sums of the nb
// workgroup-sized slices, blocks[nb + q * nb + b] the nine pose sums of
// slice b.  The member kernels keep them in m.blocks (device memory, read by
// later workgroups and launches); k_finish_small keeps them in LDS.

// The slice functions below run as groups of kBlock threads: kG groups per
// workgroup (1 in the member kernels, 4 in k_finish_small), group g =
// threadIdx.x / kBlock working on its own slice; every thread of the
// workgroup calls them (their barriers are workgroup-wide, and every lane
// of every wavefront is active there).  A group sum is formed in one fixed
// order whatever kG is (the row sums below, then the group's 16-lane rows
// in order from 0.0), so a slice gives the same bits in either kernel.

// v from the lane that DPP control kCtrl selects (both 32-bit halves).
template <int kCtrl>
__device__ __forceinline__ double dpp_d(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// The sum of v over each 16-lane row, in every lane of the row, without the
// LDS: the xor butterfly with partners 1 and 2 (quad permutations), then 4
// and 8 through the half-row and row mirrors.  Addition commutes, so after
// the first two steps the four lanes of a quad hold the same value and
// after the third the eight lanes of a half-row do; the mirrors therefore
// pair every lane with a lane that holds exactly what its xor-4 (xor-8)
// partner holds.
__device__ __forceinline__ double row_sum_d(double v)
{
    v += dpp_d<0xB1>(v);   // quad_perm [1, 0, 3, 2]: lane ^ 1
    v += dpp_d<0x4E>(v);   // quad_perm [2, 3, 0, 1]: lane ^ 2
    v += dpp_d<0x141>(v);  // row_half_mirror: lane 7 - l of its eight
    v += dpp_d<0x140>(v);  // row_mirror: lane 15 - l of its row
    return v;
}

constexpr int kRows = kWaves * 4;  // 16-lane rows per group

template <int kG>
__device__ __forceinline__ double group_sum_d(double v, double* sh /* kRows * kG */)
{
    v = row_sum_d(v);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if ((lane & 15) == 0) sh[wave * 4 + (lane >> 4)] = v;
    __syncthreads();
    double r = 0.0;
    if ((threadIdx.x & (kBlock - 1)) == 0)
        for (int k = 0; k < kRows; ++k) r += sh[wave * 4 + k];
    __syncthreads();
    return r;  // in the group's thread 0
}

template <int kG>
__device__ __forceinline__ double group_sum9(const double (&v)[9], double* sh /* kRows * kG * 9 */)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double w[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) w[k] = row_sum_d(v[k]);
    if ((lane & 15) == 0) {
#pragma unroll
        for (int k = 0; k < 9; ++k) sh[(wave * 4 + (lane >> 4)) * 9 + k] = w[k];
    }
    __syncthreads();
    double r = 0.0;
    const uint32_t lt = threadIdx.x & (kBlock - 1);
    if (lt < 9)
        for (int k = 0; k < kRows; ++k) r += sh[(wave * 4 + k) * 9 + lt];
    __syncthreads();
    return r;  // sum k in the group's thread k < 9
}

// raw[p] = sum over chunks (ascending) of partial[c][p], in double, stored as
// float (and returned); the partial sum of raw in double over slice b
// (particles [b kBlock, (b+1) kBlock)) into blocks[b] (a slice past nb
// contributes and writes nothing).
template <int kG>
__device__ __forceinline__ float reduce_slice(const Member& m, uint32_t P, uint32_t nb, uint32_t b, double* sh,
                                              double* blocks)
{
    const uint32_t lt = threadIdx.x & (kBlock - 1);
    const uint32_t i = b * kBlock + lt;
    double w = 0.0;
    float wf = 0.0f;
    if (b < nb && i < P) {
        const uint32_t slot = m.inv ? m.inv[i] : i;
        // the chunk partials in chunk order, four loads in flight at a time
        const float* part = m.partial + slot;
        double acc = 0.0;
        uint32_t c = 0;
        for (; c + 4 <= m.chunks; c += 4) {
            const float v0 = part[(size_t)c * m.stride], v1 = part[(size_t)(c + 1) * m.stride],
                        v2 = part[(size_t)(c + 2) * m.stride], v3 = part[(size_t)(c + 3) * m.stride];
            acc += (double)v0;
            acc += (double)v1;
            acc += (double)v2;
            acc += (double)v3;
        }
        for (; c < m.chunks; ++c) acc += (double)part[(size_t)c * m.stride];
        wf = (float)acc;
        m.raw[i] = wf;
        w = (double)wf;
    }
    const double t = group_sum_d<kG>(w, sh);
    if ((lt == 0) & (b < nb)) blocks[b] = t;
    return wf;
}

__global__ __launch_bounds__(kBlock) void k_reduce_m(const Member* __restrict__ members)
{
    const Member& m = members[blockIdx.y];
    __shared__ double sh[kRows];
    reduce_slice<1>(m, m.st->count, gridDim.x, blockIdx.x, sh, m.blocks);
}

// sum = total of the nb raw-weight partials (thread k of a group adds
// partials k, k + kBlock, ... in order, then the group sum; every group and
// every workgroup forms the same value), returned to every thread;
// published to stats[0] (and the status "no valid particle" when it is 0
// as a float) by thread 0 of a publishing workgroup.
template <int kG>
__device__ __forceinline__ float weight_total(const Member& m, uint32_t nb, bool publish, double* sh,
                                              double* total_sh /* kG */, const double* blocks)
{
    const uint32_t lt = threadIdx.x & (kBlock - 1);
    double t = 0.0;
    for (uint32_t k = lt; k < nb; k += kBlock) t += blocks[k];
    const double total = group_sum_d<kG>(t, sh);
    if (lt == 0) total_sh[threadIdx.x / kBlock] = total;
    if (threadIdx.x == 0 && publish) {
        m.st->stats[0] = total;
        m.st->status = ((float)total == 0.0f) ? kStNoValid : kStOk;
    }
    __syncthreads();
    return (float)total_sh[threadIdx.x / kBlock];
}

// The pose of particle i of slice b (zeros outside the member), read ahead
// of the weight so that its load overlaps other work.
struct SlicePose {
    float f[6];
};

__device__ __forceinline__ SlicePose slice_pose(const Member& m, uint32_t P, uint32_t nb, uint32_t b)
{
    const uint32_t i = b * kBlock + (threadIdx.x & (kBlock - 1));
    SlicePose sp;
#pragma unroll
    for (int k = 0; k < 6; ++k) sp.f[k] = 0.0f;
    if (b < nb && i < P) {
#pragma unroll
        for (int k = 0; k < 6; ++k) sp.f[k] = m.particles[i].first[k];
    }
    return sp;
}

// Slice b: w_i = raw_i / sum written to particles[i].second (the
// reference's particles[index].second /= weight_sum, float division) and
// over raw_i, and the slice's partials of the nine weighted pose sums
// (tsdf_evaluator.cpp:198-217) in double into blocks[nb + q nb + b].  raw_i
// and the pose come from the caller.
template <int kG>
__device__ __forceinline__ void normalize_slice(const Member& m, uint32_t P, uint32_t nb, uint32_t b, float sum,
                                                double* sh, double* blocks, float raw_i, const SlicePose& sp)
{
    const uint32_t lt = threadIdx.x & (kBlock - 1);
    const uint32_t i = b * kBlock + lt;
    double q[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) q[k] = 0.0;
    if (b < nb && i < P && sum != 0.0f) {
        const float w = raw_i / sum;
        m.particles[i].second = w;
        m.raw[i] = w;  // a contiguous copy of the normalised weights for host readers
        const double wd = (double)w;
        q[0] = (double)sp.f[0] * wd;
        q[1] = (double)sp.f[1] * wd;
        q[2] = (double)sp.f[2] * wd;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float s, c;
            sincosf_mode(m.trig, sp.f[3 + a], &s, &c);
            q[3 + 2 * a] = (double)s * wd;
            q[4 + 2 * a] = (double)c * wd;
        }
    }
    const double t = group_sum9<kG>(q, sh);
    if ((lt < 9) & (b < nb)) blocks[nb + (size_t)lt * nb + b] = t;
}

// The nine pose sums from nb slice partials into stats[1..9] (every group
// forms them; group 0 stores).
template <int kG>
__device__ __forceinline__ void pose_sums(const Member& m, uint32_t nb, double* sh, const double* blocks)
{
    const uint32_t lt = threadIdx.x & (kBlock - 1);
    double t[9];
#pragma unroll
    for (int qq = 0; qq < 9; ++qq) t[qq] = 0.0;
    const double* part = blocks + nb;
    for (uint32_t k = lt; k < nb; k += kBlock) {
        double v[9];
#pragma unroll
        for (int qq = 0; qq < 9; ++qq) v[qq] = part[(size_t)qq * nb + k];
#pragma unroll
        for (int qq = 0; qq < 9; ++qq) t[qq] += v[qq];
    }
    const double total = group_sum9<kG>(t, sh);
    if (threadIdx.x < 9) m.st->stats[1 + threadIdx.x] = total;
}

// The reductions avoid a "last workgroup finishes" step: on this GPU the
// device-scope fence it needs writes back and invalidates the L2 of the
// XCD, once per wavefront.  Instead every workgroup that needs a total
// re-forms it from the slice partials in one fixed order, and a small final
// kernel forms the pose sums.  A lone filter of at most
// kFinishSmallSlices slices runs all three steps in k_finish_small, four
// slices at a time, with the same functions and therefore the same bits.
__global__ __launch_bounds__(kBlock) void k_normalize_m(const Member* __restrict__ members)
{
    const Member& m = members[blockIdx.y];
    __shared__ double sh[kRows * 9];
    __shared__ double total_sh[1];
    const uint32_t nb = gridDim.x;
    const uint32_t P = m.st->count;
    const uint32_t b = blockIdx.x, i = b * kBlock + threadIdx.x;
    const SlicePose sp = slice_pose(m, P, nb, b);
    const float raw_i = i < P ? m.raw[i] : 0.0f;
    const float sum = weight_total<1>(m, nb, blockIdx.x == 0, sh, total_sh, m.blocks);
    normalize_slice<1>(m, P, nb, b, sum, sh, m.blocks, raw_i, sp);
}

__global__ __launch_bounds__(kBlock) void k_pose_sums_m(const Member* __restrict__ members, uint32_t nb)
{
    __shared__ double sh[kRows * 9];
    pose_sums<1>(members[blockIdx.y], nb, sh, members[blockIdx.y].blocks);
}

constexpr int kFinishGroups = 4;
constexpr int kFinishRounds = (kFinishSmallSlices + kFinishGroups - 1) / kFinishGroups;
__global__ __launch_bounds__(kBlock * kFinishGroups) void k_finish_small(const Member* __restrict__ members,
                                                                           uint32_t nb)
{
    const Member& m = members[blockIdx.y];
    __shared__ double sh[kRows * kFinishGroups * 9];
    __shared__ double total_sh[kFinishGroups];
    __shared__ double blocks[kFinishSmallSlices * 10];  // the slice partials, in LDS
    const uint32_t P = m.st->count;
    const uint32_t g = threadIdx.x / kBlock;
    // every slice this thread normalises: pose read ahead, raw weight kept
    SlicePose sp[kFinishRounds];
    float raw[kFinishRounds];
#pragma unroll
    for (int r = 0; r < kFinishRounds; ++r) sp[r] = slice_pose(m, P, nb, r * kFinishGroups + g);
#pragma unroll
    for (int r = 0; r < kFinishRounds; ++r) {
        raw[r] = 0.0f;
        if ((uint32_t)(r * kFinishGroups) < nb) raw[r] = reduce_slice<kFinishGroups>(m, P, nb, r * kFinishGroups + g, sh, blocks);
    }
    __syncthreads();  // the slice partials are written
    const float sum = weight_total<kFinishGroups>(m, nb, true, sh, total_sh, blocks);
#pragma unroll
    for (int r = 0; r < kFinishRounds; ++r)
        if ((uint32_t)(r * kFinishGroups) < nb)
            normalize_slice<kFinishGroups>(m, P, nb, r * kFinishGroups + g, sum, sh, blocks, raw[r], sp[r]);
    __syncthreads();  // the pose partials are written
    pose_sums<kFinishGroups>(m, nb, sh, blocks);
}

// Restore every member from its snapshot.
__global__ void k_restore_m(const Member* __restrict__ members)
{
    const Member& m = members[blockIdx.y];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m.snapshot_count) m.particles[i] = m.snapshot[i];
    if (i == 0) {
        m.st->count = m.snapshot_count;
        m.st->status = kStOk;
    }
}

// Work plan of a batched evaluation: member f owns items [prefix[f],
// prefix[f+1]), groups x chunks of them in (particle group, chunk) order;
// the item counter starts at zero.
__global__ __launch_bounds__(1024) void k_batch_plan(const Member* __restrict__ members, uint32_t n, uint32_t K,
                                                     BatchPlan* __restrict__ plan)
{
    __shared__ uint32_t wsum[16];
    const uint32_t f = threadIdx.x;
    uint32_t items = 0;
    if (f < n) {
        const Member& m = members[f];
        const uint32_t per = kWaves * K;
        items = m.N ? (m.st->count + per - 1) / per * m.chunks : 0u;
    }
    // block exclusive scan
    const uint32_t lane = f & 63, wave = f >> 6;
    uint32_t incl = items;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off, 64);
        if (lane >= (uint32_t)off) incl += o;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (uint32_t w = 0; w < wave; ++w) woff += wsum[w];
    const uint32_t excl = woff + incl - items;
    if (f < n) plan->prefix[f] = excl;
    if (f == n - 1) {
        plan->prefix[n] = excl + items;
        plan->total = excl + items;
        plan->n = n;
    }
    if (f < kPlanXcds) plan->next[f * 32] = 0;
}

// Copy the device count into the filter state (count comes from the host).
__global__ void k_set_count(FilterState* st, uint32_t n)
{
    st->count = n;
    st->status = kStOk;
}

// Restore the resident set from the snapshot and reset the count.
__global__ void k_restore(const tsl_particle* __restrict__ src, tsl_particle* __restrict__ dst, uint32_t n,
                          FilterState* st)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
    if (i == 0) {
        st->count = n;
        st->status = kStOk;
    }
}

// Tile shape: K particles per wavefront (template instantiations 1, 2)
// and the scan chunk a wavefront walks.  Automatic choice: K = 2 (the
// point load, tile walk and tail handling are shared by two particles; the
// kernel is bound by VALU issue), block offsets from the LDS rank table when
// the map has one, spatial particle order from 8192 particles, and enough
// point chunks to give the 256 CUs at least four full rounds of wavefronts.
}  // namespace

Tile choose_tile(const tsl_filter* f, size_t P, size_t N)
{
    Tile t;
    t.K = f->tile_k ? f->tile_k : 2;
    t.P = f->tile_p ? f->tile_p : 1;
    t.order = f->order_mode >= 0 ? f->order_mode : (P >= 8192 ? 1 : 0);
    t.rank = f->rank_mode >= 0 ? f->rank_mode : 1;
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

uint32_t batch_chunk(size_t N)
{
    // 2048-point chunks: items of kWaves*K particles x 2048 points keep the
    // workgroups of an XCD on neighbouring particle groups and scan chunks
    // while they are claimed dynamically (measured best on the batched
    // filters workload against 8192 points and whole scans)
    const uint64_t len = std::min<uint64_t>(2048, (N + 63) / 64 * 64);
    return (uint32_t)(len ? len : 64);
}

namespace {

// Launch the evaluation kernel for a tile shape.  The trace modes always use
// one point per lane per iteration (the lookup code is the same).
template <int kMode>
int launch_eval_kernel(const Tile& tile, const tsl_filter* f, size_t P, size_t N, hipStream_t stream,
                       float* partial, void* trace_out)
{
    const tsl_map* map = f->ev->map;
    const size_t groups = (P + tile.K - 1) / tile.K;
    const size_t items = (groups + kWaves - 1) / kWaves * tile.chunks;
    if (items > 0x7fffffffull) return set_error(TSL_ERR_UNSUPPORTED, "evaluation grid too large");
    const dim3 grid((unsigned)items);
    const bool rank = tile.rank && map->dev.occ_nwords != 0;
    const size_t lds = lds_bytes(map->dev, rank);
    const bool u24 = map->dev.up_dim_x < (1u << 23) && map->dev.up_dim_2 < (1u << 23) &&
                     map->dev.sub_dim_2 < (1 << 23);
    const bool brick = map->dev.brick != 0;
    const int pp = kMode == 0 ? tile.P : 1;
    const int sel = tile.K * 8 + pp;
#define TSL_EVAL_FLAGS kMode, u24, brick, rank, grid, lds, stream, f, tile, N, partial, trace_out
    switch (sel) {
    case 9: return launch_eval_tile_1_1(TSL_EVAL_FLAGS);
    case 17: return launch_eval_tile_2_1(TSL_EVAL_FLAGS);
    default: break;
    }
    if (kMode == 0) {
        switch (sel) {
        case 10: return launch_eval_tile_1_2(TSL_EVAL_FLAGS);
        case 12: return launch_eval_tile_1_4(TSL_EVAL_FLAGS);
        case 18: return launch_eval_tile_2_2(TSL_EVAL_FLAGS);
        case 20: return launch_eval_tile_2_4(TSL_EVAL_FLAGS);
        default: break;
        }
    }
#undef TSL_EVAL_FLAGS
    return set_error(TSL_ERR_INVALID_ARGUMENT, "tile must have K in {1, 2} and points per lane in {1, 2, 4} (got %d, %d)",
                     tile.K, tile.P);
}

}  // namespace

namespace {
__global__ void k_trig(const float* __restrict__ x, uint32_t n, int mode, float* __restrict__ s,
                       float* __restrict__ c)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float sv, cv;
    sincosf_mode(mode, x[i], &sv, &cv);
    s[i] = sv;
    c[i] = cv;
}
}  // namespace

int launch_trig(const float* d_x, uint32_t n, int mode, float* d_s, float* d_c, hipStream_t stream)
{
    if (n == 0) return TSL_OK;
    hipLaunchKernelGGL(k_trig, dim3((n + 255) / 256), dim3(256), 0, stream, d_x, n, mode, d_s, d_c);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

int launch_pose_matrices(const tsl_particle* d_particles, uint32_t n, const float tf[16], int trig, float* d_out,
                         hipStream_t stream)
{
    if (n == 0) return TSL_OK;
    Tf12 t;
    for (int k = 0; k < 16; ++k) t.m[k] = tf[k];
    hipLaunchKernelGGL(k_pose, dim3((n + 255) / 256), dim3(256), 0, stream, d_particles, nullptr, n, t, trig,
                       d_out);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

// Pose matrices for an evaluation: caller order, or spatial order with the
// inverse map for the reduction.
static int launch_pose_for_tile(tsl_filter* f, const Tile& tile, hipStream_t stream)
{
    const size_t P = f->count_bound;
    FilterState* st = f->state.as<FilterState>();
    Tf12 t;
    for (int k = 0; k < 16; ++k) t.m[k] = f->tf[k];
    const unsigned g = (unsigned)((P + 255) / 256);
    if (tile.order) {
        TSL_TRY(launch_particle_order(f, stream));
        hipLaunchKernelGGL(k_pose_ordered, dim3(g), dim3(256), 0, stream, f->particles.as<tsl_particle>(),
                           &st->count, f->porder.as<uint32_t>(), t, f->ev->trig, f->L.as<float>(),
                           f->pinv.as<uint32_t>());
    } else {
        hipLaunchKernelGGL(k_pose, dim3(g), dim3(256), 0, stream, f->particles.as<tsl_particle>(), &st->count,
                           (uint32_t)P, t, f->ev->trig, f->L.as<float>());
    }
    TSL_HIP_CHECK(hipGetLastError());
    f->last_order = tile.order;
    return TSL_OK;
}

int launch_pose_spatial(tsl_filter* f, hipStream_t stream)
{
    Tile t{};
    t.order = 1;
    return launch_pose_for_tile(f, t, stream);
}

static int launch_eval_core(tsl_filter* f, hipStream_t stream, Tile& tile)
{
    const size_t P = f->count_bound;
    const size_t N = f->num_points;
    tile = choose_tile(f, P, N);
    TSL_TRY(launch_pose_for_tile(f, tile, stream));
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
// the fastest.  Candidates: particle order {caller, spatial} x block lookup
// x K in {1, 2} x points per lane in {1, 2, 4} x {automatic chunks, 1/2/4/8
// chunks of the scan}.  A candidate's time is its evaluation kernel plus the
// per-update cost of its particle order (pose matrices, and for the spatial
// order the key computation and sort).
int autotune_tile(tsl_filter* f, hipStream_t stream, int reps, int* best_order, int* best_k, int* best_p,
                  uint32_t* best_chunk, float* best_ms)
{
    const size_t P = f->count_bound;
    const size_t N = f->num_points;
    if (P == 0 || N == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "autotune needs particles and a scan");
    const int saved_k = f->tile_k, saved_p = f->tile_p, saved_order = f->order_mode, saved_rank = f->rank_mode;
    const uint32_t saved_chunk = f->tile_chunk;
    hipEvent_t a, b;
    TSL_HIP_CHECK(hipEventCreate(&a));
    TSL_HIP_CHECK(hipEventCreate(&b));
    float best = 1e30f;
    int bo = 0, br = 0, bk = 1, bp = 1;
    uint32_t bc = 0;
    const int ks[2] = {1, 2};
    const int ps[3] = {1, 2, 4};
    const uint32_t div[5] = {0, 1, 2, 4, 8};
    int rc = TSL_OK;
    const int ranks = f->ev->map->dev.occ_nwords ? 2 : 1;
    for (int order = 0; order < 2 && rc == TSL_OK; ++order) {
        f->order_mode = order;
        Tile pt = choose_tile(f, P, N);
        rc = launch_pose_for_tile(f, pt, stream);  // warm-up
        float pose_ms = 0.0f;
        if (rc == TSL_OK) {
            (void)hipEventRecord(a, stream);
            for (int r = 0; r < reps && rc == TSL_OK; ++r) rc = launch_pose_for_tile(f, pt, stream);
            (void)hipEventRecord(b, stream);
            (void)hipEventSynchronize(b);
            (void)hipEventElapsedTime(&pose_ms, a, b);
            pose_ms /= (float)(reps > 0 ? reps : 1);
        }
        for (int rk = 0; rk < ranks && rc == TSL_OK; ++rk) {
        f->rank_mode = rk;
        for (int ki = 0; ki < 2 && rc == TSL_OK; ++ki) {
            for (int pi = 0; pi < 3 && rc == TSL_OK; ++pi) {
                for (int ci = 0; ci < 5 && rc == TSL_OK; ++ci) {
                    f->tile_k = ks[ki];
                    f->tile_p = ps[pi];
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
                    ms = ms / (float)(reps > 0 ? reps : 1) + pose_ms;
                    if (ms < best) {
                        best = ms;
                        bo = order;
                        br = rk;
                        bk = ks[ki];
                        bp = ps[pi];
                        bc = f->tile_chunk;
                    }
                }
            }
        }
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc != TSL_OK) {
        f->tile_k = saved_k;
        f->tile_p = saved_p;
        f->tile_chunk = saved_chunk;
        f->order_mode = saved_order;
        f->rank_mode = saved_rank;
        return rc;
    }
    f->order_mode = bo;
    f->rank_mode = br;
    f->tile_k = bk;
    f->tile_p = bp;
    f->tile_chunk = bc;
    if (best_order) *best_order = bo;
    if (best_k) *best_k = bk;
    if (best_p) *best_p = bp;
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
    const unsigned nb = (unsigned)((P + kBlock - 1) / kBlock);
    TSL_TRY(f->blocks.reserve(sizeof(double) * kReducePartials * (size_t)nb));
    Member* dm;
    TSL_TRY(filter_member(f, &dm, tile.chunk_len, tile.chunks, stream));
    if (nb <= kFinishSmallSlices) {
        hipLaunchKernelGGL(k_finish_small, dim3(1), dim3(kBlock * kFinishGroups), 0, stream, dm, nb);
        TSL_HIP_CHECK(hipGetLastError());
        return TSL_OK;
    }
    hipLaunchKernelGGL(k_reduce_m, dim3(nb), dim3(kBlock), 0, stream, dm);
    TSL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_normalize_m, dim3(nb), dim3(kBlock), 0, stream, dm);
    TSL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_pose_sums_m, dim3(1), dim3(kBlock), 0, stream, dm, nb);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

int launch_raw(tsl_filter* f, hipStream_t stream)
{
    const size_t P = f->count_bound;
    if (P == 0) return TSL_OK;
    Tile tile;
    TSL_TRY(launch_eval_core(f, stream, tile));
    const unsigned nb = (unsigned)((P + kBlock - 1) / kBlock);
    TSL_TRY(f->blocks.reserve(sizeof(double) * kReducePartials * (size_t)nb));
    Member* dm;
    TSL_TRY(filter_member(f, &dm, tile.chunk_len, tile.chunks, stream));
    hipLaunchKernelGGL(k_reduce_m, dim3(nb), dim3(kBlock), 0, stream, dm);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

int launch_trace(tsl_filter* f, hipStream_t stream, void* out_device, int values)
{
    const size_t P = f->count_bound;
    const size_t N = f->num_points;
    if (P == 0 || N == 0) return TSL_OK;
    TSL_TRY(launch_pose_matrices(f->particles.as<tsl_particle>(), (uint32_t)P, f->tf, f->ev->trig,
                                 f->L.as<float>(), stream));
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

namespace {
__global__ void k_store_member(Member m, Member* __restrict__ dst) { *dst = m; }
}  // namespace

int launch_store_member(const Member& m, Member* dst, hipStream_t stream)
{
    hipLaunchKernelGGL(k_store_member, dim3(1), dim3(1), 0, stream, m, dst);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

int launch_set_count(tsl_filter* f, uint32_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, stream, f->state.as<FilterState>(), n);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

// ---------------------------------------------------------------------------
// Batched update: every member's pose matrices, the item plan, one
// persistent evaluation launch over all members, reduction and
// normalisation, each a single launch for the whole batch.
// ---------------------------------------------------------------------------
namespace {

int dispatch_eval_batch(const BatchLaunch& b, hipStream_t stream, uint32_t* grid_out)
{
    switch (b.K * 8 + b.P) {
    case 9: return launch_eval_batch_tile_1_1(b, stream, grid_out);
    case 10: return launch_eval_batch_tile_1_2(b, stream, grid_out);
    case 12: return launch_eval_batch_tile_1_4(b, stream, grid_out);
    case 17: return launch_eval_batch_tile_2_1(b, stream, grid_out);
    case 18: return launch_eval_batch_tile_2_2(b, stream, grid_out);
    case 20: return launch_eval_batch_tile_2_4(b, stream, grid_out);
    default: return set_error(TSL_ERR_INVALID_ARGUMENT, "batch tile K=%d points per lane=%d unsupported", b.K, b.P);
    }
}

}  // namespace

int batch_grid(const BatchLaunch& b, uint32_t* grid) { return dispatch_eval_batch(b, nullptr, grid); }

int launch_batch_restore(const BatchLaunch& b, hipStream_t stream)
{
    const unsigned gx = (b.max_cap + 255) / 256 ? (b.max_cap + 255) / 256 : 1;
    hipLaunchKernelGGL(k_restore_m, dim3(gx, b.n), dim3(256), 0, stream, b.members);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

int launch_batch_update(const BatchLaunch& b, hipStream_t stream)
{
    if (b.n == 0) return TSL_OK;
    if (b.n > kMaxMembers) return set_error(TSL_ERR_UNSUPPORTED, "at most %u filters per batch", kMaxMembers);
    const unsigned gx = (b.max_cap + 255) / 256 ? (b.max_cap + 255) / 256 : 1;
    hipLaunchKernelGGL(k_pose_m, dim3(gx, b.n), dim3(256), 0, stream, b.members);
    TSL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_batch_plan, dim3(1), dim3(1024), 0, stream, b.members, b.n, (uint32_t)b.K, b.plan);
    TSL_HIP_CHECK(hipGetLastError());
    if (b.grid == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "batch grid not sized");
    TSL_TRY(dispatch_eval_batch(b, stream, nullptr));
    hipLaunchKernelGGL(k_reduce_m, dim3(gx, b.n), dim3(kBlock), 0, stream, b.members);
    TSL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_normalize_m, dim3(gx, b.n), dim3(kBlock), 0, stream, b.members);
    TSL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_pose_sums_m, dim3(1, b.n), dim3(kBlock), 0, stream, b.members, gx);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

}  // namespace tsl
