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

#include <type_traits>

namespace tsl {

namespace {

constexpr int kBlock = 256;           // threads per workgroup (4 wavefronts)
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kXcds = 8;         // accelerator complex dies of the MI355X (one L2 each)
constexpr uint32_t kTile = 512;       // scan points staged in LDS per step (8 KB), a multiple of 64 * 4
constexpr uint32_t kFinishSmallSlices = 8;  // filters of up to 8 * kBlock particles finish in one workgroup

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
template <bool kU24, bool kRank, bool kWide>
__device__ __forceinline__ uint64_t slow_index(const DevMap& map, const SubEntry* __restrict__ tab,
                                               const uint64_t* __restrict__ neg, const OccWord* __restrict__ occ,
                                               float ox, float oy, float oz)
{
    if (wrap_ok(ox, oy, oz, map.range_bits)) return wrap_index<kU24, kRank, kWide>(map, tab, neg, occ, ox, oy, oz);
    return lookup_index_general(*map.full, ox, oy, oz);
}

// Pose matrices in the evaluation order: slot s holds particle order[s];
// inv[order[s]] = s lets the reduction read each particle's partial
// weights back in the caller's order.
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

// LDS copies of the per-map tables every lookup reads: occupancy-rank
// words, wrap-band table, sub-voxel thresholds, value palette.  The
// scoring kernels keep the palette pre-multiplied by a_hit (each entry is
// the float product a_hit * value the score would form per pair, so the
// sums are unchanged); the tracing kernels keep the raw values.
struct LdsTables {
    const AxisEntry* atab;
    const OccWord* occ;
    const uint64_t* neg;
    const SubEntry* tab;
    const float* lut;
    float4* tile;       // kTile staged scan points
    uint32_t* scratch;  // 4 words for workgroup-wide values
};

// Palette value at byte offset off (a voxel code).  The palette starts at
// LDS address 0 (the kernels declare no static LDS and stage it first), so
// the code is the LDS address itself.
__device__ __forceinline__ float palette_at(const float* lut, uint32_t off)
{
    (void)lut;
    return *reinterpret_cast<const __attribute__((address_space(3))) float*>((uintptr_t)off);
}

template <bool kRank>
__device__ __forceinline__ LdsTables stage_tables(const DevMap& map, bool scaled, float a_hit)
{
    extern __shared__ uint4 smem_raw[];
    // palette first (voxel codes are byte offsets from the start of the
    // dynamic LDS), padded to 16 bytes for the 16-byte axis entries
    float* lut = reinterpret_cast<float*>(smem_raw);
    AxisEntry* atab = reinterpret_cast<AxisEntry*>(smem_raw + (map.palette_size + 3) / 4);
    OccWord* occ = reinterpret_cast<OccWord*>(atab + map.axis_n);
    uint64_t* neg = reinterpret_cast<uint64_t*>(occ + (kRank ? map.occ_nwords : 0u));
    SubEntry* tab = reinterpret_cast<SubEntry*>(neg + 8);
    // the point tile and the workgroup words follow, 16-byte aligned (the
    // kernels declare no static LDS, so the palette sits at address 0)
    const uint32_t tables_end = (uint32_t)(reinterpret_cast<char*>(tab + map.sub_entries) -
                                           reinterpret_cast<char*>(smem_raw));
    float4* tile = reinterpret_cast<float4*>(reinterpret_cast<char*>(smem_raw) + ((tables_end + 15u) & ~15u));
    uint32_t* scratch = reinterpret_cast<uint32_t*>(tile + kTile);
    {
        const uint4* src = reinterpret_cast<const uint4*>(map.axis_tab);
        uint4* dst = reinterpret_cast<uint4*>(atab);
        for (uint32_t i = threadIdx.x; i < map.axis_n; i += blockDim.x) dst[i] = src[i];
    }
    if (kRank) {
        const uint2* src = reinterpret_cast<const uint2*>(map.occ_words);
        uint2* dst = reinterpret_cast<uint2*>(occ);
        for (uint32_t i = threadIdx.x; i < map.occ_nwords; i += blockDim.x) dst[i] = src[i];
    }
    if (threadIdx.x < 8) neg[threadIdx.x] = map.neg_tab[threadIdx.x];
    for (uint32_t i = threadIdx.x; i < map.sub_entries; i += blockDim.x) tab[i] = map.sub_tab[i];
    for (uint32_t i = threadIdx.x; i < map.palette_size; i += blockDim.x)
        lut[i] = scaled ? a_hit * map.palette[i] : map.palette[i];
    __syncthreads();
    return LdsTables{atab, occ, neg, tab, lut, tile, scratch};
}

// LDS bytes stage_tables needs for a map.
inline size_t lds_bytes(const DevMap& d, bool rank)
{
    const size_t tables = 16 * ((d.palette_size + 3) / 4) + sizeof(AxisEntry) * d.axis_n +
                          (rank ? sizeof(OccWord) * d.occ_nwords : 0) + 8 * sizeof(uint64_t) +
                          sizeof(SubEntry) * d.sub_entries;
    return (tables + 15) / 16 * 16 + sizeof(float4) * kTile + 16;
}

// One wavefront: K particles (pose matrices wavefront-uniform) x the points
// [j0, j1) of one chunk.  Each iteration covers 64*kP consecutive points,
// lane l taking points base + l + 64u (u < kP), so kP independent gather
// chains per particle are in flight per lane; lane l still adds its points in
// ascending order, so the sums do not depend on kP.  Each iteration picks,
// by wavefront-uniform ballot, the fast tier if every pair qualifies, else
// the exact wrap/general tiers per pair.  kMode 1 writes the voxel index of
// every pair, kMode 2 the voxel value, instead of accumulating the score,
// with the same tiling, tier selection and addressing as the production
// mode (kMode 0); trace_out is indexed by the caller's point order (perm).
// kMode 0 writes the chunk's partial weight of particle p0 + k to
// partial_row[p0 + k].  kWide: the map's block table holds slots and its
// voxel and byte offsets need 64 bits (bricked layout, no rank table).
// kF32: the float form (the voxel array holds the values themselves, read at
// twice the code byte offsets; no palette, no rank table).
template <int K, int kP, int kMode, bool kU24, bool kBrick, bool kRank, bool kWide = false, bool kF32 = false>
__device__ __forceinline__ void eval_wave(const DevMap& map, const LdsTables& t, const float* __restrict__ L,
                                          uint32_t P, uint32_t p0, bool active, const float4* __restrict__ pts,
                                          uint32_t N, uint32_t j0, uint32_t j1, float a_hit,
                                          float* __restrict__ partial_row, void* __restrict__ trace_out,
                                          const uint32_t* __restrict__ perm, float4* __restrict__ tile)
{
    constexpr bool kTrace = kMode != 0;
    const OccWord* occ = t.occ;
    const uint64_t* neg = t.neg;
    const SubEntry* tab = t.tab;
    const float* lut = t.lut;
    const uint32_t lane = threadIdx.x & 63;
    float m[K][12];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t p = active ? min(p0 + k, P - 1) : 0u;
        // wavefront-uniform: held in scalar registers
#pragma unroll
        for (int e = 0; e < 12; ++e)
            m[k][e] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(L[(size_t)p * 12 + e])));
    }
    float acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0f;

    // The workgroup's waves walk the same chunk: its points are staged in
    // LDS one tile at a time (every wave helps and waits, active or not).
    for (uint32_t tb = j0; tb < j1; tb += kTile) {
    const uint32_t te = min(j1, tb + kTile);
    __syncthreads();  // the previous tile is consumed
    for (uint32_t i = threadIdx.x; i < te - tb; i += blockDim.x) tile[i] = pts[tb + i];
    __syncthreads();
    if (!active) continue;
    // one iteration: 64*kP consecutive points from base; kFull: all of
    // them lie before te (no per-lane tail checks)
    auto step = [&](uint32_t base, auto full_tag) {
        constexpr bool kFull = decltype(full_tag)::value;
        float4 q[kP];
        bool act[kP];
#pragma unroll
        for (int u = 0; u < kP; ++u) {
            const uint32_t j = base + 64u * u + lane;
            act[u] = kFull || j < te;
            q[u] = tile[act[u] ? j - tb : 0u];
        }
        float ox[kP][K], oy[kP][K], oz[kP][K];
        bool fast = true;
#pragma unroll
        for (int u = 0; u < kP; ++u) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                // tsdf_evaluator.cpp:44-46: ((L0*x + L1*y) + L2*z) + L3, no FMA
                const float tx = m[k][0] * q[u].x + m[k][1] * q[u].y + m[k][2] * q[u].z + m[k][3];
                const float ty = m[k][4] * q[u].x + m[k][5] * q[u].y + m[k][6] * q[u].z + m[k][7];
                const float tz = m[k][8] * q[u].x + m[k][9] * q[u].y + m[k][10] * q[u].z + m[k][11];
                ox[u][k] = tx - map.min_x;
                oy[u][k] = ty - map.min_y;
                oz[u][k] = tz - map.min_z;
                fast = fast & fast_ok(ox[u][k], oy[u][k], oz[u][k], map.range_bits);
            }
        }
        if (__builtin_amdgcn_ballot_w64(!fast) == 0) {
            // byte offsets into the code array: 32 bits, or 64 for wide maps
            using Off = std::conditional_t<kWide, uint64_t, uint32_t>;
            float v[kP][K];
            uint64_t full[kP][K];
            Off addr[kP][K];
            bool ok[kP][K];
            Off sb[kP][K];
            bool blk[kP][K];
            bool any = false;
#pragma unroll
            for (int u = 0; u < kP; ++u) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if constexpr (kWide)
                        fast_block_wide<kU24>(map, ox[u][k], oy[u][k], oz[u][k], sb[u][k], blk[u][k]);
                    else
                        fast_block<kU24, kRank>(map, occ, ox[u][k], oy[u][k], oz[u][k], sb[u][k], blk[u][k]);
                    any = any | blk[u][k];
                }
            }
            // no pair of the iteration in an allocated block inside the grid:
            // every pair reads init_value (the sentinel code's value), and
            // the in-block part is skipped
            const bool none = !kTrace && __builtin_amdgcn_ballot_w64(any) == 0;
#pragma unroll
            for (int u = 0; u < kP; ++u) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    ok[u][k] = false;
                    addr[u][k] = 0u;
                    if (kTrace) full[u][k] = map.data_size;
                }
            }
            if (!none) {
#pragma unroll
            for (int u = 0; u < kP; ++u) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    uint32_t pc, kx, ky, kz;
                    fast_inblock(map, t.atab, ox[u][k], oy[u][k], oz[u][k], pc, kx, ky, kz);
                    const Off s = sb[u][k];
                    // byte address of the code; a lane whose reference index
                    // leaves its block row (carry: an axis poisoned its part)
                    // recomputes it from the linear index, and only such a
                    // lane can fall past data_size (a whole row lies inside
                    // the array)
                    Off a = s + pc;
                    bool valid = blk[u][k];
                    const bool carry = kWide ? pc >= kCarryPoison : a >= map.carry_limit;
                    if (kTrace || __builtin_expect(carry, 0)) {
                        const Off s_el = s >> 1;
                        const uint32_t small = kx + mul_u<kU24>(ky, (uint32_t)map.sub_dim) +
                                               mul_u<kU24>(kz, (uint32_t)map.sub_dim_2);
                        const Off sum = s_el + small;
                        if (carry) {
                            if constexpr (kWide) {
                                valid = valid & (sum < map.data_size);
                                a = phys_from_linear64(map, sum) << 1;
                            } else {
                                valid = valid & (sum < map.ds);
                                a = (kBrick ? phys_from_linear(map, sum) : sum) << 1;
                            }
                        }
                        if (kTrace) full[u][k] = blk[u][k] ? (uint64_t)s_el + small : map.data_size;
                    }
                    addr[u][k] = a;
                    ok[u][k] = valid;
                }
            }
            }
            if constexpr (kF32) {
                // float form: the value itself at twice the code's byte
                // offset; rejected pairs take init_value without a load
#pragma unroll
                for (int u = 0; u < kP; ++u) {
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        float val = map.init;
                        if (ok[u][k])
                            val = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(map.codes) +
                                                                  2 * (uint64_t)addr[u][k]);
                        v[u][k] = kTrace ? val : a_hit * val;
                    }
                }
            } else {
            // rejected pairs take the code of init_value without a load
            uint32_t code[kP][K];
#pragma unroll
            for (int u = 0; u < kP; ++u) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    code[u][k] = map.sentinel_code;
                    // byte offset from the array base (32-bit unless wide)
                    if (ok[u][k])
                        code[u][k] = *reinterpret_cast<const uint16_t*>(reinterpret_cast<const char*>(map.codes) +
                                                                        addr[u][k]);
                }
            }
#pragma unroll
            for (int u = 0; u < kP; ++u) {
#pragma unroll
                for (int k = 0; k < K; ++k) v[u][k] = palette_at(lut, code[u][k]);
            }
            }
#pragma unroll
            for (int u = 0; u < kP; ++u) {
                const uint32_t j = base + 64u * u + lane;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if (kMode == 1) {
                        if (act[u] && p0 + k < P)
                            static_cast<uint64_t*>(trace_out)[(size_t)(p0 + k) * N + perm[j]] = full[u][k];
                    } else if (kMode == 2) {
                        if (act[u] && p0 + k < P)
                            static_cast<float*>(trace_out)[(size_t)(p0 + k) * N + perm[j]] = v[u][k];
                    } else if (act[u]) {
                        acc[k] += v[u][k] + q[u].w;  // tsdf_evaluator.cpp:56-67 (v = a_hit * value)
                    }
                }
            }
        } else {
            // exact tiers on the same offsets; same per-lane summation order
#pragma unroll
            for (int u = 0; u < kP; ++u) {
                const uint32_t j = base + 64u * u + lane;
                const bool a = act[u];
                const float4 qq = q[u];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const uint64_t full =
                        slow_index<kU24, kRank, kWide>(map, tab, neg, occ, ox[u][k], oy[u][k], oz[u][k]);
                    float v = kTrace ? map.init : a_hit * map.init;
                    if (full < map.data_size) {
                        if constexpr (kWide) {
                            v = palette_at(lut, map.codes[phys_from_linear64(map, full)]);
                        } else {
                            const uint32_t x = (uint32_t)full;
                            const uint32_t e = kBrick ? phys_from_linear(map, x) : x;
                            if constexpr (kF32) {
                                const float val = reinterpret_cast<const float*>(map.codes)[e];
                                v = kTrace ? val : a_hit * val;
                            } else {
                                v = palette_at(lut, map.codes[e]);
                            }
                        }
                    }
                    if (kMode == 1) {
                        if (a && p0 + k < P) static_cast<uint64_t*>(trace_out)[(size_t)(p0 + k) * N + perm[j]] = full;
                    } else if (kMode == 2) {
                        if (a && p0 + k < P) static_cast<float*>(trace_out)[(size_t)(p0 + k) * N + perm[j]] = v;
                    } else if (a) {
                        acc[k] += v + qq.w;
                    }
                }
            }
        }
    };
    uint32_t base = tb;
    for (; base + 64u * kP <= te; base += 64u * kP) step(base, std::true_type{});
    if (base < te) step(base, std::false_type{});
    }
    if (kTrace || !active) return;
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = wave_sum(acc[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (p0 + k < P) partial_row[p0 + k] = acc[k];
    }
}

template <int K, int kP, int kMode, bool kU24, bool kBrick, bool kRank, bool kWide = false, bool kF32 = false>
__global__ __launch_bounds__(kBlock) void k_eval(DevMap map, const float* __restrict__ L,
                                                 const uint32_t* __restrict__ count,
                                                 const float4* __restrict__ pts, uint32_t N,
                                                 uint32_t chunk_len, uint32_t chunks, float a_hit,
                                                 float* __restrict__ partial, uint32_t stride,
                                                 void* __restrict__ trace_out, const uint32_t* __restrict__ perm)
{
    const LdsTables t = stage_tables<kRank>(map, kMode == 0, a_hit);
    const uint32_t P = *count;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // Workgroups are dealt to the 8 XCDs round-robin by index; give each XCD
    // a contiguous range of (particle group, chunk) items so that the
    // particles one XCD works on together are neighbours in the evaluation
    // order (spatially sorted particles then share L2 lines).
    const uint32_t items = gridDim.x;
    const uint32_t xcd = blockIdx.x % kXcds, r = blockIdx.x / kXcds;
    const uint32_t q = items / kXcds, rem = items % kXcds;
    const uint32_t item = xcd * q + min(xcd, rem) + r;
    const uint32_t g = item / chunks, c = item - g * chunks;
    const uint32_t p0 = (g * kWaves + wave) * K;
    if (g * kWaves * K >= P) return;  // the whole workgroup is idle (uniform across its waves)
    const uint32_t j0 = c * chunk_len;
    const uint32_t j1 = min(N, j0 + chunk_len);
    eval_wave<K, kP, kMode, kU24, kBrick, kRank, kWide, kF32>(map, t, L, P, p0, p0 < P, pts, N, j0, j1, a_hit,
                                                        kMode == 0 ? partial + (size_t)c * stride : nullptr,
                                                        trace_out, perm, t.tile);
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
// slice b.

// The slice functions below run as groups of kBlock threads: kG groups per
// workgroup (1 in the member kernels, 4 in k_finish_small), group g =
// threadIdx.x / kBlock working on its own slice; every thread of the
// workgroup calls them (their barriers are workgroup-wide).  A group sum is
// formed in one fixed order whatever kG is (the fixed butterfly of each
// wave, then the group's waves in order from 0.0), so a slice gives the
// same bits in either kernel.
template <int kG>
__device__ __forceinline__ double group_sum_d(double v, double* sh /* kWaves * kG */)
{
    v = wave_sum_d(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double r = 0.0;
    if ((threadIdx.x & (kBlock - 1)) == 0)
        for (int w = 0; w < kWaves; ++w) r += sh[wave + w];
    __syncthreads();
    return r;  // in the group's thread 0
}

template <int kG>
__device__ __forceinline__ double group_sum9(const double (&v)[9], double* sh /* kWaves * kG * 9 */)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const double w = wave_sum_d(v[k]);
        if (lane == 0) sh[wave * 9 + k] = w;
    }
    __syncthreads();
    double r = 0.0;
    const uint32_t lt = threadIdx.x & (kBlock - 1);
    if (lt < 9)
        for (int w = 0; w < kWaves; ++w) r += sh[(wave + w) * 9 + lt];
    __syncthreads();
    return r;  // sum k in the group's thread k < 9
}

// raw[p] = sum over chunks (ascending) of partial[c][p], in double, stored as
// float; the partial sum of raw in double over slice b (particles
// [b kBlock, (b+1) kBlock)) into blocks[b] (a slice past nb contributes and
// writes nothing).
template <int kG>
__device__ __forceinline__ void reduce_slice(const Member& m, uint32_t P, uint32_t nb, uint32_t b, double* sh)
{
    const uint32_t lt = threadIdx.x & (kBlock - 1);
    const uint32_t i = b * kBlock + lt;
    double w = 0.0;
    if (b < nb && i < P) {
        const uint32_t slot = m.inv ? m.inv[i] : i;
        double acc = 0.0;
        for (uint32_t c = 0; c < m.chunks; ++c) acc += (double)m.partial[(size_t)c * m.stride + slot];
        const float wf = (float)acc;
        m.raw[i] = wf;
        w = (double)wf;
    }
    const double t = group_sum_d<kG>(w, sh);
    if (lt == 0 && b < nb) m.blocks[b] = t;
}

__global__ __launch_bounds__(kBlock) void k_reduce_m(const Member* __restrict__ members)
{
    const Member& m = members[blockIdx.y];
    __shared__ double sh[kWaves];
    reduce_slice<1>(m, m.st->count, gridDim.x, blockIdx.x, sh);
}

// sum = total of the nb raw-weight partials (thread k of a group adds
// partials k, k + kBlock, ... in order, then the group sum; every group and
// every workgroup forms the same value), returned to every thread;
// published to stats[0] (and the status "no valid particle" when it is 0
// as a float) by thread 0 of a publishing workgroup.
template <int kG>
__device__ __forceinline__ float weight_total(const Member& m, uint32_t nb, bool publish, double* sh,
                                              double* total_sh /* kG */)
{
    const uint32_t lt = threadIdx.x & (kBlock - 1);
    double t = 0.0;
    for (uint32_t k = lt; k < nb; k += kBlock) t += m.blocks[k];
    const double total = group_sum_d<kG>(t, sh);
    if (lt == 0) total_sh[threadIdx.x / kBlock] = total;
    if (threadIdx.x == 0 && publish) {
        m.st->stats[0] = total;
        m.st->status = ((float)total == 0.0f) ? kStNoValid : kStOk;
    }
    __syncthreads();
    return (float)total_sh[threadIdx.x / kBlock];
}

// Slice b: w_i = raw_i / sum written to particles[i].second (the
// reference's particles[index].second /= weight_sum, float division) and
// over raw_i, and
// the slice's partials of the nine weighted pose sums
// (tsdf_evaluator.cpp:198-217) in double into blocks[nb + q nb + b].
template <int kG>
__device__ __forceinline__ void normalize_slice(const Member& m, uint32_t P, uint32_t nb, uint32_t b, float sum,
                                                double* sh)
{
    const uint32_t lt = threadIdx.x & (kBlock - 1);
    const uint32_t i = b * kBlock + lt;
    double q[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) q[k] = 0.0;
    if (b < nb && i < P && sum != 0.0f) {
        tsl_particle* particles = m.particles;
        const float w = m.raw[i] / sum;
        particles[i].second = w;
        m.raw[i] = w;  // a contiguous copy of the normalised weights for host readers
        const double wd = (double)w;
        q[0] = (double)particles[i].first[0] * wd;
        q[1] = (double)particles[i].first[1] * wd;
        q[2] = (double)particles[i].first[2] * wd;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float s, c;
            sincosf_mode(m.trig, particles[i].first[3 + a], &s, &c);
            q[3 + 2 * a] = (double)s * wd;
            q[4 + 2 * a] = (double)c * wd;
        }
    }
    const double t = group_sum9<kG>(q, sh);
    if (lt < 9 && b < nb) m.blocks[nb + (size_t)lt * nb + b] = t;
}

// The nine pose sums from nb slice partials into stats[1..9] (every group
// forms them; group 0 stores).
template <int kG>
__device__ __forceinline__ void pose_sums(const Member& m, uint32_t nb, double* sh)
{
    const uint32_t lt = threadIdx.x & (kBlock - 1);
    double t[9];
#pragma unroll
    for (int qq = 0; qq < 9; ++qq) t[qq] = 0.0;
    const double* part = m.blocks + nb;
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
    __shared__ double sh[kWaves * 9];
    __shared__ double total_sh[1];
    const uint32_t nb = gridDim.x;
    const float sum = weight_total<1>(m, nb, blockIdx.x == 0, sh, total_sh);
    normalize_slice<1>(m, m.st->count, nb, blockIdx.x, sum, sh);
}

__global__ __launch_bounds__(kBlock) void k_pose_sums_m(const Member* __restrict__ members, uint32_t nb)
{
    __shared__ double sh[kWaves * 9];
    pose_sums<1>(members[blockIdx.y], nb, sh);
}

constexpr int kFinishGroups = 4;
__global__ __launch_bounds__(kBlock * kFinishGroups) void k_finish_small(const Member* __restrict__ members,
                                                                           uint32_t nb)
{
    const Member& m = members[blockIdx.y];
    __shared__ double sh[kWaves * kFinishGroups * 9];
    __shared__ double total_sh[kFinishGroups];
    const uint32_t P = m.st->count;
    const uint32_t g = threadIdx.x / kBlock;
    for (uint32_t b0 = 0; b0 < nb; b0 += kFinishGroups) reduce_slice<kFinishGroups>(m, P, nb, b0 + g, sh);
    __syncthreads();  // the slice partials are written
    const float sum = weight_total<kFinishGroups>(m, nb, true, sh, total_sh);
    for (uint32_t b0 = 0; b0 < nb; b0 += kFinishGroups) normalize_slice<kFinishGroups>(m, P, nb, b0 + g, sum, sh);
    __syncthreads();  // the pose partials are written
    pose_sums<kFinishGroups>(m, nb, sh);
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

// Batched evaluation: a persistent grid claims items one at a time, the
// workgroups of XCD x (dealt round-robin by index) from the items x, x + 8,
// ... through their XCD's counter, so all workgroups sweep the members
// together, as plain per-filter launches would, and the tail balances
// itself; per item the same wave body as k_eval.  One launch covers the
// slice [item0, item0 + count) of the plan's items (its claim counters
// start at zero), so a long evaluation can be cut into launches that other
// streams' kernels interleave with; every item writes only its own partial
// weights, so the slicing changes no bit.
template <int K, int kP, bool kU24, bool kBrick, bool kRank, bool kWide = false, bool kF32 = false>
__global__ __launch_bounds__(kBlock) void k_eval_batch(DevMap map, const Member* __restrict__ members,
                                                       BatchPlan* __restrict__ plan, float a_hit, uint32_t item0,
                                                       uint32_t count)
{
    const LdsTables t = stage_tables<kRank>(map, true, a_hit);
    float4* tile = t.tile;
    uint32_t* claim = t.scratch;
    const uint32_t n = plan->n;
    const uint32_t total = min(plan->total, item0 + min(count, 0xffffffffu - item0));
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t* counter = &plan->next[(blockIdx.x % kPlanXcds) * 32];
    const uint32_t xcd = blockIdx.x % kPlanXcds;
    for (;;) {
        if (threadIdx.x == 0) {
            const uint32_t item = item0 + atomicAdd(counter, 1u) * kPlanXcds + xcd;
            uint32_t lo = 0, hi = n;  // member: prefix[lo] <= item < prefix[lo + 1]
            if (item < total) {
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (plan->prefix[mid] <= item) lo = mid; else hi = mid;
                }
            }
            claim[0] = item;
            claim[1] = lo;
        }
        __syncthreads();
        const uint32_t item = __builtin_amdgcn_readfirstlane(claim[0]);
        const uint32_t f = __builtin_amdgcn_readfirstlane(claim[1]);
        __syncthreads();  // claim[] is read before thread 0 overwrites it
        if (item >= total) break;  // uniform across the workgroup
        const Member& m = members[f];
        const uint32_t local = item - plan->prefix[f];
        const uint32_t g = local / m.chunks, c = local - g * m.chunks;
        const uint32_t P = m.st->count;
        const uint32_t p0 = (g * kWaves + wave) * K;
        const uint32_t j0 = c * m.chunk_len;
        const uint32_t j1 = min(m.N, j0 + m.chunk_len);
        eval_wave<K, kP, 0, kU24, kBrick, kRank, kWide, kF32>(map, t, m.L, P, p0, p0 < P, m.points, m.N, j0, j1,
                                                              a_hit,
                                                        m.partial + (size_t)c * m.stride, nullptr, nullptr, tile);
    }
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

// The scoring kernels read the palette at LDS address 0 (palette_at): they
// must declare no static LDS, which would sit before the dynamic block.
// Checked once per kernel.
bool palette_at_zero(const void* kernel)
{
    hipFuncAttributes a{};
    return hipFuncGetAttributes(&a, kernel) == hipSuccess && a.sharedSizeBytes == 0;
}

#define TSL_CHECK_PALETTE_AT_ZERO(kernel)                                                              \
    do {                                                                                               \
        static const bool tsl_lds0_ = palette_at_zero(reinterpret_cast<const void*>(kernel));          \
        if (!tsl_lds0_) return set_error(TSL_ERR_UNSUPPORTED, "evaluation kernel declares static LDS"); \
    } while (0)

template <int K, int kP, int kMode, bool kU24, bool kBrick, bool kRank, bool kWide = false, bool kF32 = false>
int launch_eval_t(dim3 grid, size_t lds, hipStream_t stream, const tsl_filter* f, const Tile& tile, size_t N,
                  float* partial, void* trace_out)
{
    TSL_CHECK_PALETTE_AT_ZERO((&k_eval<K, kP, kMode, kU24, kBrick, kRank, kWide, kF32>));
    const tsl_map* map = f->ev->map;
    FilterState* st = f->state.as<FilterState>();
    hipLaunchKernelGGL((k_eval<K, kP, kMode, kU24, kBrick, kRank, kWide, kF32>), grid, dim3(kBlock), lds, stream,
                       map->dev,
                       f->L.as<float>(), &st->count, f->points.as<float4>(), (uint32_t)N, tile.chunk_len,
                       tile.chunks, f->ev->a_hit, partial, (uint32_t)f->capacity, trace_out,
                       f->perm.as<uint32_t>());
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

template <int K, int kP, int kMode>
int launch_eval_flags(bool u24, bool brick, bool rank, dim3 grid, size_t lds, hipStream_t stream,
                      const tsl_filter* f, const Tile& tile, size_t N, float* partial, void* trace_out)
{
#define TSL_EVAL_ARGS grid, lds, stream, f, tile, N, partial, trace_out
    if (f->ev->map->dev.wide) {  // bricked, slot table, no rank table
        if (u24) return launch_eval_t<K, kP, kMode, true, true, false, true>(TSL_EVAL_ARGS);
        return launch_eval_t<K, kP, kMode, false, true, false, true>(TSL_EVAL_ARGS);
    }
    if (f->ev->map->dev.f32) {  // float form: int32 block table, no rank table
        if (u24) return brick ? launch_eval_t<K, kP, kMode, true, true, false, false, true>(TSL_EVAL_ARGS)
                              : launch_eval_t<K, kP, kMode, true, false, false, false, true>(TSL_EVAL_ARGS);
        return brick ? launch_eval_t<K, kP, kMode, false, true, false, false, true>(TSL_EVAL_ARGS)
                     : launch_eval_t<K, kP, kMode, false, false, false, false, true>(TSL_EVAL_ARGS);
    }
    if (u24) {
        if (brick) {
            if (rank) return launch_eval_t<K, kP, kMode, true, true, true>(TSL_EVAL_ARGS);
            return launch_eval_t<K, kP, kMode, true, true, false>(TSL_EVAL_ARGS);
        }
        if (rank) return launch_eval_t<K, kP, kMode, true, false, true>(TSL_EVAL_ARGS);
        return launch_eval_t<K, kP, kMode, true, false, false>(TSL_EVAL_ARGS);
    }
    if (brick) {
        if (rank) return launch_eval_t<K, kP, kMode, false, true, true>(TSL_EVAL_ARGS);
        return launch_eval_t<K, kP, kMode, false, true, false>(TSL_EVAL_ARGS);
    }
    if (rank) return launch_eval_t<K, kP, kMode, false, false, true>(TSL_EVAL_ARGS);
    return launch_eval_t<K, kP, kMode, false, false, false>(TSL_EVAL_ARGS);
#undef TSL_EVAL_ARGS
}

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
#define TSL_EVAL_FLAGS u24, brick, rank, grid, lds, stream, f, tile, N, partial, trace_out
    switch (sel) {
    case 9: return launch_eval_flags<1, 1, kMode>(TSL_EVAL_FLAGS);
    case 17: return launch_eval_flags<2, 1, kMode>(TSL_EVAL_FLAGS);
    default: break;
    }
    if (kMode == 0) {
        switch (sel) {
        case 10: return launch_eval_flags<1, 2, 0>(TSL_EVAL_FLAGS);
        case 12: return launch_eval_flags<1, 4, 0>(TSL_EVAL_FLAGS);
        case 18: return launch_eval_flags<2, 2, 0>(TSL_EVAL_FLAGS);
        case 20: return launch_eval_flags<2, 4, 0>(TSL_EVAL_FLAGS);
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

// stream == nullptr && grid_out: report the persistent grid size only.
__global__ void k_batch_claims_reset(BatchPlan* __restrict__ plan)
{
    if (threadIdx.x < kPlanXcds) plan->next[threadIdx.x * 32] = 0;
}

template <int K, int kP, bool kU24, bool kBrick, bool kRank, bool kWide = false, bool kF32 = false>
int launch_eval_batch_t(const BatchLaunch& b, hipStream_t stream, uint32_t* grid_out)
{
    TSL_CHECK_PALETTE_AT_ZERO((&k_eval_batch<K, kP, kU24, kBrick, kRank, kWide, kF32>));
    const size_t lds = lds_bytes(b.map->dev, kRank);
    if (grid_out) {
        int per_cu = 0;
        TSL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&k_eval_batch<K, kP, kU24, kBrick, kRank, kWide, kF32>), kBlock,
            lds));
        if (per_cu < 1) per_cu = 1;
        int cus = 0;
        TSL_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b.map->device));
        uint64_t grid = (uint64_t)per_cu * (uint64_t)cus;
        const uint64_t need = (b.max_items + kXcds - 1) / kXcds * kXcds;
        if (grid > need) grid = need;
        grid = (grid + kXcds - 1) / kXcds * kXcds;
        if (grid < kXcds) grid = kXcds;
        *grid_out = (uint32_t)grid;
        return TSL_OK;
    }
    const uint64_t slice = b.slice_items ? b.slice_items : (uint64_t)0xffffffffu;
    const uint64_t end = b.slice_items ? std::min<uint64_t>(b.max_items, 0xffffffffull) : 1;
    for (uint64_t i0 = 0; i0 < end; i0 += slice) {
        if (i0) {
            hipLaunchKernelGGL(k_batch_claims_reset, dim3(1), dim3(64), 0, stream, b.plan);
            TSL_HIP_CHECK(hipGetLastError());
        }
        hipLaunchKernelGGL((k_eval_batch<K, kP, kU24, kBrick, kRank, kWide, kF32>), dim3(b.grid), dim3(kBlock), lds,
                           stream, b.map->dev, b.members, b.plan, b.a_hit, (uint32_t)i0,
                           (uint32_t)std::min<uint64_t>(slice, 0xffffffffu));
        TSL_HIP_CHECK(hipGetLastError());
    }
    return TSL_OK;
}

template <int K, int kP>
int launch_eval_batch_flags(const BatchLaunch& b, hipStream_t stream, uint32_t* grid_out)
{
    const DevMap& d = b.map->dev;
    const bool u24 = d.up_dim_x < (1u << 23) && d.up_dim_2 < (1u << 23) && d.sub_dim_2 < (1 << 23);
    const bool brick = d.brick != 0, rank = b.rank && d.occ_nwords != 0;
    if (d.wide)
        return u24 ? launch_eval_batch_t<K, kP, true, true, false, true>(b, stream, grid_out)
                   : launch_eval_batch_t<K, kP, false, true, false, true>(b, stream, grid_out);
    if (d.f32) {  // float form: int32 block table, no rank table
        if (u24) return brick ? launch_eval_batch_t<K, kP, true, true, false, false, true>(b, stream, grid_out)
                              : launch_eval_batch_t<K, kP, true, false, false, false, true>(b, stream, grid_out);
        return brick ? launch_eval_batch_t<K, kP, false, true, false, false, true>(b, stream, grid_out)
                     : launch_eval_batch_t<K, kP, false, false, false, false, true>(b, stream, grid_out);
    }
    if (u24) {
        if (brick) return rank ? launch_eval_batch_t<K, kP, true, true, true>(b, stream, grid_out)
                               : launch_eval_batch_t<K, kP, true, true, false>(b, stream, grid_out);
        return rank ? launch_eval_batch_t<K, kP, true, false, true>(b, stream, grid_out)
                    : launch_eval_batch_t<K, kP, true, false, false>(b, stream, grid_out);
    }
    if (brick) return rank ? launch_eval_batch_t<K, kP, false, true, true>(b, stream, grid_out)
                           : launch_eval_batch_t<K, kP, false, true, false>(b, stream, grid_out);
    return rank ? launch_eval_batch_t<K, kP, false, false, true>(b, stream, grid_out)
                : launch_eval_batch_t<K, kP, false, false, false>(b, stream, grid_out);
}

int dispatch_eval_batch(const BatchLaunch& b, hipStream_t stream, uint32_t* grid_out)
{
    switch (b.K * 8 + b.P) {
    case 9: return launch_eval_batch_flags<1, 1>(b, stream, grid_out);
    case 10: return launch_eval_batch_flags<1, 2>(b, stream, grid_out);
    case 12: return launch_eval_batch_flags<1, 4>(b, stream, grid_out);
    case 17: return launch_eval_batch_flags<2, 1>(b, stream, grid_out);
    case 18: return launch_eval_batch_flags<2, 2>(b, stream, grid_out);
    case 20: return launch_eval_batch_flags<2, 4>(b, stream, grid_out);
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
