// Evaluation kernels of the sensor update and their launchers, shared by
// eval.hip (dispatch, reductions, pose matrices) and eval_tiles.hip, which
// is compiled once per tile shape (particles per wavefront x points per
// lane) so that the many instantiations build in parallel.  Everything here
// has internal linkage: each instantiation lives in the one translation unit
// that launches it.
//
// Work decomposition: a wavefront owns K particles and walks a chunk of the
// scan with its 64 lanes on consecutive points, so the K pose matrices are
// wavefront-uniform (scalar registers) and neighbouring lanes transform
// neighbouring scan points, which land in neighbouring voxels: the block
// table and voxel gathers of a wavefront hit few cache lines.  The voxel
// value palette sits in LDS.
#pragma once

#include "tsl_lookup.h"

#include <type_traits>

namespace tsl {

namespace {

constexpr int kBlock = 256;           // threads per workgroup (4 wavefronts)
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kXcds = 8;         // accelerator complex dies of the MI355X (one L2 each)
#ifndef TSL_TILE_POINTS
#define TSL_TILE_POINTS 512
#endif
constexpr uint32_t kTile = TSL_TILE_POINTS;  // scan points staged in LDS per step (8 KB), a multiple of 64 * 4
static_assert(kTile % 256 == 0, "the point tile holds whole iterations of 64 lanes x 4 points");
constexpr uint32_t kFinishSmallSlices = 8;  // filters of up to 8 * kBlock particles finish in one workgroup

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

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
    uint64_t* neg = reinterpret_cast<uint64_t*>(occ + (kRank ? map.occ_used : 0u));
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
        for (uint32_t i = threadIdx.x; i < map.occ_used; i += blockDim.x) dst[i] = src[i];
    }
    if (threadIdx.x < 8) neg[threadIdx.x] = map.neg_tab[threadIdx.x];
    for (uint32_t i = threadIdx.x; i < map.sub_entries; i += blockDim.x) tab[i] = map.sub_tab[i];
    for (uint32_t i = threadIdx.x; i < map.palette_size; i += blockDim.x)
        lut[i] = scaled ? a_hit * map.palette[i] : map.palette[i];
    __syncthreads();
    return LdsTables{atab, occ, neg, tab, lut, tile, scratch};
}

// LDS bytes stage_tables needs for a map.  The rank table's used words are
// staged; the allocation still reaches the end of the whole power-of-two
// table, so that the word a discarded lane addresses lies inside it.
inline size_t lds_bytes(const DevMap& d, bool rank)
{
    const size_t head = 16 * ((d.palette_size + 3) / 4) + sizeof(AxisEntry) * d.axis_n;
    const size_t tables = head + (rank ? sizeof(OccWord) * d.occ_used : 0) + 8 * sizeof(uint64_t) +
                          sizeof(SubEntry) * d.sub_entries;
    const size_t total = (tables + 15) / 16 * 16 + sizeof(float4) * kTile + 16;
    const size_t reach = rank ? head + sizeof(OccWord) * d.occ_nwords : 0;
    return total > reach ? total : reach;
}

// Instances held to eight wavefronts per SIMD: the compiler then fits them
// in 80 scalar registers (parking a few wavefront-uniform values in
// vector-register lanes), and with only the rank table's used words in LDS
// eight workgroups fit on a CU; left to itself it takes 106 scalar
// registers, which allow seven.  On the global-localisation workload:
// k_eval 4.12 -> 3.87 ms (1 particle x 2 points per wavefront iteration),
// 4.08 -> 3.86 ms (2 particles x 1 point, the second particle's matrix in
// vector registers).  Wider tiles, the 64-bit form with two particles and
// the float form spill to memory under that limit and keep the compiler's
// choice.
template <int K, int kP, bool kWide, bool kF32>
constexpr bool kEightWaves = !kF32 && (K == 1 ? kP <= 2 : (K == 2 && kP == 1 && !kWide));

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
        // wavefront-uniform: held in scalar registers, but in the
        // eight-wavefront instances the second particle's in vector
        // registers (the empty asm hides the uniformity), which keeps the
        // scalar registers within that budget
#pragma unroll
        for (int e = 0; e < 12; ++e) {
            float v = L[(size_t)p * 12 + e];
            if (k == 1 && kEightWaves<K, kP, kWide, kF32>)
                asm volatile("" : "+v"(v));
            else
                v = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
            m[k][e] = v;
        }
    }
    float acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0f;

    // The workgroup's waves walk the same chunk: its points are staged in
    // LDS one tile at a time (every wave helps and waits, active or not).
    // tb, te, base and active are wavefront-uniform; readfirstlane says so to
    // the compiler, which then keeps the point loop's control in scalar
    // registers instead of running it as a divergent loop on the VALU
    for (uint32_t tb = j0; tb < j1; tb += kTile) {
    const uint32_t te = __builtin_amdgcn_readfirstlane(min(j1, tb + kTile));
    __syncthreads();  // the previous tile is consumed
    for (uint32_t i = threadIdx.x; i < te - tb; i += blockDim.x) tile[i] = pts[tb + i];
    __syncthreads();
    if (!__builtin_amdgcn_readfirstlane(active ? 1u : 0u)) continue;
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
    uint32_t base = __builtin_amdgcn_readfirstlane(tb);
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

#define TSL_EVAL_PARAMS                                                                                    \
    DevMap map, const float* __restrict__ L, const uint32_t* __restrict__ count,                           \
        const float4* __restrict__ pts, uint32_t N, uint32_t chunk_len, uint32_t chunks, float a_hit,      \
        float* __restrict__ partial, uint32_t stride, void* __restrict__ trace_out,                        \
        const uint32_t* __restrict__ perm
#define TSL_EVAL_PARAM_NAMES map, L, count, pts, N, chunk_len, chunks, a_hit, partial, stride, trace_out, perm

template <int K, int kP, int kMode, bool kU24, bool kBrick, bool kRank, bool kWide, bool kF32>
__device__ __forceinline__ void eval_items(TSL_EVAL_PARAMS)
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


template <int K, int kP, int kMode, bool kU24, bool kBrick, bool kRank, bool kWide = false, bool kF32 = false>
__global__ __launch_bounds__(kBlock) void k_eval(TSL_EVAL_PARAMS)
{
    eval_items<K, kP, kMode, kU24, kBrick, kRank, kWide, kF32>(TSL_EVAL_PARAM_NAMES);
}

template <int K, int kP, int kMode, bool kU24, bool kBrick, bool kRank, bool kWide = false, bool kF32 = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_eval8(TSL_EVAL_PARAMS)
{
    eval_items<K, kP, kMode, kU24, kBrick, kRank, kWide, kF32>(TSL_EVAL_PARAM_NAMES);
}

// The evaluation kernel an instance launches.
template <int K, int kP, int kMode, bool kU24, bool kBrick, bool kRank, bool kWide, bool kF32>
constexpr auto eval_kernel()
{
    if constexpr (kEightWaves<K, kP, kWide, kF32>)
        return &k_eval8<K, kP, kMode, kU24, kBrick, kRank, kWide, kF32>;
    else
        return &k_eval<K, kP, kMode, kU24, kBrick, kRank, kWide, kF32>;
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
#define TSL_BATCH_PARAMS                                                                                   \
    DevMap map, const Member* __restrict__ members, BatchPlan* __restrict__ plan, float a_hit, uint32_t item0, \
        uint32_t count
#define TSL_BATCH_PARAM_NAMES map, members, plan, a_hit, item0, count

template <int K, int kP, bool kU24, bool kBrick, bool kRank, bool kWide, bool kF32>
__device__ __forceinline__ void batch_items(TSL_BATCH_PARAMS)
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


template <int K, int kP, bool kU24, bool kBrick, bool kRank, bool kWide = false, bool kF32 = false>
__global__ __launch_bounds__(kBlock) void k_eval_batch(TSL_BATCH_PARAMS)
{
    batch_items<K, kP, kU24, kBrick, kRank, kWide, kF32>(TSL_BATCH_PARAM_NAMES);
}

template <int K, int kP, bool kU24, bool kBrick, bool kRank, bool kWide = false, bool kF32 = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_eval_batch8(TSL_BATCH_PARAMS)
{
    batch_items<K, kP, kU24, kBrick, kRank, kWide, kF32>(TSL_BATCH_PARAM_NAMES);
}

// The batched evaluation kernel an instance launches.
template <int K, int kP, bool kU24, bool kBrick, bool kRank, bool kWide, bool kF32>
constexpr auto batch_kernel()
{
    if constexpr (kEightWaves<K, kP, kWide, kF32>)
        return &k_eval_batch8<K, kP, kU24, kBrick, kRank, kWide, kF32>;
    else
        return &k_eval_batch<K, kP, kU24, kBrick, kRank, kWide, kF32>;
}

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
    constexpr auto kernel = eval_kernel<K, kP, kMode, kU24, kBrick, kRank, kWide, kF32>();
    TSL_CHECK_PALETTE_AT_ZERO(kernel);
    const tsl_map* map = f->ev->map;
    FilterState* st = f->state.as<FilterState>();
    hipLaunchKernelGGL(kernel, grid, dim3(kBlock), lds, stream,
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

// stream == nullptr && grid_out: report the persistent grid size only.
__global__ void k_batch_claims_reset(BatchPlan* __restrict__ plan)
{
    if (threadIdx.x < kPlanXcds) plan->next[threadIdx.x * 32] = 0;
}

template <int K, int kP, bool kU24, bool kBrick, bool kRank, bool kWide = false, bool kF32 = false>
int launch_eval_batch_t(const BatchLaunch& b, hipStream_t stream, uint32_t* grid_out)
{
    constexpr auto kernel = batch_kernel<K, kP, kU24, kBrick, kRank, kWide, kF32>();
    TSL_CHECK_PALETTE_AT_ZERO(kernel);
    const size_t lds = lds_bytes(b.map->dev, kRank);
    if (grid_out) {
        int per_cu = 0;
        TSL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(kernel), kBlock,
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
        hipLaunchKernelGGL(kernel, dim3(b.grid), dim3(kBlock), lds,
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

}  // namespace

// Launchers of one tile shape, defined in eval_tiles.hip (one object per
// shape).  launch_eval_tile_K_P: mode 0 scores (any P), modes 1 and 2 trace
// voxel indices or values (P = 1 only).  launch_eval_batch_tile_K_P: the
// persistent batch kernel (grid_out != NULL: report its grid size only).
int launch_eval_tile_1_1(int mode, bool u24, bool brick, bool rank, dim3 grid, size_t lds, hipStream_t stream,
                         const tsl_filter* f, const Tile& tile, size_t N, float* partial, void* trace_out);
int launch_eval_tile_1_2(int mode, bool u24, bool brick, bool rank, dim3 grid, size_t lds, hipStream_t stream,
                         const tsl_filter* f, const Tile& tile, size_t N, float* partial, void* trace_out);
int launch_eval_tile_1_4(int mode, bool u24, bool brick, bool rank, dim3 grid, size_t lds, hipStream_t stream,
                         const tsl_filter* f, const Tile& tile, size_t N, float* partial, void* trace_out);
int launch_eval_tile_2_1(int mode, bool u24, bool brick, bool rank, dim3 grid, size_t lds, hipStream_t stream,
                         const tsl_filter* f, const Tile& tile, size_t N, float* partial, void* trace_out);
int launch_eval_tile_2_2(int mode, bool u24, bool brick, bool rank, dim3 grid, size_t lds, hipStream_t stream,
                         const tsl_filter* f, const Tile& tile, size_t N, float* partial, void* trace_out);
int launch_eval_tile_2_4(int mode, bool u24, bool brick, bool rank, dim3 grid, size_t lds, hipStream_t stream,
                         const tsl_filter* f, const Tile& tile, size_t N, float* partial, void* trace_out);
int launch_eval_batch_tile_1_1(const BatchLaunch& b, hipStream_t stream, uint32_t* grid_out);
int launch_eval_batch_tile_1_2(const BatchLaunch& b, hipStream_t stream, uint32_t* grid_out);
int launch_eval_batch_tile_1_4(const BatchLaunch& b, hipStream_t stream, uint32_t* grid_out);
int launch_eval_batch_tile_2_1(const BatchLaunch& b, hipStream_t stream, uint32_t* grid_out);
int launch_eval_batch_tile_2_2(const BatchLaunch& b, hipStream_t stream, uint32_t* grid_out);
int launch_eval_batch_tile_2_4(const BatchLaunch& b, hipStream_t stream, uint32_t* grid_out);

}  // namespace tsl
