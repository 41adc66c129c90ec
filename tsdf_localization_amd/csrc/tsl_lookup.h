// Device-side voxel lookup with the reference's exact index semantics.
//
// Reference: CudaSubVoxelMap::getIndex / getEntry,
// include/tsdf_localization/cuda/cuda_sub_voxel_map.tcc:51-158 (the lookup the
// CPU evaluator uses).  The index produced is the reference's linear voxel
// index (block offset + sx + sy*sub_dim + sz*sub_dim^2, modulo 2^64);
// data_size means "rejected, read init_value".
//
// Per axis the reference computes, with o = t - min (float):
//   g = (size_t)o;  grid check (size_t)((float)g / res) < dim;
//   u = (size_t)(o / 1.0f) (== g);  sub = o - (float)u;  s = (size_t)(sub / res)
// and then up = ux + uy*up_dim_x + uz*up_dim_2, offset = grid_occ[up],
// idx = offset + sx + sy*sub_dim + sz*sub_dim^2.  The kernels evaluate this
// in three exact tiers:
//   fast:    every offset in (-1, 2^23): 32-bit integers; the grid check is
//            an integer compare with a host-derived threshold; s comes from
//            the sub-voxel threshold table instead of a division.
//   wrap:    offsets in (-2^23, -1] on some axis: gcc's size_t conversion
//            wraps (u = 2^64 + trunc(o)); that axis's grid result and index
//            contribution are per-map constants; 64-bit index arithmetic.
//   general: |offset| >= 2^23 or NaN: every conversion spelled out with x86-64
//            semantics and IEEE division.
#pragma once

#include "tsl_internal.h"

namespace tsl {

// General tier: the reference's arithmetic step by step.
__device__ __forceinline__ uint64_t lookup_index_general(const DevMapFull& m, float ox, float oy, float oz)
{
    // :53-62 integer-metre bounds check
    const uint64_t gx = cvt_f32_u64(ox), gy = cvt_f32_u64(oy), gz = cvt_f32_u64(oz);
    const uint64_t qx = cvt_f32_u64(cvt_u64_f32(gx) / m.res);
    const uint64_t qy = cvt_f32_u64(cvt_u64_f32(gy) / m.res);
    const uint64_t qz = cvt_f32_u64(cvt_u64_f32(gz) / m.res);
    if (qx >= m.dim_x || qy >= m.dim_y || qz >= m.dim_z) return m.data_size;
    // :67-84: x_offset / 1.0f == x_offset, so the block coordinate equals gx
    const float spx = ox - cvt_u64_f32(gx) * 1.0f;
    const float spy = oy - cvt_u64_f32(gy) * 1.0f;
    const float spz = oz - cvt_u64_f32(gz) * 1.0f;
    const uint64_t up = gx + gy * m.up_dim_x + gz * m.up_dim_2;
    if (up >= m.grid_occ_size) return m.data_size;
    int64_t s;
    if (m.wide) {
        const uint32_t slot = m.slots[up];
        s = slot == kNoSlot ? -1 : (int64_t)((uint64_t)slot * m.sd3);
    } else {
        s = m.grid_occ[up];
    }
    if (s < 0) return m.data_size;
    const uint64_t sx = cvt_f32_u64(spx / m.res);
    const uint64_t sy = cvt_f32_u64(spy / m.res);
    const uint64_t sz = cvt_f32_u64(spz / m.res);
    return (uint64_t)s + sx + sy * m.sub_dim + sz * m.sub_dim_2;
}

// trunc(fl(a / res)) for a in [0, 1): bucket of a's bit pattern in the
// LDS-resident threshold table, one compare.
__device__ __forceinline__ uint32_t sub_index(uint32_t sub_base, uint32_t sub_shift, const SubEntry* __restrict__ tab,
                                              float a)
{
    const uint32_t key = __float_as_uint(a);
    const uint32_t b = (key > sub_base ? key - sub_base : 0u) >> sub_shift;
    const SubEntry e = tab[b];
    return e.count + (key >= e.thr ? 1u : 0u);
}

// 32-bit products of operands known to fit in 24 bits (full-rate
// v_mul_u32_u24 / v_mul_i32_i24); maps whose block-table strides exceed 2^23
// use the general 32-bit multiply.
template <bool kU24>
__device__ __forceinline__ uint32_t mul_u(uint32_t a, uint32_t b)
{
    if (kU24) return __umul24(a, b);
    return a * b;
}

template <bool kU24>
__device__ __forceinline__ int32_t mul_i(int32_t a, int32_t b)
{
    if (kU24) return __mul24(a, b);
    return a * b;
}

// Occupancy-rank form of the block table (LDS copy, 2^log2w words): for
// block up (< grid_occ_size), set = allocated and the return value = rank *
// sub_dim^3, which equals grid_occ[up] for maps whose offsets follow setData
// order.  Any up reads a word inside the table (word index = bits
// [5, 5 + log2w) of up), so lanes outside the grid need no clamp.  The bit
// field extracts use the low five bits of their offset and width operands,
// i.e. up mod 32.  The rank is below 2^17 and sub_dim^3 below 2^24, so the
// product is a 24-bit multiply.
__device__ __forceinline__ uint32_t rank_block(const OccWord* __restrict__ occ, uint32_t log2w, uint32_t up,
                                               uint32_t sd3, bool& set)
{
    const OccWord w = occ[__builtin_amdgcn_ubfe(up, 5u, log2w)];
    set = __builtin_amdgcn_ubfe(w.bits, up, 1u) != 0u;
    return __umul24((uint32_t)__popc(__builtin_amdgcn_ubfe(w.bits, 0u, up)) + w.prefix, sd3);
}

__device__ __forceinline__ int32_t rank_offset(const OccWord* __restrict__ occ, uint32_t log2w, uint32_t up,
                                               uint32_t sd3)
{
    bool set;
    const uint32_t off = rank_block(occ, log2w, up, sd3, set);
    return set ? (int32_t)off : -1;
}

template <bool kRank, bool kWide>
__device__ __forceinline__ int64_t block_offset(const DevMap& m, const OccWord* __restrict__ occ, uint32_t up)
{
    if (kWide) {
        const uint32_t slot = m.slots[up];
        return slot == kNoSlot ? -1 : (int64_t)((uint64_t)slot * m.sd3);
    }
    if (kRank) return rank_offset(occ, m.occ_log2, up, m.sd3);
    return m.grid_occ[up];
}

// Fast tier: 0 <= o < R on every axis (one unsigned compare of the bit
// pattern per axis, since floats >= +0 order like their bits).
__device__ __forceinline__ bool fast_ok(float ox, float oy, float oz, uint32_t range_bits)
{
    return max(max(__float_as_uint(ox), __float_as_uint(oy)), __float_as_uint(oz)) < range_bits;
}

// Wrap tier: |o| < R on every axis.
__device__ __forceinline__ bool wrap_ok(float ox, float oy, float oz, uint32_t range_bits)
{
    return max(max(__float_as_uint(ox) & 0x7fffffffu, __float_as_uint(oy) & 0x7fffffffu),
               __float_as_uint(oz) & 0x7fffffffu) < range_bits;
}

// Offset of local voxel (lx, ly, lz), each < sub_dim, inside a bricked block:
// ((lx>>2) + (ly>>2)*B + (lz>>2)*B^2)*64 + (lx&3) + (ly&3)*4 + (lz&3)*16.
__device__ __forceinline__ uint32_t brick_offset(const DevMap& m, uint32_t lx, uint32_t ly, uint32_t lz)
{
    return lx + 60u * (lx >> 2) + 4u * ly + m.brick_cy * (ly >> 2) + 16u * lz + m.brick_cz * (lz >> 2);
}

// Physical position of reference index x (x < 2^31) in the bricked array.
__device__ __forceinline__ uint32_t phys_from_linear(const DevMap& m, uint32_t x)
{
    const uint32_t slot = __umulhi(x, m.mag_sd3) >> m.sh_sd3;
    const uint32_t rem = x - slot * m.sd3;
    const uint32_t lz = __umulhi(rem, m.mag_sd2) >> m.sh_sd2;
    const uint32_t r2 = rem - lz * (uint32_t)m.sub_dim_2;
    const uint32_t ly = __umulhi(r2, m.mag_sd) >> m.sh_sd;
    const uint32_t lx = r2 - ly * (uint32_t)m.sub_dim;
    return slot * m.sd3 + brick_offset(m, lx, ly, lz);
}

// The same for any 64-bit reference index (wide maps).
__device__ __forceinline__ uint64_t phys_from_linear64(const DevMap& m, uint64_t x)
{
    const uint64_t slot = x / m.sd3;
    const uint32_t rem = (uint32_t)(x - slot * m.sd3);
    const uint32_t lz = __umulhi(rem, m.mag_sd2) >> m.sh_sd2;
    const uint32_t r2 = rem - lz * (uint32_t)m.sub_dim_2;
    const uint32_t ly = __umulhi(r2, m.mag_sd) >> m.sh_sd;
    const uint32_t lx = r2 - ly * (uint32_t)m.sub_dim;
    return slot * m.sd3 + brick_offset(m, lx, ly, lz);
}

// One axis of the fast tier: the in-block coordinate k of fraction f and
// its contribution to the physical in-block offset, from one LDS entry.
__device__ __forceinline__ uint32_t axis_part(uint32_t sub_base, uint32_t sub_shift,
                                              const AxisEntry* __restrict__ t, float f, uint32_t& k)
{
    const uint32_t key = __float_as_uint(f);
    const uint32_t b = (key > sub_base ? key - sub_base : 0u) >> sub_shift;
    const AxisEntry e = t[b];
    const bool hi = key >= e.thr;
    k = e.count + (hi ? 1u : 0u);
    return hi ? e.c_hi : e.c_lo;
}

// Fast tier for one lane, in two parts.  o >= 0 on every axis here, so
// trunc is a plain conversion and the fraction o - trunc(o) is exactly
// fract(o).
//
// Block part: the block's offset s in bytes of the code array, with valid
// = block allocated and inside the grid.
template <bool kU24, bool kRank>
__device__ __forceinline__ void fast_block(const DevMap& m, const OccWord* __restrict__ occ, float ox, float oy,
                                           float oz, uint32_t& s, bool& valid)
{
    const uint32_t ix = (uint32_t)ox, iy = (uint32_t)oy, iz = (uint32_t)oz;
    const bool grid = (ix < m.gthr_x) & (iy < m.gthr_y) & (iz < m.gthr_z);
    const uint32_t up = ix + mul_u<kU24>(iy, m.up_dim_x) + mul_u<kU24>(iz, m.up_dim_2);
    if (kRank) {
        // branch-free: lanes outside the grid read some word and discard
        // it; inside the grid every up lies in the table (padded past
        // grid_occ_size with unallocated blocks)
        bool set;
        s = rank_block(occ, m.occ_log2, up, m.sd3b, set);
        valid = grid & set;
    } else {
        int32_t g = -1;
        if (grid & (up < m.occ)) g = m.grid_occ[up];
        valid = g >= 0;
        s = valid ? (uint32_t)g << 1 : 0u;
    }
}

// Block part of a wide map: the block's byte offset in 64 bits from its slot.
template <bool kU24>
__device__ __forceinline__ void fast_block_wide(const DevMap& m, float ox, float oy, float oz, uint64_t& s,
                                                bool& valid)
{
    const uint32_t ix = (uint32_t)ox, iy = (uint32_t)oy, iz = (uint32_t)oz;
    const bool grid = (ix < m.gthr_x) & (iy < m.gthr_y) & (iz < m.gthr_z);
    const uint32_t up = ix + mul_u<kU24>(iy, m.up_dim_x) + mul_u<kU24>(iz, m.up_dim_2);
    uint32_t slot = kNoSlot;
    if (grid & (up < m.occ)) slot = m.slots[up];
    valid = slot != kNoSlot;
    s = valid ? (uint64_t)slot * m.sd3b : 0ull;
}

// In-block part: the sum pc of the three per-axis in-block byte-offset
// contributions and the in-block coordinates.  An axis whose coordinate
// equals sub_dim contributes kCarryPoison, so s + pc >= carry_limit flags a
// lane whose reference index leaves the block row (its address is then
// recomputed from the linear index).
__device__ __forceinline__ void fast_inblock(const DevMap& m, const AxisEntry* __restrict__ atab, float ox, float oy,
                                             float oz, uint32_t& pc, uint32_t& kx, uint32_t& ky, uint32_t& kz)
{
    const float fx = __builtin_amdgcn_fractf(ox), fy = __builtin_amdgcn_fractf(oy), fz = __builtin_amdgcn_fractf(oz);
    const uint32_t n = m.sub_entries;
    pc = axis_part(m.sub_base, m.sub_shift, atab, fx, kx) + axis_part(m.sub_base, m.sub_shift, atab + n, fy, ky) +
         axis_part(m.sub_base, m.sub_shift, atab + 2u * n, fz, kz);
}

// Wrap tier for one lane (|o| < R on every axis): the reference's index,
// data_size when rejected.  neg is the LDS copy of the wrap-band table.
template <bool kU24, bool kRank, bool kWide>
__device__ __forceinline__ uint64_t wrap_index(const DevMap& m, const SubEntry* __restrict__ tab,
                                               const uint64_t* __restrict__ neg, const OccWord* __restrict__ occ,
                                               float ox, float oy, float oz)
{
    const int32_t ix = (int32_t)ox, iy = (int32_t)oy, iz = (int32_t)oz;
    const bool nx = ox <= -1.0f, ny = oy <= -1.0f, nz = oz <= -1.0f;
    const bool grid = (((uint32_t)ix < m.gthr_x) | (nx & ((m.gneg & 1u) != 0u))) &
                      (((uint32_t)iy < m.gthr_y) | (ny & ((m.gneg & 2u) != 0u))) &
                      (((uint32_t)iz < m.gthr_z) | (nz & ((m.gneg & 4u) != 0u)));
    const int32_t up = ix + mul_i<kU24>(iy, (int32_t)m.up_dim_x) + mul_i<kU24>(iz, (int32_t)m.up_dim_2);
    int64_t s = -1;
    if (grid & ((uint32_t)up < m.occ)) s = block_offset<kRank, kWide>(m, occ, (uint32_t)up);
    const float fx = ox - (float)ix, fy = oy - (float)iy, fz = oz - (float)iz;
    int32_t ex = (int32_t)sub_index(m.sub_base, m.sub_shift, tab, fabsf(fx));
    int32_t ey = (int32_t)sub_index(m.sub_base, m.sub_shift, tab, fabsf(fy));
    int32_t ez = (int32_t)sub_index(m.sub_base, m.sub_shift, tab, fabsf(fz));
    ex = nx ? 0 : (fx < 0.0f ? -ex : ex);
    ey = ny ? 0 : (fy < 0.0f ? -ey : ey);
    ez = nz ? 0 : (fz < 0.0f ? -ez : ez);
    const int32_t small = ex + mul_i<kU24>(ey, m.sub_dim) + mul_i<kU24>(ez, m.sub_dim_2);
    const uint32_t mask = (nx ? 1u : 0u) | (ny ? 2u : 0u) | (nz ? 4u : 0u);
    const uint64_t idx = (uint64_t)(s + (int64_t)small) + neg[mask];
    return s >= 0 ? idx : m.data_size;
}

}  // namespace tsl
