// Host-side tables that let the device lookup reproduce the reference's
// float divisions and float->size_t conversions without dividing.
//
// 1. Sub-voxel index.  The reference computes, per axis,
//        k = (size_t)(sub_pos / resolution)     (cuda_sub_voxel_map.tcc:130-134)
//    with sub_pos in [0, 1) (or its negative for offsets in (-1, 0)).  k is a
//    non-decreasing step function of sub_pos, so it is fixed by the
//    thresholds T_j = min{x : trunc(fl(x / res)) >= j}, found here by binary
//    search over float bit patterns using the host's IEEE division.  Floats
//    >= 0 order like their bit patterns, so the device maps |sub_pos| to a
//    bucket b = sat(bits - base_bits) >> shift, chosen so that no bucket
//    holds two thresholds, and reads (threshold bits, count below the bucket)
//    from a small table: k = count + (bits >= threshold).
// 2. Offsets in (-2^23, -1].  There gcc's size_t conversion wraps:
//    u = 2^64 + trunc(o), and fl(u) = 2^64 exactly, so the grid check and
//    the in-block index of such an axis do not depend on o; both are
//    per-map constants computed here with the same conversions.
#include "tsl_internal.h"

#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

namespace tsl {

static inline uint32_t fbits(float f)
{
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

static inline float bitsf(uint32_t b)
{
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// trunc(fl(a / res)) as the reference computes it (a >= 0, small result)
static inline uint64_t ref_sub(float a, float res) { return cvt_f32_u64(a / res); }

int build_sub_table(float res, SubTable* out)
{
    const uint32_t one = fbits(1.0f);
    std::vector<uint32_t> thr;  // bit patterns of T_1, T_2, ... below 1.0f
    for (uint64_t k = 1;; ++k) {
        if (ref_sub(bitsf(one - 1), res) < k) break;  // no x < 1 reaches k
        uint32_t lo = 0, hi = one - 1;                // smallest bits in [lo, hi] with ref_sub >= k
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (ref_sub(bitsf(mid), res) >= k) hi = mid; else lo = mid + 1;
        }
        if (!thr.empty() && lo <= thr.back())
            return set_error(TSL_ERR_UNSUPPORTED, "sub-voxel thresholds not increasing (res %g)", (double)res);
        thr.push_back(lo);
        if (k > 4096) return set_error(TSL_ERR_UNSUPPORTED, "resolution %g too fine", (double)res);
    }
    SubTable t;
    // base: largest power of two <= T_1 (every threshold lies at or above it)
    uint32_t base = 0;
    if (!thr.empty()) base = thr[0] & 0x7f800000u;  // exponent bits of T_1 == 2^floor(log2 T_1)
    int shift = 23;
    for (; shift >= 0; --shift) {
        bool ok = true;
        for (size_t j = 1; j < thr.size() && ok; ++j)
            ok = ((thr[j] - base) >> shift) > ((thr[j - 1] - base) >> shift);
        const uint32_t nb = ((one - 1 - base) >> shift) + 1;
        if (ok && nb <= kMaxSubTable) break;
    }
    if (shift < 0) return set_error(TSL_ERR_UNSUPPORTED, "no bucket table for resolution %g", (double)res);
    const uint32_t nb = ((one - 1 - base) >> shift) + 1;
    t.shift = (uint32_t)shift;
    t.base_bits = base;
    t.entries.assign(nb, SubEntry{0x7f800000u, 0});  // threshold +inf
    size_t j = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        const uint64_t lo_bits = (uint64_t)base + ((uint64_t)b << shift);  // bucket 0 also covers [0, base)
        while (j < thr.size() && thr[j] < lo_bits) ++j;
        t.entries[b].count = (uint32_t)j;
        const uint64_t hi_bits = (uint64_t)base + ((uint64_t)(b + 1) << shift);
        if (j < thr.size() && thr[j] < hi_bits) t.entries[b].thr = thr[j];
    }
    *out = std::move(t);
    return TSL_OK;
}

uint32_t sub_lookup_host(const SubTable& t, float a)
{
    const uint32_t key = fbits(a);
    const uint32_t b = (key > t.base_bits ? key - t.base_bits : 0u) >> t.shift;
    const SubEntry& e = t.entries[b];
    return e.count + (key >= e.thr ? 1u : 0u);
}

// floor(x / d) == umulhi(x, mag) >> sh for every x < 2^31 (d >= 2):
// p = 31 + ceil(log2 d), mag = ceil(2^p / d) < 2^32; the rounding error of
// x*mag/2^p over x/d is below x/2^p < 1/d, so the floor is exact.
void magic_divisor(uint32_t d, uint32_t* mag, uint32_t* sh)
{
    uint32_t lg = 0;
    while ((1ull << lg) < d) ++lg;
    const uint32_t p = 31 + lg;
    const unsigned __int128 two_p = (unsigned __int128)1 << p;
    *mag = (uint32_t)((two_p + d - 1) / d);
    *sh = p - 32;
}

void negative_band_constants(const tsl_map_coef& c, DevMapFull& d)
{
    const float two64 = 18446744073709551616.0f;
    d.gneg_x = cvt_f32_u64(two64 / c.resolution) < c.dim_x ? 1u : 0u;
    d.gneg_y = cvt_f32_u64(two64 / c.resolution) < c.dim_y ? 1u : 0u;
    d.gneg_z = cvt_f32_u64(two64 / c.resolution) < c.dim_z ? 1u : 0u;
    // sub_pos = fl(o - fl(u)) = fl(o - 2^64) = -2^64 for |o| < 2^39
    const uint64_t s_neg = cvt_f32_u64(-two64 / c.resolution);
    d.neg_x = s_neg;
    d.neg_y = s_neg * c.sub_dim;
    d.neg_z = s_neg * c.sub_dim_2;
}

}  // namespace tsl

extern "C" {

// Host-only self-check of the sub-voxel table for one resolution: compares
// the table lookup with the reference's division for every float in
// [0, 1) whose bit pattern is a multiple of `stride` (stride 1 = all
// 1065353216 of them).  Writes the number of mismatches and returns TSL_OK,
// or the table builder's error.  No counterpart in the reference.
int tsl_check_sub_table(float res, uint32_t stride, uint64_t* mismatches, uint32_t* entries)
{
    using namespace tsl;
    if (!mismatches || stride == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "bad argument");
    SubTable t;
    TSL_TRY(build_sub_table(res, &t));
    if (entries) *entries = (uint32_t)t.entries.size();
    const uint32_t one = fbits(1.0f);
    std::atomic<uint64_t> bad{0};
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nt; ++w) {
        pool.emplace_back([&, w]() {
            uint64_t local = 0;
            for (uint64_t b = (uint64_t)w * stride; b < one; b += (uint64_t)nt * stride) {
                const float a = bitsf((uint32_t)b);
                if ((uint64_t)sub_lookup_host(t, a) != ref_sub(a, res)) ++local;
            }
            bad += local;
        });
    }
    for (auto& th : pool) th.join();
    *mismatches = bad.load();
    return TSL_OK;
}

// Host-only self-check of the constant-divisor multipliers used by the
// brick layout: compares umulhi(x, mag) >> sh with x / d for every x < 2^31
// whose value is a multiple of stride.  No counterpart in the reference.
int tsl_check_magic_divisor(uint32_t d, uint32_t stride, uint64_t* mismatches)
{
    using namespace tsl;
    if (!mismatches || d < 2 || stride == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "bad argument");
    uint32_t mag, sh;
    magic_divisor(d, &mag, &sh);
    std::atomic<uint64_t> bad{0};
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nt; ++w) {
        pool.emplace_back([&, w]() {
            uint64_t local = 0;
            for (uint64_t x = (uint64_t)w * stride; x < 0x80000000ull; x += (uint64_t)nt * stride) {
                const uint32_t q = (uint32_t)(((uint64_t)(uint32_t)x * mag) >> 32) >> sh;
                if (q != (uint32_t)x / d) ++local;
            }
            bad += local;
        });
    }
    for (auto& th : pool) th.join();
    *mismatches = bad.load();
    return TSL_OK;
}

}  // extern "C"
