// Arithmetic shared by host and device code of the MI355X sensor update.
//
// Everything on the sensor-update path here gives bit-identical results on
// the host (x86-64, gcc or clang) and on gfx950: IEEE float/double +,-,*,/
// only, no FMA contraction (all sources are compiled with -ffp-contract=off),
// no library math calls.  The motion model at the end is the exception: the
// double sin and cos of its noise draws and the log, sqrt and cos that make
// the draws come from each side's math library (their results are rounded
// to float before they meet the pose, which hides almost every last-bit
// difference); its float functions are the restatements below.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TSL_HD __host__ __device__ __forceinline__
#else
#define TSL_HD inline
#endif

namespace tsl {

// ---------------------------------------------------------------------------
// float <-> size_t as compiled by gcc for x86-64 (the reference build,
// CMakeLists.txt:5-14).  The reference assigns float values to size_t in its
// lookup (cuda_sub_voxel_map.tcc:53-55,58-60,74-76,132-134); for values the
// C++ standard leaves undefined (negative beyond -1, >= 2^64, NaN) the x86-64
// code gcc emits is
//     x >= 2^63 ? cvttss2si(x - 2^63) ^ 2^63 : cvttss2si(x)
// with cvttss2si returning trunc(x) when it fits in int64 and
// 0x8000000000000000 otherwise.  These helpers spell that out so that the GPU
// reproduces the reference's indices for every input.
// ---------------------------------------------------------------------------
TSL_HD uint64_t cvtt_f32_i64(float x)
{
    if (x >= -9223372036854775808.0f && x < 9223372036854775808.0f)
        return (uint64_t)(int64_t)x;
    return 0x8000000000000000ull;
}

TSL_HD uint64_t cvt_f32_u64(float x)
{
    if (x >= 9223372036854775808.0f)
        return cvtt_f32_i64(x - 9223372036854775808.0f) ^ 0x8000000000000000ull;
    return cvtt_f32_i64(x);
}

// size_t -> float, round to nearest even.
TSL_HD float cvt_u64_f32(uint64_t v) { return (float)v; }

// ---------------------------------------------------------------------------
// sin / cos of a float angle, evaluated in double and rounded once to float.
// Cody-Waite reduction by pi/2 (pi/2 = P1 + P2 + P3 with P1, P2 of 33
// significant bits, so k*P1 and k*P2 are exact for |k| < 2^20), then the
// Taylor series of sin (to r^19) and cos (to r^20) on |r| <= pi/4 in Horner
// form.  The same routine runs in the pose kernel and in the CPU oracle, so
// pose matrices agree bit for bit.  Domain |x| < 1.6e6 rad; NaN outside.
// ---------------------------------------------------------------------------
TSL_HD double rint_d(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_rint(x);
#else
    return __builtin_rint(x);
#endif
}

TSL_HD void sincosf_exact(float xf, float* s_out, float* c_out)
{
    const double x = (double)xf;
    const double ax = x < 0.0 ? -x : x;
    if (!(ax < 1.6e6)) {
        *s_out = __builtin_nanf("");
        *c_out = __builtin_nanf("");
        return;
    }
    if (x == 0.0) {  // sin(+-0) = +-0
        *s_out = xf;
        *c_out = 1.0f;
        return;
    }
    const double kd = rint_d(x * 0x1.45f306dc9c883p-1);
    const double r = ((x - kd * 0x1.921fb54400000p+0) - kd * 0x1.0b4611a600000p-34) -
                     kd * 0x1.3198a2e037073p-69;
    const double r2 = r * r;
    double ps = -0x1.2f49b46814157p-57;          // -1/19!
    ps = r2 * ps + 0x1.952c77030ad4ap-49;        // +1/17!
    ps = r2 * ps - 0x1.ae7f3e733b81fp-41;        // -1/15!
    ps = r2 * ps + 0x1.6124613a86d09p-33;        // +1/13!
    ps = r2 * ps - 0x1.ae64567f544e4p-26;        // -1/11!
    ps = r2 * ps + 0x1.71de3a556c734p-19;        // +1/9!
    ps = r2 * ps - 0x1.a01a01a01a01ap-13;        // -1/7!
    ps = r2 * ps + 0x1.1111111111111p-7;         // +1/5!
    ps = r2 * ps - 0x1.5555555555555p-3;         // -1/3!
    const double s = r + (r * r2) * ps;
    double pc = 0x1.e542ba4020225p-62;           // +1/20!
    pc = r2 * pc - 0x1.6827863b97d97p-53;        // -1/18!
    pc = r2 * pc + 0x1.ae7f3e733b81fp-45;        // +1/16!
    pc = r2 * pc - 0x1.93974a8c07c9dp-37;        // -1/14!
    pc = r2 * pc + 0x1.1eed8eff8d898p-29;        // +1/12!
    pc = r2 * pc - 0x1.27e4fb7789f5cp-22;        // -1/10!
    pc = r2 * pc + 0x1.a01a01a01a01ap-16;        // +1/8!
    pc = r2 * pc - 0x1.6c16c16c16c17p-10;        // -1/6!
    pc = r2 * pc + 0x1.5555555555555p-5;         // +1/4!
    pc = r2 * pc - 0x1.0000000000000p-1;         // -1/2!
    const double c = 1.0 + r2 * pc;
    const int q = (int)((int64_t)kd & 3);
    if (q == 0) { *s_out = (float)s;  *c_out = (float)c; }
    else if (q == 1) { *s_out = (float)c;  *c_out = (float)-s; }
    else if (q == 2) { *s_out = (float)-s; *c_out = (float)-c; }
    else { *s_out = (float)-c; *c_out = (float)s; }
}

// ---------------------------------------------------------------------------
// sin / cos of a float angle as glibc (2.28 and later) computes sinf / cosf.
// The reference's CPU evaluator calls unqualified sin(float) / cos(float)
// (tsdf_evaluator.cpp:108-113, 206-213), which libstdc++ resolves to
// std::sin(float), i.e. the C library's sinf (gcc may fuse a sin/cos pair
// into sincosf, which gives the same bits).  glibc's sinf/cosf is the
// algorithm of ARM's optimized-routines (Szabolcs Nagy, 2018): evaluate in
// double; |x| < 0.75 (top 12 bits below those of pi/4): odd/even polynomial
// directly (|x| < 2^-12: sin = x, cos = 1); |x| < 120: n = round(x 2/pi)
// from a 2^24-scaled product, r = x - n pi/2; otherwise Payne-Hanek
// reduction by 2/pi from a table of 24-bit windows of its bits; then a
// degree-7 sin or degree-8 cos polynomial on r, chosen and signed by n mod 4,
// rounded once to float.  On x86-64 glibc picks at load time between a build
// of that C code with FMA contraction (CPUs with FMA3) and one without; the
// two differ on 34 of the 2^32 inputs.  kFma selects the variant; both are
// checked bit for bit against glibc on every non-NaN float.
// ---------------------------------------------------------------------------
enum TrigMode : int {
    kTrigGlibcFma = 0,    // glibc sinf/cosf on x86-64 with FMA3 (the default)
    kTrigGlibcNoFma = 1,  // glibc sinf/cosf on x86-64 without FMA3
    kTrigDouble = 2,      // (float)sin((double)x): sincosf_exact
};

template <bool kFma>
TSL_HD double glibc_mad(double a, double b, double c)
{
    if (kFma) return __builtin_fma(a, b, c);
    return a * b + c;
}

TSL_HD uint32_t f32_top12(float f)
{
    union { float f; uint32_t u; } v;
    v.f = f;
    return (v.u >> 20) & 0x7ffu;
}

// polynomial on the reduced argument: n even -> sin, odd -> cos; neg picks
// the cos coefficients negated (n & 2 in the caller)
template <bool kFma>
TSL_HD float glibc_trig_poly(double x, double x2, bool neg, int n)
{
    if ((n & 1) == 0) {
        const double s1 = -0x1.555545995a603p-3, s2 = 0x1.1107605230bc4p-7, s3 = -0x1.994eb3774cf24p-13;
        const double x3 = x * x2;
        const double t = glibc_mad<kFma>(x2, s3, s2);
        const double x7 = x3 * x2;
        const double s = glibc_mad<kFma>(x3, s1, x);
        return (float)glibc_mad<kFma>(x7, t, s);
    }
    const double sg = neg ? -1.0 : 1.0;
    const double c0 = sg * 0x1p0, c1 = sg * -0x1.ffffffd0c621cp-2, c2 = sg * 0x1.55553e1068f19p-5,
                 c3 = sg * -0x1.6c087e89a359dp-10, c4 = sg * 0x1.99343027bf8c3p-16;
    const double x4 = x2 * x2;
    const double t2 = glibc_mad<kFma>(x2, c4, c3);
    const double t1 = glibc_mad<kFma>(x2, c1, c0);
    const double x6 = x4 * x2;
    const double c = glibc_mad<kFma>(x4, c2, t1);
    return (float)glibc_mad<kFma>(x6, t2, c);
}

// 2/pi in hex: 0.A2F9836E 4E441529 FC2757D1 F534DDC0 DB629599 3C439041 ...;
// word i is the 32 bits ending 8(i + 1) bits after the point
TSL_HD uint32_t two_over_pi_word(int i)
{
    const uint32_t b[7] = {0u, 0xa2f9836eu, 0x4e441529u, 0xfc2757d1u, 0xf534ddc0u, 0xdb629599u, 0x3c439041u};
    const int end = 8 * (i + 1);         // bit position after the point
    const int w = end >> 5, sh = end & 31;  // the window is the low bits of word w plus word w-1's
    const uint64_t hi = b[w], lo = (w + 1 < 7) ? b[w + 1] : 0u;
    const uint64_t two = (hi << 32) | lo;
    return (uint32_t)(two >> (32 - sh));
}

// Payne-Hanek reduction of |x| >= 120: r in double and the quadrant n
TSL_HD double glibc_reduce_large(uint32_t xi, int* np)
{
    const int base = (int)((xi >> 26) & 15);
    const int shift = (int)((xi >> 23) & 7);
    xi = (xi & 0xffffffu) | 0x800000u;
    xi <<= shift;
    uint64_t r0 = (uint32_t)(xi * two_over_pi_word(base));
    const uint64_t r1 = (uint64_t)xi * two_over_pi_word(base + 4);
    const uint64_t r2 = (uint64_t)xi * two_over_pi_word(base + 8);
    r0 = (r2 >> 32) | (r0 << 32);
    r0 += r1;
    const uint64_t n = (r0 + (1ull << 61)) >> 62;
    r0 -= n << 62;
    *np = (int)n;
    return (double)(int64_t)r0 * 0x1.921fb54442d18p-62;
}

template <bool kFma>
TSL_HD void sincosf_glibc(float y, float* s_out, float* c_out)
{
    double x = (double)y;
    const uint32_t top = f32_top12(y);
    if (top < 0x3f4u) {  // |y| < 0.75
        const double x2 = x * x;
        if (top < 0x398u) {  // |y| < 2^-12
            *s_out = y;
            *c_out = 1.0f;
            return;
        }
        *s_out = glibc_trig_poly<kFma>(x, x2, false, 0);
        *c_out = glibc_trig_poly<kFma>(x, x2, false, 1);
        return;
    }
    int n;
    double sgn;
    bool neg;
    if (top < 0x42fu) {  // |y| < 120
        const double r = x * 0x1.45f306dc9c883p+23;
        n = (int)(((int32_t)r + 0x800000) >> 24);
        x = glibc_mad<kFma>(-(double)n, 0x1.921fb54442d18p0, x);
        sgn = ((n & 3) == 1 || (n & 3) == 2) ? -1.0 : 1.0;
        neg = (n & 2) != 0;
    } else if (top < 0x7f8u) {
        union { float f; uint32_t u; } v;
        v.f = y;
        const int sign = (int)(v.u >> 31);
        x = glibc_reduce_large(v.u, &n);
        const int q = (n + sign) & 3;
        sgn = (q == 1 || q == 2) ? -1.0 : 1.0;
        neg = (q & 2) != 0;
    } else {  // inf or NaN
        *s_out = (y - y) / (y - y);
        *c_out = *s_out;
        return;
    }
    const double xs = x * sgn;
    *s_out = glibc_trig_poly<kFma>(xs, x * x, neg, n);
    *c_out = glibc_trig_poly<kFma>(xs, x * x, neg, n ^ 1);
}

TSL_HD void sincosf_mode(int mode, float x, float* s, float* c)
{
    if (mode == kTrigGlibcNoFma)
        sincosf_glibc<false>(x, s, c);
    else if (mode == kTrigDouble)
        sincosf_exact(x, s, c);
    else
        sincosf_glibc<true>(x, s, c);
}

// ---------------------------------------------------------------------------
// asinf, atanf and atan2f as glibc 2.35 computes them on x86-64
// (sysdeps/ieee754/flt-32/e_asinf.c, s_atanf.c, e_atan2f.c): the float
// versions of Sun's fdlibm algorithms (fdlibm 5.3, 1993), asinf with the
// degree-4 polynomial of Moshier's Cephes library.  The reference's
// getAngleFromMat (src/util/util.cpp:86-111) applies asin, atan2 and cos to
// floats; with libstdc++ those are the float functions.  Float arithmetic,
// no contraction; checked bit for bit against the host's glibc (asinf and
// atanf on every float).
// ---------------------------------------------------------------------------
TSL_HD uint32_t f32_bits(float x)
{
    union { float f; uint32_t u; } v;
    v.f = x;
    return v.u;
}

TSL_HD float f32_from_bits(uint32_t u)
{
    union { uint32_t u; float f; } v;
    v.u = u;
    return v.f;
}

TSL_HD float asinf_glibc(float x)
{
    const float pio2_hi = 1.57079637050628662109375f, pio2_lo = -4.37113900018624283e-8f,
                pio4_hi = 0.785398185253143310546875f;
    const float p0 = 1.666675248e-1f, p1 = 7.495297643e-2f, p2 = 4.547037598e-2f, p3 = 2.417951451e-2f,
                p4 = 4.216630880e-2f;
    const uint32_t hx = f32_bits(x), ix = hx & 0x7fffffffu;
    if (ix == 0x3f800000u) return x * pio2_hi + x * pio2_lo;  // +-pi/2
    if (ix > 0x3f800000u) return (x - x) / (x - x);           // |x| > 1 or NaN: NaN
    if (ix < 0x3f000000u) {                                   // |x| < 0.5
        if (ix < 0x32000000u) return x;                       // |x| < 2^-27
        const float t = x * x;
        const float w = t * (p0 + t * (p1 + t * (p2 + t * (p3 + t * p4))));
        return x + x * w;
    }
    // 0.5 <= |x| < 1
    float w = 1.0f - fabsf(x);
    float t = w * 0.5f;
    float p = t * (p0 + t * (p1 + t * (p2 + t * (p3 + t * p4))));
    const float s = sqrtf(t);
    if (ix >= 0x3f79999au) {  // |x| > 0.975
        t = pio2_hi - (2.0f * (s + s * p) - pio2_lo);
    } else {
        w = f32_from_bits(f32_bits(s) & 0xfffff000u);
        const float c = (t - w * w) / (s + w);
        const float r = p;
        p = 2.0f * s * r - (pio2_lo - 2.0f * c);
        const float q = pio4_hi - 2.0f * w;
        t = pio4_hi - (p - q);
    }
    return (int32_t)hx > 0 ? t : -t;
}

TSL_HD float atanf_glibc(float x)
{
    const float atanhi[4] = {4.6364760399e-01f, 7.8539812565e-01f, 9.8279368877e-01f, 1.5707962513e+00f};
    const float atanlo[4] = {5.0121582440e-09f, 3.7748947079e-08f, 3.4473217170e-08f, 7.5497894159e-08f};
    const float aT[11] = {3.3333334327e-01f,  -2.0000000298e-01f, 1.4285714924e-01f, -1.1111110449e-01f,
                          9.0908870101e-02f,  -7.6918758452e-02f, 6.6610731184e-02f, -5.8335702866e-02f,
                          4.9768779427e-02f,  -3.6531571299e-02f, 1.6285819933e-02f};
    const uint32_t hx = f32_bits(x), ix = hx & 0x7fffffffu;
    int id;
    if (ix >= 0x4c000000u) {  // |x| >= 2^25
        if (ix > 0x7f800000u) return x + x;  // NaN
        return (int32_t)hx > 0 ? atanhi[3] + atanlo[3] : -atanhi[3] - atanlo[3];
    }
    if (ix < 0x3ee00000u) {  // |x| < 0.4375
        if (ix < 0x31000000u) return x;  // |x| < 2^-29
        id = -1;
    } else {
        x = fabsf(x);
        if (ix < 0x3f980000u) {    // |x| < 1.1875
            if (ix < 0x3f300000u) {  // 7/16 <= |x| < 11/16
                id = 0;
                x = (2.0f * x - 1.0f) / (2.0f + x);
            } else {  // 11/16 <= |x| < 19/16
                id = 1;
                x = (x - 1.0f) / (x + 1.0f);
            }
        } else if (ix < 0x401c0000u) {  // |x| < 2.4375
            id = 2;
            x = (x - 1.5f) / (1.0f + 1.5f * x);
        } else {  // 2.4375 <= |x| < 2^25
            id = 3;
            x = -1.0f / x;
        }
    }
    const float z = x * x;
    const float w = z * z;
    const float s1 = z * (aT[0] + w * (aT[2] + w * (aT[4] + w * (aT[6] + w * (aT[8] + w * aT[10])))));
    const float s2 = w * (aT[1] + w * (aT[3] + w * (aT[5] + w * (aT[7] + w * aT[9]))));
    if (id < 0) return x - x * (s1 + s2);
    const float r = atanhi[id] - ((x * (s1 + s2) - atanlo[id]) - x);
    return (int32_t)hx < 0 ? -r : r;
}

TSL_HD float atan2f_glibc(float y, float x)
{
    const float tiny = 1.0e-30f, pi_o_4 = 7.8539818525e-01f, pi_o_2 = 1.5707963705e+00f, pi = 3.1415927410e+00f,
                pi_lo = -8.7422776573e-08f;
    const uint32_t hx = f32_bits(x), hy = f32_bits(y);
    const uint32_t ix = hx & 0x7fffffffu, iy = hy & 0x7fffffffu;
    if (ix > 0x7f800000u || iy > 0x7f800000u) return x + y;  // NaN
    if (hx == 0x3f800000u) return atanf_glibc(y);            // x = 1
    const int m = (int)((hy >> 31) & 1u) | (int)((hx >> 30) & 2u);  // 2 sign(x) + sign(y)
    if (iy == 0) {
        if (m < 2) return y;  // atan(+-0, +anything) = +-0
        return m == 2 ? pi + tiny : -pi - tiny;
    }
    if (ix == 0) return (int32_t)hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
    if (ix == 0x7f800000u) {
        if (iy == 0x7f800000u) {
            switch (m) {
            case 0: return pi_o_4 + tiny;
            case 1: return -pi_o_4 - tiny;
            case 2: return 3.0f * pi_o_4 + tiny;
            default: return -3.0f * pi_o_4 - tiny;
            }
        }
        switch (m) {
        case 0: return 0.0f;
        case 1: return -0.0f;
        case 2: return pi + tiny;
        default: return -pi - tiny;
        }
    }
    if (iy == 0x7f800000u) return (int32_t)hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
    const int k = ((int32_t)iy - (int32_t)ix) >> 23;
    float z;
    if (k > 60)
        z = pi_o_2 + 0.5f * pi_lo;  // |y/x| > 2^60
    else if ((int32_t)hx < 0 && k < -60)
        z = 0.0f;  // |y|/x < -2^60
    else
        z = atanf_glibc(fabsf(y / x));
    switch (m) {
    case 0: return z;
    case 1: return f32_from_bits(f32_bits(z) ^ 0x80000000u);
    case 2: return pi - (z - pi_lo);
    default: return (z - pi_lo) - pi;
    }
}

// ---------------------------------------------------------------------------
// Pose matrix of a particle composed with the scanner->base transform
// (tsdf_evaluator.cpp:102-145): L = [R | t] * T, R = Rz(yaw) Ry(pitch)
// Rx(roll), left-to-right float evaluation exactly as the reference writes it.
// pose = (x, y, z, roll, pitch, yaw); tf = row-major, entries 0..11 used.
// ---------------------------------------------------------------------------
TSL_HD void pose_matrix(const float* pose, const float* tf, float* L, int trig = kTrigGlibcFma)
{
    float sa, ca, sb, cb, sg, cg;
    sincosf_mode(trig, pose[3], &sa, &ca);
    sincosf_mode(trig, pose[4], &sb, &cb);
    sincosf_mode(trig, pose[5], &sg, &cg);
    float p[12];
    p[0] = cb * cg;
    p[4] = cb * sg;
    p[8] = -sb;
    p[3] = pose[0];
    p[1] = sa * sb * cg - ca * sg;
    p[5] = sa * sb * sg + ca * cg;
    p[9] = sa * cb;
    p[7] = pose[1];
    p[2] = ca * sb * cg + sa * sg;
    p[6] = ca * sb * sg - sa * cg;
    p[10] = ca * cb;
    p[11] = pose[2];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        L[4 * r + 0] = p[4 * r] * tf[0] + p[4 * r + 1] * tf[4] + p[4 * r + 2] * tf[8];
        L[4 * r + 1] = p[4 * r] * tf[1] + p[4 * r + 1] * tf[5] + p[4 * r + 2] * tf[9];
        L[4 * r + 2] = p[4 * r] * tf[2] + p[4 * r + 1] * tf[6] + p[4 * r + 2] * tf[10];
        L[4 * r + 3] = p[4 * r] * tf[3] + p[4 * r + 1] * tf[7] + p[4 * r + 2] * tf[11] + p[4 * r + 3];
    }
}

// ---------------------------------------------------------------------------
// Motion noise: counter-based draws, so that particle i of a filter gets the
// same noise whatever else runs with it.  SplitMix64 (Steele, Lea, Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014) mixes
// (seed, particle, draw, half) into 64 bits; 53 of them give a uniform in
// (0, 1); Box-Muller (cosine branch) turns two uniforms into one standard
// normal.
// ---------------------------------------------------------------------------
TSL_HD uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

TSL_HD double motion_uniform(uint64_t seed, uint64_t particle, uint32_t draw, uint32_t half)
{
    const uint64_t h = splitmix64(splitmix64(seed ^ splitmix64(particle)) + (uint64_t)(2u * draw + half));
    return ((double)(h >> 11) + 0.5) * 0x1.0p-53;
}

TSL_HD double motion_normal(uint64_t seed, uint64_t particle, uint32_t draw)
{
    const double u = motion_uniform(seed, particle, draw, 0);
    const double v = motion_uniform(seed, particle, draw, 1);
    return sqrt(-2.0 * log(u)) * cos(6.283185307179586 * v);
}

// ParticleCloud::apply_model (src/particle_cloud.cpp:499-617) for one
// particle given its six draws d = (dx, dy, dz, roll, pitch, yaw) (double,
// as std::normal_distribution<> returns them): the noisy odometry step is
// rotated into the particle's frame (translation added, rotation composed)
// and the new angles are read back as getAngleFromMat does
// (src/util/util.cpp:86-111).  Float arithmetic left to right, no FMA.
TSL_HD void apply_motion(const float last[6], const double d[6], float out[6], int trig = kTrigGlibcFma)
{
    const float sin_roll = (float)sin(d[3]), cos_roll = (float)cos(d[3]);
    const float sin_pitch = (float)sin(d[4]), cos_pitch = (float)cos(d[4]);
    const float sin_yaw = (float)sin(d[5]), cos_yaw = (float)cos(d[5]);
    float o[12];
    o[0] = cos_pitch * cos_yaw;
    o[4] = cos_pitch * sin_yaw;
    o[8] = -sin_pitch;
    o[3] = (float)d[0];
    o[1] = sin_roll * sin_pitch * cos_yaw - cos_roll * sin_yaw;
    o[5] = sin_roll * sin_pitch * sin_yaw + cos_roll * cos_yaw;
    o[9] = sin_roll * cos_pitch;
    o[7] = (float)d[1];
    o[2] = cos_roll * sin_pitch * cos_yaw + sin_roll * sin_yaw;
    o[6] = cos_roll * sin_pitch * sin_yaw - sin_roll * cos_yaw;
    o[10] = cos_roll * cos_pitch;
    o[11] = (float)d[2];
    float sa, ca, sb, cb, sg, cg;
    sincosf_mode(trig, last[3], &sa, &ca);
    sincosf_mode(trig, last[4], &sb, &cb);
    sincosf_mode(trig, last[5], &sg, &cg);
    float p[12];
    p[0] = cb * cg;
    p[4] = cb * sg;
    p[8] = -sb;
    p[3] = 0.0f;
    p[1] = sa * sb * cg - ca * sg;
    p[5] = sa * sb * sg + ca * cg;
    p[9] = sa * cb;
    p[7] = 0.0f;
    p[2] = ca * sb * cg + sa * sg;
    p[6] = ca * sb * sg - sa * cg;
    p[10] = ca * cb;
    p[11] = 0.0f;
    float t[12];
    for (int r = 0; r < 3; ++r) {
        t[4 * r + 0] = p[4 * r] * o[0] + p[4 * r + 1] * o[4] + p[4 * r + 2] * o[8];
        t[4 * r + 1] = p[4 * r] * o[1] + p[4 * r + 1] * o[5] + p[4 * r + 2] * o[9];
        t[4 * r + 2] = p[4 * r] * o[2] + p[4 * r + 1] * o[6] + p[4 * r + 2] * o[10];
        t[4 * r + 3] = p[4 * r] * o[3] + p[4 * r + 1] * o[7] + p[4 * r + 2] * o[11] + p[4 * r + 3];
    }
    // getAngleFromMat on the 3x4 part (entries 0, 4, 8, 9, 10 of the 4x4):
    // float arguments, so asin, atan2 and cos bind to the float functions
    // (glibc's asinf/atan2f, cosf of the trig mode), the divisions are float;
    // in the double mode, the C library's double functions
    float roll, pitch, yaw;
    if (trig == kTrigDouble) {
        if (fabs((double)t[8]) >= 1.0) {
            yaw = 0.0f;
            const double delta = atan2((double)t[9], (double)t[10]);
            pitch = t[8] < 0.0f ? (float)(3.141592653589793 / 2.0) : (float)(-3.141592653589793 / 2.0);
            roll = (float)delta;
        } else {
            pitch = (float)(-asin((double)t[8]));
            const double cpd = cos((double)pitch);
            roll = (float)atan2((double)t[9] / cpd, (double)t[10] / cpd);
            yaw = (float)atan2((double)t[4] / cpd, (double)t[0] / cpd);
        }
    } else if (fabsf(t[8]) >= 1.0f) {
        yaw = 0.0f;
        roll = atan2f_glibc(t[9], t[10]);
        pitch = t[8] < 0.0f ? (float)(3.141592653589793 / 2.0) : (float)(-3.141592653589793 / 2.0);
    } else {
        pitch = -asinf_glibc(t[8]);
        float sp, cp;
        sincosf_mode(trig, pitch, &sp, &cp);
        roll = atan2f_glibc(t[9] / cp, t[10] / cp);
        yaw = atan2f_glibc(t[4] / cp, t[0] / cp);
    }
    out[0] = last[0] + t[3];
    out[1] = last[1] + t[7];
    out[2] = last[2] + t[11];
    out[3] = roll;
    out[4] = pitch;
    out[5] = yaw;
}

}  // namespace tsl
