// Systematic resampling on the device, bit-exact with
// SystematicResampler::resample (include/tsdf_localization/resampling/
// novel_resampling.h:41-72):
//
//   inverse_M = 1.0 / M                      (double)
//   U = draw from [0, inverse_M)              (float)
//   s = 0.0;  for m: s += w_m                 (double running sum of float weights)
//             while (s > U) { emit m; U += inverse_M; }   (U stays a float)
//
// Equivalently output j is particle m(j) = min{m : s_m > U_j} and the number
// emitted is J = #{j : U_j < s_{M-1}}.  Both sequences are sequential
// recurrences with rounding, so the kernel
//   1. computes a candidate for each in parallel (a block scan of the
//      weights in double; U_j from a per-binade table of constant float
//      steps built on the host),
//   2. checks every element against one step of the recurrence from its
//      predecessor (s_m == s_{m-1} + w_m, U_j == float(U_{j-1} + 1/M)); if all
//      checks pass, the candidate equals the sequential result by induction,
//   3. otherwise recomputes serially from the first failing element,
// then assigns outputs by binary search and gathers the particles.
#include "tsl_internal.h"

#include <cmath>
#include <vector>

namespace tsl {

struct USeg {
    uint32_t j0;   // first index of the run
    float u;       // U_{j0}
    double step;   // constant difference inside the run (exact)
};

namespace {

constexpr int kRsThreads = 1024;
constexpr int kMaxSeg = 128;

// Passed by value as a kernel argument (2 KB), so no host buffer has to
// outlive the launch.
struct USegTable {
    USeg seg[kMaxSeg];
};

__device__ __forceinline__ double wave_incl_scan(double v, int lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    return v;
}

// Phase A (one workgroup): s_m, the double running sum, candidate by block
// scan, checked element by element, repaired serially from the first
// mismatch.
__global__ __launch_bounds__(kRsThreads) void k_rs_scan(const tsl_particle* __restrict__ in, uint32_t M,
                                                        double* __restrict__ s)
{
    __shared__ double wave_tot[kRsThreads / 64];
    __shared__ double carry_sh;
    __shared__ uint32_t first_bad;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int kItems = 4;
    constexpr uint32_t kChunk = kRsThreads * kItems;
    if (tid == 0) {
        carry_sh = 0.0;
        first_bad = 0xffffffffu;
    }
    __syncthreads();
    for (uint32_t base = 0; base < M; base += kChunk) {
        double loc[kItems];
        double run = 0.0;
#pragma unroll
        for (int e = 0; e < kItems; ++e) {
            const uint32_t m = base + tid * kItems + e;
            const double w = m < M ? (double)in[m].second : 0.0;
            run += w;
            loc[e] = run;
        }
        const double incl = wave_incl_scan(run, lane);
        if (lane == 63) wave_tot[wave] = incl;
        __syncthreads();
        double woff = 0.0;
        for (int w = 0; w < wave; ++w) woff += wave_tot[w];
        const double carry = carry_sh;
        const double excl = carry + (woff + (incl - run));
#pragma unroll
        for (int e = 0; e < kItems; ++e) {
            const uint32_t m = base + tid * kItems + e;
            if (m < M) s[m] = excl + loc[e];
        }
        __syncthreads();
        if (tid == kRsThreads - 1) {
            double tot = 0.0;
            for (int w = 0; w < kRsThreads / 64; ++w) tot += wave_tot[w];
            carry_sh = carry + tot;
        }
        __syncthreads();
    }
    // check s_m == s_{m-1} + w_m
    for (uint32_t m = tid; m < M; m += kRsThreads) {
        const double prev = m == 0 ? 0.0 : s[m - 1];
        if (s[m] != prev + (double)in[m].second) atomicMin(&first_bad, m);
    }
    __syncthreads();
    if (first_bad != 0xffffffffu && tid == 0) {
        double acc = first_bad == 0 ? 0.0 : s[first_bad - 1];
        for (uint32_t m = first_bad; m < M; ++m) {
            acc += (double)in[m].second;
            s[m] = acc;
        }
    }
}

__device__ __forceinline__ float u_candidate(const USegTable& segs, uint32_t nseg, uint32_t j)
{
    uint32_t lo = 0, hi = nseg;  // last run with j0 <= j
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (segs.seg[mid].j0 <= j) lo = mid; else hi = mid;
    }
    const USeg sg = segs.seg[lo];
    return (float)((double)sg.u + (double)(j - sg.j0) * sg.step);
}

// Phase B (grid): U_j candidates from the run table, each checked against one
// step of the recurrence from its predecessor.
__global__ void k_rs_u(const USegTable segs, uint32_t nseg, uint32_t cap, double inv_m, float u0,
                       float* __restrict__ U, uint32_t* __restrict__ first_bad)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cap) return;
    const float u = u_candidate(segs, nseg, j);
    U[j] = u;
    const float expect = j == 0 ? u0 : (float)((double)u_candidate(segs, nseg, j - 1) + inv_m);
    if (u != expect) atomicMin(first_bad, j);
}

// Phase C (one thread): serial repair of U from the first mismatch, then the
// number emitted J = #{j : U_j < s_{M-1}} (U is non-decreasing).
__global__ void k_rs_count(uint32_t M, const double* __restrict__ s, float* __restrict__ U, uint32_t cap,
                           double inv_m, float u0, uint32_t* __restrict__ first_bad, uint32_t* __restrict__ count_out,
                           uint32_t* __restrict__ status_out)
{
    if (threadIdx.x != 0) return;
    const uint32_t fb = *first_bad;
    if (fb != 0xffffffffu) {
        float u = fb == 0 ? u0 : (float)((double)U[fb - 1] + inv_m);
        for (uint32_t j = fb; j < cap; ++j) {
            U[j] = u;
            u = (float)((double)u + inv_m);
        }
    }
    *first_bad = 0xffffffffu;
    const double s_final = s[M - 1];
    if (cap == 0 || (double)U[cap - 1] < s_final) {
        *status_out = 2u;
        *count_out = 0u;
        return;
    }
    uint32_t lo = 0, hi = cap;  // first j with U_j >= s_final
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((double)U[mid] < s_final) lo = mid + 1; else hi = mid;
    }
    *status_out = 0u;
    *count_out = lo;
}

// Phase D (grid): output j is particle min{m : s_m > U_j}.
__global__ void k_rs_assign(const tsl_particle* __restrict__ in, uint32_t M, const double* __restrict__ s,
                            const float* __restrict__ U, const uint32_t* __restrict__ count,
                            const uint32_t* __restrict__ status, int32_t* __restrict__ idx_out,
                            tsl_particle* __restrict__ out)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (*status != 0u || j >= *count) return;
    const double u = (double)U[j];
    uint32_t lo = 0, hi = M;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s[mid] > u) hi = mid; else lo = mid + 1;
    }
    const uint32_t m = lo < M ? lo : M - 1;
    if (idx_out) idx_out[j] = (int32_t)m;
    if (out) out[j] = in[m];
}

float rec_step(float u, double inv_m) { return (float)((double)u + inv_m); }

}  // namespace

// Run table of the U recurrence: inside a binade the float step is constant
// (rounding of U + 1/M depends only on the binade once U is on the float
// grid), so U_j = U_{j0} + (j - j0)*step exactly.  Runs end before the next
// power of two; a step that differs from its successor (round-to-even ties,
// U = 0, subnormals) becomes a run of length 1.  The kernel re-checks every
// element, so this table only needs to be right to be fast, not to be exact.
static int build_segments(size_t M, float u0, uint32_t cap, std::vector<USeg>& segs)
{
    const double inv_m = 1.0 / (double)M;
    segs.clear();
    uint32_t j = 0;
    float u = u0;
    while (j < cap) {
        if ((int)segs.size() >= kMaxSeg) break;  // remainder repaired serially on the device
        const float n1 = rec_step(u, inv_m);
        const double d = (double)n1 - (double)u;
        if (!(d > 0.0)) return set_error(TSL_ERR_UNSUPPORTED, "systematic draw does not advance (M=%zu)", M);
        uint64_t L = 1;
        if (u >= 1.1754943508222875e-38f) {
            int e;
            std::frexp((double)u, &e);
            const double top = std::ldexp(1.0, e);  // next power of two above u
            const double n2 = (double)rec_step(n1, inv_m) - (double)n1;
            if (n2 == d) {
                double steps = std::ceil((top - (double)u) / d) - 1.0;
                if (steps < 1.0) steps = 1.0;
                if (steps > (double)(cap - j)) steps = (double)(cap - j);
                L = (uint64_t)steps;
                const double last = (double)u + (double)(L - 1) * d;
                if ((double)rec_step((float)last, inv_m) != last + d) L = 1;
            }
        }
        segs.push_back(USeg{j, u, d});
        j += (uint32_t)L;
        u = (float)((double)u + (double)L * d);
    }
    return TSL_OK;
}

// Number of U values below 1 + 1e-5, plus one: an upper bound on J for
// normalised weights (their double sum is within ~1e-7 of 1).
int resample_capacity(size_t M, float u0, size_t* cap_out)
{
    if (M == 0) {
        *cap_out = 0;
        return TSL_OK;
    }
    const double inv_m = 1.0 / (double)M;
    const double stop = 1.0 + 1e-5;
    // walk the runs without materialising them
    size_t j = 0;
    float u = u0;
    while ((double)u < stop) {
        const float n1 = rec_step(u, inv_m);
        const double d = (double)n1 - (double)u;
        if (!(d > 0.0)) return set_error(TSL_ERR_UNSUPPORTED, "systematic draw does not advance (M=%zu)", M);
        uint64_t L = 1;
        if (u >= 1.1754943508222875e-38f) {
            int e;
            std::frexp((double)u, &e);
            const double top = std::ldexp(1.0, e);
            double steps = std::ceil((top - (double)u) / d) - 1.0;
            if (steps < 1.0) steps = 1.0;
            const double lim = std::ceil((stop - (double)u) / d);
            if (steps > lim) steps = lim < 1.0 ? 1.0 : lim;
            L = (uint64_t)steps;
            const double last = (double)u + (double)(L - 1) * d;
            if ((double)rec_step((float)last, inv_m) != last + d || (double)rec_step(n1, inv_m) - (double)n1 != d) L = 1;
        }
        j += L;
        u = (float)((double)u + (double)L * d);
        if (j > 4 * M + 1024) break;
    }
    *cap_out = j + 1;
    return TSL_OK;
}

int launch_resample_buffers(const tsl_particle* d_in, size_t M, float u0, size_t cap, double* d_s,
                            float* d_u, int32_t* d_idx, tsl_particle* d_out, uint32_t* d_count,
                            uint32_t* d_status, uint32_t* d_first_bad, hipStream_t stream)
{
    if (M == 0 || M > 0x7fffffffu || cap > 0xffffffffu)
        return set_error(TSL_ERR_INVALID_ARGUMENT, "resample: M=%zu out of range", M);
    std::vector<USeg> segs;
    TSL_TRY(build_segments(M, u0, (uint32_t)cap, segs));
    USegTable table{};
    for (size_t k = 0; k < segs.size(); ++k) table.seg[k] = segs[k];
    const double inv_m = 1.0 / (double)M;
    const unsigned gc = (unsigned)((cap + 255) / 256);
    hipLaunchKernelGGL(k_rs_scan, dim3(1), dim3(kRsThreads), 0, stream, d_in, (uint32_t)M, d_s);
    TSL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_rs_u, dim3(gc ? gc : 1), dim3(256), 0, stream, table, (uint32_t)segs.size(), (uint32_t)cap,
                       inv_m, u0, d_u, d_first_bad);
    TSL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_rs_count, dim3(1), dim3(64), 0, stream, (uint32_t)M, d_s, d_u, (uint32_t)cap, inv_m, u0,
                       d_first_bad, d_count, d_status);
    TSL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_rs_assign, dim3(gc ? gc : 1), dim3(256), 0, stream, d_in, (uint32_t)M, d_s, d_u, d_count,
                       d_status, d_idx, d_out);
    TSL_HIP_CHECK(hipGetLastError());
    return TSL_OK;
}

}  // namespace tsl
