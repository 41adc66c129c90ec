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
// recurrences with rounding, so the kernels
//   1. compute a candidate for each in parallel (a scan of the weights in
//      double, split over workgroups that start from the sum of the earlier
//      workgroups' totals; U_j from a per-binade table of constant float
//      steps),
//   2. check every element against one step of the recurrence from its
//      predecessor (s_m == s_{m-1} + w_m, U_j == float(U_{j-1} + 1/M)); if all
//      checks pass, the candidate equals the sequential result by induction,
//   3. otherwise recompute serially from the first failing element,
// then assign outputs by binary search and gather the particles.
//
// Every kernel is indexed by filter ("member"): the plan (M, the U run
// table, the number of candidates) is built on the device from the
// filter's device-side count, so a chain of updates and resamplings needs
// no host round trip and can be captured in a graph.
#include "tsl_internal.h"

#include <cmath>
#include <vector>

namespace tsl {

namespace {

constexpr int kRsThreads = 1024;
constexpr uint32_t kRsSmall = 2048;  // a lone filter with at most this many candidates resamples in one launch

__host__ __device__ __forceinline__ float rec_step(float u, double inv_m) { return (float)((double)u + inv_m); }

// Candidates needed for weights summing to at most 1 + 1e-5: the number of
// U_j below 1 + 1e-5, plus one; 0 if the draw does not advance.  Walks the
// runs of the U recurrence without materialising them: inside a binade the
// float step is constant (rounding of U + 1/M depends only on the binade
// once U is on the float grid), so U_j = U_{j0} + (j - j0)*step exactly.
// Runs end before the next power of two; a step that differs from its
// successor (round-to-even ties, U = 0, subnormals) becomes a run of length
// 1.  With seg, the first kMaxSeg runs are also recorded as the run table
// the kernels rebuild U from; they re-check every element, so the table
// only needs to be right to be fast, not to be exact.
__host__ __device__ uint64_t cap_walk(uint64_t M, float u0, double inv_m, USeg* seg = nullptr, uint32_t* nseg = nullptr)
{
    const double stop = 1.0 + 1e-5;
    uint64_t j = 0;
    uint32_t n = 0;
    float u = u0;
    while ((double)u < stop) {
        const float n1 = rec_step(u, inv_m);
        const double d = (double)n1 - (double)u;
        if (!(d > 0.0)) return 0;
        uint64_t L = 1;
        if (u >= 1.1754943508222875e-38f) {
            int e;
            (void)frexp((double)u, &e);
            const double top = ldexp(1.0, e);  // next power of two above u
            double steps = ceil((top - (double)u) / d) - 1.0;
            if (steps < 1.0) steps = 1.0;
            const double lim = ceil((stop - (double)u) / d);
            if (steps > lim) steps = lim < 1.0 ? 1.0 : lim;
            L = (uint64_t)steps;
            const double last = (double)u + (double)(L - 1) * d;
            if ((double)rec_step((float)last, inv_m) != last + d || (double)rec_step(n1, inv_m) - (double)n1 != d)
                L = 1;
        }
        if (seg && n < (uint32_t)kMaxSeg) {
            seg[n].j0 = (uint32_t)j;
            seg[n].u = u;
            seg[n].step = d;
            ++n;
        }
        j += L;
        u = (float)((double)u + (double)L * d);
        if (j > 4 * M + 1024) break;
    }
    if (nseg) *nseg = n;
    return j + 1;
}

// Plan (one thread per member): M from the device count, cap, run table.
// A member whose update failed (status set) or that holds no particles is
// skipped (cap = 0) and keeps its particles.
__device__ void plan_member(const Member& m, float u0_arg, USeg* seg_out = nullptr)
{
    FilterState* st = m.st;
    RsPlan* p = m.plan;
    const uint32_t M = st->count;
    const float u0 = m.u0_dev ? *m.u0_dev : u0_arg;
    st->first_bad = 0xffffffffu;
    p->s_bad = 0xffffffffu;
    p->M = M;
    p->u0 = u0;
    p->cap = 0;
    p->nseg = 0;
    if (st->status != kStOk) return;
    if (M == 0) {
        st->status = kStEmpty;
        return;
    }
    const double inv_m = 1.0 / (double)M;
    uint32_t nseg = 0;
    const uint64_t cap = cap_walk(M, u0, inv_m, seg_out ? seg_out : p->seg, &nseg);
    if (cap == 0) {
        st->status = kStNoAdvance;
        return;
    }
    if (cap > m.cap_limit) {
        st->status = kStCapacity;
        return;
    }
    p->inv_m = inv_m;
    p->nseg = nseg;
    p->cap = (uint32_t)cap;
}

__global__ void k_rs_plan(Member* __restrict__ members, float u0_arg)
{
    if (threadIdx.x != 0) return;
    plan_member(members[blockIdx.x], u0_arg);
}

__device__ __forceinline__ double wave_incl_scan(double v, int lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    return v;
}

// Span of scan workgroup b of B for M particles: whole chunks, the last
// workgroups possibly empty.
constexpr int kRsItems = 4;
constexpr uint32_t kRsChunk = kRsThreads * kRsItems;

__device__ __forceinline__ uint32_t rs_span(uint32_t M, uint32_t B)
{
    const uint32_t chunks = (M + kRsChunk - 1) / kRsChunk;
    return (chunks + B - 1) / B * kRsChunk;
}

__device__ __forceinline__ double block_sum_rs(double v, double* sh)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kRsThreads / 64; ++w) t += sh[w];
    __syncthreads();
    return t;  // valid in thread 0
}

// Phase A1 (grid x scan workgroups, y members; launched when a member's
// scan spans several workgroups): the weight total of each workgroup's
// span, in any order (it only seeds the candidate offsets of phase A2).
__global__ __launch_bounds__(kRsThreads) void k_rs_blk(const Member* __restrict__ members)
{
    const Member& mb = members[blockIdx.y];
    RsPlan* p = mb.plan;
    if (p->cap == 0) return;
    const uint32_t M = p->M;
    const uint32_t span = rs_span(M, gridDim.x);
    const uint32_t b0 = blockIdx.x * span, b1 = min(M, b0 + span);
    __shared__ double sh[kRsThreads / 64];
    double acc = 0.0;
    for (uint32_t m = b0 + threadIdx.x; m < b1; m += kRsThreads) acc += (double)mb.particles[m].second;
    const double t = block_sum_rs(acc, sh);
    if (threadIdx.x == 0) p->blk[blockIdx.x] = t;
}

// Phase A2 (grid x scan workgroups, y members): s_m, the double running
// sum, as a candidate: workgroup b scans its span from the sum of the
// earlier spans' totals, then checks s_m == s_{m-1} + w_m for every
// element of its span but the first (whose predecessor another workgroup
// writes; k_rs_count checks those and repairs serially from the first
// mismatch).
// Scan workgroup b of B (blockDim.x == kRsThreads).
__device__ void scan_span(const Member& mb, uint32_t b, uint32_t B)
{
    RsPlan* p = mb.plan;
    const uint32_t M = p->M;
    const uint32_t span = rs_span(M, B);
    const uint32_t lo = b * span;
    if (lo >= M) return;
    const uint32_t hi = min(M, lo + span);
    const tsl_particle* __restrict__ in = mb.particles;
    double* __restrict__ s = mb.cumsum;
    __shared__ double wave_tot[kRsThreads / 64];
    __shared__ double carry_sh;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    if (tid == 0) {
        double off = 0.0;
        for (uint32_t k = 0; k < b; ++k) off += p->blk[k];
        carry_sh = off;
    }
    __syncthreads();
    for (uint32_t base = lo; base < hi; base += kRsChunk) {
        double loc[kRsItems];
        double run = 0.0;
#pragma unroll
        for (int e = 0; e < kRsItems; ++e) {
            const uint32_t m = base + tid * kRsItems + e;
            const double w = m < hi ? (double)in[m].second : 0.0;
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
        for (int e = 0; e < kRsItems; ++e) {
            const uint32_t m = base + tid * kRsItems + e;
            if (m < hi) s[m] = excl + loc[e];
        }
        __syncthreads();
        if (tid == kRsThreads - 1) {
            double tot = 0.0;
            for (int w = 0; w < kRsThreads / 64; ++w) tot += wave_tot[w];
            carry_sh = carry + tot;
        }
        __syncthreads();
    }
    // check s_m == s_{m-1} + w_m inside the span (element 0: s_0 == 0 + w_0)
    for (uint32_t m = lo + tid; m < hi; m += kRsThreads) {
        if (m == lo && m != 0) continue;
        const double prev = m == 0 ? 0.0 : s[m - 1];
        if (s[m] != prev + (double)in[m].second) atomicMin(&p->s_bad, m);
    }
}

__global__ __launch_bounds__(kRsThreads) void k_rs_scan(const Member* __restrict__ members)
{
    const Member& mb = members[blockIdx.y];
    if (mb.plan->cap == 0) return;
    scan_span(mb, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ float u_candidate(const USeg* __restrict__ seg, uint32_t nseg, uint32_t j)
{
    uint32_t lo = 0, hi = nseg;  // last run with j0 <= j
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (seg[mid].j0 <= j) lo = mid; else hi = mid;
    }
    const USeg sg = seg[lo];
    return (float)((double)sg.u + (double)(j - sg.j0) * sg.step);
}

// Phase B (grid x candidates, y members): U_j candidates from the run
// table, each checked against one step of the recurrence from its
// predecessor.
__device__ __forceinline__ void u_step(const Member& mb, const RsPlan* p, uint32_t j)
{
    const float u = u_candidate(p->seg, p->nseg, j);
    mb.ubuf[j] = u;
    const float expect = j == 0 ? p->u0 : (float)((double)u_candidate(p->seg, p->nseg, j - 1) + p->inv_m);
    if (u != expect) atomicMin(&mb.st->first_bad, j);
}

__global__ void k_rs_u(const Member* __restrict__ members)
{
    const Member& mb = members[blockIdx.y];
    const RsPlan* p = mb.plan;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= p->cap) return;
    u_step(mb, p, j);
}

// Phase C (one thread per member): the running-sum checks left to it and
// serial repairs of s and U from their first mismatches, then the number
// emitted J = #{j : U_j < s_{M-1}} (U is non-decreasing).
// Run by the first wavefront (threads 0..63) of a workgroup.
// kSearch false (the one-workgroup form): stop after the capacity check and
// leave s_{M-1} in *s_final_out (thread 0); the caller counts J in parallel.
template <bool kSearch = true>
__device__ void count_member(const Member& mb, uint32_t scan_blocks, double* s_final_out = nullptr)
{
    RsPlan* p = mb.plan;
    const uint32_t cap = p->cap;
    if (cap == 0) return;
    FilterState* st = mb.st;
    const uint32_t M = p->M;
    const tsl_particle* __restrict__ in = mb.particles;
    double* __restrict__ sum = mb.cumsum;
    // running sum: the first element of each scan workgroup's span (one
    // wavefront, a span per lane), then a serial repair from the first
    // mismatch
    // (both first-mismatch words are read as atomics: a fused launch sets
    // them with atomics earlier in the same kernel)
    uint32_t bad = __hip_atomic_load(&p->s_bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    {
        const uint32_t span = rs_span(M, scan_blocks);
        for (uint32_t b = 1 + threadIdx.x; b < scan_blocks; b += 64) {
            const uint32_t m = b * span;
            if (m < M && sum[m] != sum[m - 1] + (double)in[m].second) bad = min(bad, m);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) bad = min(bad, (uint32_t)__shfl_xor((int)bad, off, 64));
    }
    if (threadIdx.x != 0) return;
    {
        if (bad < M) {
            double acc = bad == 0 ? 0.0 : sum[bad - 1];
            for (uint32_t m = bad; m < M; ++m) {
                acc += (double)in[m].second;
                sum[m] = acc;
            }
        }
    }
    float* __restrict__ U = mb.ubuf;
    const double inv_m = p->inv_m;
    const uint32_t fb = __hip_atomic_load(&st->first_bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (fb != 0xffffffffu) {
        float u = fb == 0 ? p->u0 : (float)((double)U[fb - 1] + inv_m);
        for (uint32_t j = fb; j < cap; ++j) {
            U[j] = u;
            u = (float)((double)u + inv_m);
        }
    }
    st->first_bad = 0xffffffffu;
    const double s_final = mb.cumsum[p->M - 1];
    if ((double)U[cap - 1] < s_final) {  // weights sum beyond 1 + 1e-5
        st->status = kStCapacity;
        p->cap = 0;
        return;
    }
    if constexpr (!kSearch) {
        *s_final_out = s_final;
        return;
    }
    uint32_t lo = 0, hi = cap;  // first j with U_j >= s_final
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((double)U[mid] < s_final) lo = mid + 1; else hi = mid;
    }
    st->resample_count = lo;
}

__global__ void k_rs_count(const Member* __restrict__ members, uint32_t scan_blocks)
{
    count_member(members[blockIdx.x], scan_blocks);
}

// Phase D (grid x outputs, y members): output j is particle
// min{m : s_m > U_j}, written to alt (and its index to idx_out).
__device__ __forceinline__ void assign_step(const Member& mb, const RsPlan* p, uint32_t j)
{
    const uint32_t M = p->M;
    const double* __restrict__ s = mb.cumsum;
    const double u = (double)mb.ubuf[j];
    uint32_t lo = 0, hi = M;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s[mid] > u) hi = mid; else lo = mid + 1;
    }
    const uint32_t m = lo < M ? lo : M - 1;
    if (mb.idx_out) mb.idx_out[j] = (int32_t)m;
    mb.alt[j] = mb.particles[m];
}

__global__ void k_rs_assign(const Member* __restrict__ members)
{
    const Member& mb = members[blockIdx.y];
    const RsPlan* p = mb.plan;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (p->cap == 0 || j >= mb.st->resample_count) return;
    assign_step(mb, p, j);
}

// Phase E (grid x outputs, y members): the emitted particles replace the
// resident set and its count.
__global__ void k_rs_copy_back(const Member* __restrict__ members)
{
    const Member& mb = members[blockIdx.y];
    if (mb.plan->cap == 0) return;
    const uint32_t J = mb.st->resample_count;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < J) mb.particles[j] = mb.alt[j];
    if (blockIdx.x == 0 && threadIdx.x == 0) mb.st->count = J;
}

// The whole resampling of one member in one workgroup (kRsThreads
// threads), for sets of up to kRsSmall particles and candidates: the phases
// above in order, separated by workgroup barriers instead of launches, on
// LDS copies of the run table, the weights, the running sum and U (the
// device-memory copies are not written; nothing reads them after a
// resampling).  Candidates are formed and checked as in the member kernels
// and repaired serially from the first mismatch, so s and U are the
// sequential recurrences and every output is the same; the particles move
// through registers instead of the alternate buffer when they replace the
// resident set.
constexpr int kRsSmallPer = (int)(kRsSmall / kRsThreads);  // particles and candidates per thread
static_assert(kRsSmall % kRsThreads == 0, "");

__global__ __launch_bounds__(kRsThreads) void k_rs_small(Member* __restrict__ members, float u0_arg, int copy_back)
{
    __shared__ USeg seg_sh[kMaxSeg];
    __shared__ double s_sh[kRsSmall];
    __shared__ float w_sh[kRsSmall];
    __shared__ float u_sh[kRsSmall];
    __shared__ double wave_tot[kRsThreads / 64];
    __shared__ uint32_t count_sh[kRsThreads / 64];
    __shared__ uint32_t s_bad_sh, u_bad_sh, go_sh, cap_sh, nseg_sh;
    __shared__ double inv_m_sh;
    __shared__ float u0_sh;
    const Member& mb = members[blockIdx.y];
    RsPlan* p = mb.plan;
    FilterState* st = mb.st;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // the weights are read while thread 0 plans the draw
    const uint32_t M = st->count;
    float w[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const uint32_t m = 2 * tid + e;
        w[e] = (m < M && M <= kRsSmall) ? mb.particles[m].second : 0.0f;
    }
    if (tid == 0) {
        s_bad_sh = 0xffffffffu;
        u_bad_sh = 0xffffffffu;
        plan_member(mb, u0_arg, seg_sh);
        uint32_t go = p->cap != 0;
        if (go && (M > kRsSmall || p->cap > kRsSmall)) {  // the launcher keeps both within kRsSmall
            st->status = kStCapacity;
            p->cap = 0;
            go = 0;
        }
        go_sh = go;
        cap_sh = p->cap;
        nseg_sh = p->nseg;
        inv_m_sh = p->inv_m;
        u0_sh = p->u0;
    }
    __syncthreads();
    if (!go_sh) return;
    const uint32_t cap = cap_sh, nseg = nseg_sh;
    const double inv_m = inv_m_sh;
    const float u0 = u0_sh;
    // running-sum candidates (one span, as k_rs_scan), checked against
    // s_m == s_{m-1} + w_m
    const double d0 = (double)w[0], d1 = (double)w[1];
    const double run = d0 + d1;
    const double incl = wave_incl_scan(run, (int)lane);
    if (lane == 63) wave_tot[wave] = incl;
    w_sh[2 * tid] = w[0];
    w_sh[2 * tid + 1] = w[1];
    __syncthreads();
    double woff = 0.0;
    for (uint32_t k = 0; k < wave; ++k) woff += wave_tot[k];
    const double excl = woff + (incl - run);
    s_sh[2 * tid] = excl + d0;
    s_sh[2 * tid + 1] = excl + run;
    // U candidates from the run table, checked against one step of the
    // recurrence
    for (uint32_t j = tid; j < cap; j += kRsThreads) {
        const float u = u_candidate(seg_sh, nseg, j);
        u_sh[j] = u;
        const float expect = j == 0 ? u0 : (float)((double)u_candidate(seg_sh, nseg, j - 1) + inv_m);
        if (u != expect) atomicMin(&u_bad_sh, j);
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const uint32_t m = 2 * tid + e;
        if (m < M) {
            const double prev = m == 0 ? 0.0 : s_sh[m - 1];
            if (s_sh[m] != prev + (double)w_sh[m]) atomicMin(&s_bad_sh, m);
        }
    }
    __syncthreads();
    if (tid == 0) {
        // serial repairs from the first mismatches, then the capacity check
        const uint32_t bad = s_bad_sh;
        if (bad < M) {
            double acc = bad == 0 ? 0.0 : s_sh[bad - 1];
            for (uint32_t m = bad; m < M; ++m) {
                acc += (double)w_sh[m];
                s_sh[m] = acc;
            }
        }
        const uint32_t fb = u_bad_sh;
        if (fb != 0xffffffffu) {
            float u = fb == 0 ? u0 : (float)((double)u_sh[fb - 1] + inv_m);
            for (uint32_t j = fb; j < cap; ++j) {
                u_sh[j] = u;
                u = (float)((double)u + inv_m);
            }
        }
        st->first_bad = 0xffffffffu;
        uint32_t go = 1;
        if ((double)u_sh[cap - 1] < s_sh[M - 1]) {  // weights sum beyond 1 + 1e-5
            st->status = kStCapacity;
            p->cap = 0;
            go = 0;
        }
        go_sh = go;
    }
    __syncthreads();
    if (!go_sh) return;
    // J = #{j : U_j < s_final} (U is non-decreasing)
    const double s_final = s_sh[M - 1];
    uint32_t below = 0;
    for (uint32_t j0 = 0; j0 < cap; j0 += kRsThreads) {
        const uint32_t j = j0 + tid;
        below += (uint32_t)__popcll(__ballot(j < cap && (double)u_sh[j] < s_final));
    }
    if (lane == 0) count_sh[wave] = below;
    __syncthreads();
    uint32_t J = 0;
    for (int k = 0; k < kRsThreads / 64; ++k) J += count_sh[k];
    // output j is particle min{m : s_m > U_j}, held in registers (its
    // seven floats)
    float out[kRsSmallPer][7];
#pragma unroll
    for (int r = 0; r < kRsSmallPer; ++r) {
        const uint32_t j = tid + (uint32_t)r * kRsThreads;
#pragma unroll
        for (int k = 0; k < 7; ++k) out[r][k] = 0.0f;
        if (j < J) {
            const double u = (double)u_sh[j];
            uint32_t lo = 0, hi = M;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_sh[mid] > u) hi = mid; else lo = mid + 1;
            }
            const uint32_t m = lo < M ? lo : M - 1;
            if (mb.idx_out) mb.idx_out[j] = (int32_t)m;
            const tsl_particle& src = mb.particles[m];
#pragma unroll
            for (int k = 0; k < 6; ++k) out[r][k] = src.first[k];
            out[r][6] = src.second;
            if (!copy_back) {
#pragma unroll
                for (int k = 0; k < 6; ++k) mb.alt[j].first[k] = out[r][k];
                mb.alt[j].second = out[r][6];
            }
        }
    }
    if (tid == 0) st->resample_count = J;
    if (!copy_back) return;
    __syncthreads();  // every source particle is read
#pragma unroll
    for (int r = 0; r < kRsSmallPer; ++r) {
        const uint32_t j = tid + (uint32_t)r * kRsThreads;
        if (j < J) {
#pragma unroll
            for (int k = 0; k < 6; ++k) mb.particles[j].first[k] = out[r][k];
            mb.particles[j].second = out[r][6];
        }
    }
    if (tid == 0) st->count = J;
}

}  // namespace

int resample_capacity(size_t M, float u0, size_t* cap_out)
{
    if (M == 0) {
        *cap_out = 0;
        return TSL_OK;
    }
    const uint64_t cap = cap_walk(M, u0, 1.0 / (double)M);
    if (cap == 0) return set_error(TSL_ERR_UNSUPPORTED, "systematic draw does not advance (M=%zu)", M);
    *cap_out = cap;
    return TSL_OK;
}

int launch_resample_members(Member* d_members, uint32_t n, uint32_t grid_cap, float u0_arg, bool copy_back,
                            hipStream_t stream)
{
    if (n == 0) return TSL_OK;
    if (n == 1 && grid_cap <= kRsSmall) {
        hipLaunchKernelGGL(k_rs_small, dim3(1, 1), dim3(kRsThreads), 0, stream, d_members, u0_arg, copy_back ? 1 : 0);
        TSL_HIP_CHECK(hipGetLastError());
        return TSL_OK;
    }
    const unsigned gx = (grid_cap + 255) / 256 ? (grid_cap + 255) / 256 : 1;
    hipLaunchKernelGGL(k_rs_plan, dim3(n), dim3(64), 0, stream, d_members, u0_arg);
    TSL_HIP_CHECK(hipGetLastError());
    // running-sum workgroups per member: one per 4096-particle chunk (the
    // largest member's), at most kMaxRsBlocks
    const uint32_t chunks = (grid_cap + kRsChunk - 1) / kRsChunk;
    const uint32_t nblk = chunks < 1 ? 1u : (chunks > (uint32_t)kMaxRsBlocks ? (uint32_t)kMaxRsBlocks : chunks);
    if (nblk > 1) {
        hipLaunchKernelGGL(k_rs_blk, dim3(nblk, n), dim3(kRsThreads), 0, stream, d_members);
        TSL_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_rs_scan, dim3(nblk, n), dim3(kRsThreads), 0, stream, d_members);
    TSL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_rs_u, dim3(gx, n), dim3(256), 0, stream, d_members);
    TSL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_rs_count, dim3(n), dim3(64), 0, stream, d_members, nblk);
    TSL_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_rs_assign, dim3(gx, n), dim3(256), 0, stream, d_members);
    TSL_HIP_CHECK(hipGetLastError());
    if (copy_back) {
        hipLaunchKernelGGL(k_rs_copy_back, dim3(gx, n), dim3(256), 0, stream, d_members);
        TSL_HIP_CHECK(hipGetLastError());
    }
    return TSL_OK;
}

}  // namespace tsl
