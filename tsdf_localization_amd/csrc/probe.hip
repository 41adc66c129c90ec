// Random-line gather probe: the rate at which the memory system returns
// independent 128-byte lines to wavefronts whose 64 lanes each read 2 bytes
// from a different line of a buffer.  It measures the ceiling the voxel-code
// gathers of the evaluation kernels work against: a buffer of the size of a
// map's code array (tens to hundreds of MB) is served from the L2 and the
// Infinity Cache, a buffer of many GB from HBM.
#include "tsl_internal.h"

namespace tsl {
namespace {

constexpr int kProbeThreads = 256;
constexpr int kProbeLoadsPerLane = 64;  // independent loads in flight per lane, in groups of 8

__global__ void k_probe_fill(uint16_t* __restrict__ buf, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) buf[i] = (uint16_t)i;
}

// Lane t of the grid reads kProbeLoadsPerLane 2-byte values at the start of
// pseudo-random lines (a multiplicative hash of its index and the load
// number, masked to the buffer's power-of-two line count) and stores their
// sum so the loads are kept.
__global__ __launch_bounds__(kProbeThreads) void k_probe_gather(const uint16_t* __restrict__ buf, uint64_t line_mask,
                                                                uint32_t seed, uint32_t* __restrict__ sink)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0;
#pragma unroll
    for (int g = 0; g < kProbeLoadsPerLane; g += 8) {
        uint32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint64_t h = (t * 0x9e3779b97f4a7c15ull) ^ ((uint64_t)(g + k + seed) * 0xc2b2ae3d27d4eb4full);
            const uint64_t line = (h ^ (h >> 29)) & line_mask;
            v[k] = buf[line * 64];  // 64 uint16 = 128 bytes per line
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) acc += v[k];
    }
    sink[t] = acc;
}

}  // namespace
}  // namespace tsl

using namespace tsl;

extern "C" int tsl_probe_gather(uint64_t buffer_bytes, uint32_t reps, double* lines_per_s, double* ms_per_launch)
{
    if (!lines_per_s || !ms_per_launch || reps == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "bad argument");
    uint64_t lines = 1;
    while (lines * 2 * 128 <= buffer_bytes) lines *= 2;  // a power of two of 128-byte lines
    if (lines < 1024) return set_error(TSL_ERR_INVALID_ARGUMENT, "buffer too small for the probe");
    const uint64_t n16 = lines * 64;
    const unsigned blocks = 256 * 8 * 16;  // 32768 workgroups of 256 lanes
    DevBuf buf, sink;
    TSL_TRY(buf.reserve(n16 * sizeof(uint16_t)));
    TSL_TRY(sink.reserve((size_t)blocks * kProbeThreads * sizeof(uint32_t)));
    hipLaunchKernelGGL(k_probe_fill, dim3(4096), dim3(256), 0, 0, buf.as<uint16_t>(), n16);
    TSL_HIP_CHECK(hipGetLastError());
    hipEvent_t a, b;
    TSL_HIP_CHECK(hipEventCreate(&a));
    TSL_HIP_CHECK(hipEventCreate(&b));
    hipLaunchKernelGGL(k_probe_gather, dim3(blocks), dim3(kProbeThreads), 0, 0, buf.as<uint16_t>(), lines - 1, 0u,
                       sink.as<uint32_t>());  // warm-up
    (void)hipEventRecord(a, 0);
    for (uint32_t r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_probe_gather, dim3(blocks), dim3(kProbeThreads), 0, 0, buf.as<uint16_t>(), lines - 1,
                           r + 1, sink.as<uint32_t>());
    (void)hipEventRecord(b, 0);
    const hipError_t e = hipEventSynchronize(b);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (e != hipSuccess) return set_error(TSL_ERR_HIP, "probe failed: %s", hipGetErrorString(e));
    const double per = ms / reps;
    *ms_per_launch = per;
    *lines_per_s = (double)blocks * kProbeThreads * kProbeLoadsPerLane / (per / 1000.0);
    return TSL_OK;
}

// Plain device-to-device copies of a buffer of known size, the reference
// point for scaling the FETCH_SIZE / WRITE_SIZE counters of a profile: each
// copy reads and writes `bytes`.
extern "C" int tsl_probe_copy(uint64_t bytes, uint32_t reps, double* ms_per_copy)
{
    if (!ms_per_copy || reps == 0 || bytes == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "bad argument");
    DevBuf a, b;
    TSL_TRY(a.reserve(bytes));
    TSL_TRY(b.reserve(bytes));
    hipLaunchKernelGGL(k_probe_fill, dim3(4096), dim3(256), 0, 0, a.as<uint16_t>(), bytes / 2);
    TSL_HIP_CHECK(hipGetLastError());
    hipEvent_t e0, e1;
    TSL_HIP_CHECK(hipEventCreate(&e0));
    TSL_HIP_CHECK(hipEventCreate(&e1));
    (void)hipEventRecord(e0, 0);
    for (uint32_t r = 0; r < reps; ++r) TSL_HIP_CHECK(hipMemcpyAsync(b.ptr, a.ptr, bytes, hipMemcpyDeviceToDevice, 0));
    (void)hipEventRecord(e1, 0);
    const hipError_t e = hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return set_error(TSL_ERR_HIP, "copy failed: %s", hipGetErrorString(e));
    *ms_per_copy = ms / reps;
    return TSL_OK;
}
