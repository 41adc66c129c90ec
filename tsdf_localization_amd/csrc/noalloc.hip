// Allocation-free raw evaluation: tsl_evaluate_raw for device-resident
// inputs, for callers that cannot allocate device memory (code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc).
//
// Every intermediate of the evaluation -- filter state, pose matrices, the
// Morton-sorted scan with its sort keys and radix-sort storage, the chunk
// partial weights, the reduction partials and the member descriptor -- is
// carved out of one file-scope device array, `scratch`, of kScratchBytes.
// Each GPU holds its own copy for the life of the process; host code reaches
// it only through the address hipGetSymbolAddress returns on that device
// (kept per device), never through the array's name.  The inputs are read
// in place and the raw weights are written straight to the caller's buffer.
// All work is enqueued on the caller's stream; the descriptor is written by
// a kernel on that stream, so a call allocates nothing, copies nothing from
// the host and does not synchronise.  Calls that overlap on one device
// share the array and race.
#include "tsl_internal.h"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace tsl {
namespace {

constexpr size_t kScratchBytes = 64ull << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array, looked up once per device
// with the device current.
int scratch_address(int device, void** out)
{
    if (device < 0 || device >= kMaxDevices) return set_error(TSL_ERR_UNSUPPORTED, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(TSL_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                             hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return TSL_OK;
}

__global__ void k_fill_zero(float* __restrict__ w, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = 0.0f;
}

inline size_t round256(size_t n) { return (n + 255) / 256 * 256; }

}  // namespace
}  // namespace tsl

using namespace tsl;

extern "C" int tsl_evaluate_raw_noalloc(tsl_evaluator* ev, const tsl_particle* particles, size_t P,
                                        const tsl_point* points, size_t N, const float tf[16], float* weights_out,
                                        void* stream)
{
    if (!ev || !tf || (P && (!particles || !weights_out)) || (N && !points))
        return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (P > 0x7fffffffull || N > 0x7fffffffull) return set_error(TSL_ERR_UNSUPPORTED, "too many particles or points");
    if (P == 0) return TSL_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    TSL_HIP_CHECK(hipSetDevice(ev->device));
    if (N == 0) {
        hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, weights_out, (uint32_t)P);
        TSL_HIP_CHECK(hipGetLastError());
        return TSL_OK;
    }

    // A filter whose buffers are views: the caller's particles, points and
    // weights, and slices of the scratch array.
    tsl_filter f;
    f.ev = ev;
    f.device = ev->device;
    f.capacity = P;
    f.count_bound = P;
    f.count_known = true;
    f.num_points = N;
    std::memcpy(f.tf, tf, sizeof(f.tf));
    f.order_mode = 0;  // caller order: the spatial order's buckets are not needed (same results in any order)
    f.member_on_stream = true;
    const Tile tile = choose_tile(&f, P, N);
    const size_t nb = (P + 255) / 256;
    const size_t sizes[] = {
        sizeof(FilterState),                              // state
        sizeof(Member),                                   // member
        sizeof(RsPlan),                                   // plan
        sizeof(float) * 12 * P,                           // pose matrices
        sizeof(float) * (size_t)tile.chunks * P,          // chunk partial weights
        sizeof(double) * kReducePartials * nb,            // reduction partials
        sizeof(float4) * N,                               // sorted points with score constants
        sizeof(uint32_t) * N,                             // sorted -> caller index
        sizeof(uint64_t) * 2 * N,                         // sort keys
        sizeof(uint32_t) * N,                             // sort values
        prep_points_temp_bytes((uint32_t)N),              // radix-sort storage
    };
    size_t need = 0;
    for (size_t b : sizes) need += round256(b);
    if (need > kScratchBytes)
        return set_error(TSL_ERR_CAPACITY,
                         "tsl_evaluate_raw_noalloc: %zu particles x %zu points need %zu bytes of scratch; "
                         "the fixed scratch holds %zu bytes",
                         P, N, need, kScratchBytes);
    void* base = nullptr;
    TSL_TRY(scratch_address(ev->device, &base));
    char* at = static_cast<char*>(base);
    DevBuf* views[] = {&f.state, &f.member, &f.plan, &f.L, &f.partial, &f.blocks,
                       &f.points, &f.perm, &f.sort_keys, &f.sort_vals, &f.sort_temp};
    for (size_t k = 0; k < sizeof(views) / sizeof(views[0]); ++k) {
        views[k]->view(at, round256(sizes[k]));
        at += round256(sizes[k]);
    }
    f.particles.view(const_cast<tsl_particle*>(particles), sizeof(tsl_particle) * P);
    f.points_in.view(const_cast<tsl_point*>(points), sizeof(tsl_point) * N);
    f.raw.view(weights_out, sizeof(float) * P);

    TSL_TRY(launch_set_count(&f, (uint32_t)P, s));
    TSL_TRY(launch_prep_points(&f, s));
    return launch_raw(&f, s);
}
