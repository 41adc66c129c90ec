// Internal declarations shared by the HIP translation units of libtsl.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/tsl.h"
#include "tsl_common.h"

namespace tsl {

constexpr uint32_t kMaxPalette = 4096;  // 16-bit voxel codes index at most this many values
constexpr uint32_t kFastLimit = 8388608u;  // 2^23: offsets below this take the 32-bit lookup

// ---------------------------------------------------------------------------
// Error state
// ---------------------------------------------------------------------------
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#define TSL_HIP_CHECK(expr)                                                              \
    do {                                                                                 \
        hipError_t tsl_e_ = (expr);                                                      \
        if (tsl_e_ != hipSuccess)                                                        \
            return ::tsl::set_error(TSL_ERR_HIP, "%s failed: %s (%s:%d)", #expr,          \
                                    hipGetErrorString(tsl_e_), __FILE__, __LINE__);      \
    } while (0)

#define TSL_TRY(expr)                  \
    do {                               \
        int tsl_rc_ = (expr);          \
        if (tsl_rc_ != TSL_OK)         \
            return tsl_rc_;            \
    } while (0)

// ---------------------------------------------------------------------------
// Grow-only device buffer (the reference's evaluator also keeps grow-only
// device buffers, src/cuda/cuda_evaluator.cu:136-209).
// ---------------------------------------------------------------------------
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    int reserve(size_t n)
    {
        if (n <= bytes && ptr) return TSL_OK;
        release();
        size_t want = n ? n : 256;
        hipError_t e = hipMalloc(&ptr, want);
        if (e != hipSuccess) {
            ptr = nullptr;
            return set_error(e == hipErrorOutOfMemory ? TSL_ERR_OUT_OF_MEMORY : TSL_ERR_HIP,
                             "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        }
        bytes = want;
        return TSL_OK;
    }
    template <typename T> T* as() const { return static_cast<T*>(ptr); }
};

// ---------------------------------------------------------------------------
// Device view of a map: the reference's MapCoef fields plus the palette-coded
// voxel array and the exact fast-path thresholds derived on the host.
// ---------------------------------------------------------------------------
struct DevMap {
    const int32_t* grid_occ;   // block table: voxel offset of the block or -1
    const uint16_t* codes;     // data_size palette codes, reference voxel order
    const float* palette;      // palette_size values
    uint32_t palette_size;
    float min_x, min_y, min_z;
    float res;
    float init;
    uint32_t gthr_x, gthr_y, gthr_z;  // u < gthr  <=>  trunc(fl(u/res)) < dim, for u < 2^23
    uint64_t dim_x, dim_y, dim_z;
    uint64_t up_dim_x, up_dim_2;
    uint64_t sub_dim, sub_dim_2;
    uint64_t grid_occ_size, data_size;
};

// Smallest u in [0, 2^23] with trunc(fl(fl(u)/res)) >= dim; host IEEE division.
uint32_t grid_threshold(uint64_t dim, float res);

}  // namespace tsl

// ---------------------------------------------------------------------------
// Handle types
// ---------------------------------------------------------------------------
struct tsl_map {
    int device = 0;
    tsl_map_coef coef{};
    tsl::DevBuf grid_occ, codes, palette;
    tsl::DevMap dev{};
};

struct tsl_evaluator {
    tsl_map* map = nullptr;
    int device = 0;
    int per_point = 0;
    float a_hit = 0.9f, a_range = 0.1f, a_max = 0.0f, max_range = 100.0f;
    float inv_max_range = 0.01f, max_range_squared = 10000.0f, range_term = 0.0f;
    tsl_filter* scratch = nullptr;  // backs the host-buffer entry points
};

struct tsl_filter {
    tsl_evaluator* ev = nullptr;
    int device = 0;
    size_t capacity = 0;     // particles the buffers hold
    size_t count_bound = 0;  // host-side upper bound of the device count
    size_t num_points = 0;
    float tf[16] = {};
    tsl::DevBuf particles;   // tsl_particle[capacity]
    tsl::DevBuf particles_alt;
    tsl::DevBuf L;           // float[12 * capacity]
    tsl::DevBuf raw;         // float[capacity]
    tsl::DevBuf partial;     // float[chunks * capacity]
    tsl::DevBuf points;      // float4[num_points]: x, y, z, range term
    tsl::DevBuf points_in;   // tsl_point staging
    tsl::DevBuf blocks;      // double block partials
    tsl::DevBuf state;       // FilterState
    tsl::DevBuf cumsum;      // double[capacity]
    tsl::DevBuf ubuf;        // float U table
    void* host_state = nullptr;  // pinned FilterState mirror
};

namespace tsl {

// Device-side bookkeeping of a filter (one per filter, in device memory).
struct FilterState {
    uint32_t count;        // particles in the resident set
    uint32_t status;       // 0 ok, 1 zero weight sum, 2 resample capacity exceeded
    uint32_t counter0;     // last-block-done counters, reset by the last block
    uint32_t counter1;
    uint32_t resample_count;
    uint32_t pad[3];
    double stats[10];      // sum, Σwx, Σwy, Σwz, Σw sin/cos roll, pitch, yaw
};

// Kernel launchers (eval.hip)
int launch_update(tsl_filter* f, hipStream_t stream);
int launch_raw(tsl_filter* f, hipStream_t stream, float* raw_out_device);
int launch_trace(tsl_filter* f, hipStream_t stream, uint64_t* out_device);
int launch_pose_matrices(const tsl_particle* d_particles, uint32_t n, const float tf[16],
                         float* d_out, hipStream_t stream);
int launch_prep_points(tsl_filter* f, hipStream_t stream);

// Resampling (resample.hip)
struct USegments;
int resample_capacity(size_t M, float u0, size_t* cap);
int launch_resample(tsl_filter* f, float u0, hipStream_t stream);

// Filter buffer management (api.hip)
int filter_reserve_particles(tsl_filter* f, size_t n);
int filter_reserve_points(tsl_filter* f, size_t n);

}  // namespace tsl
