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
constexpr uint32_t kMaxSubTable = 1024;    // buckets of the sub-voxel threshold table

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
struct SubEntry {
    uint32_t thr;    // bit pattern of the threshold inside the bucket (+inf if none)
    uint32_t count;  // thresholds below the bucket
};

struct SubTable {
    uint32_t shift = 23;
    uint32_t base_bits = 0;
    std::vector<SubEntry> entries;
};

// Every field the exact tiers of the lookup need (64-bit reference sizes,
// the wrap-band constants).  Lives in device memory; the kernels read it
// only when a wavefront leaves the fast tier.
struct DevMapFull {
    const int32_t* grid_occ;
    float res;
    uint32_t gneg_x, gneg_y, gneg_z;  // grid check result of an axis offset in (-2^23, -1]
    uint64_t dim_x, dim_y, dim_z;
    uint64_t up_dim_x, up_dim_2;
    uint64_t sub_dim, sub_dim_2;
    uint64_t grid_occ_size, data_size;
    uint64_t neg_x, neg_y, neg_z;     // index contribution of an axis offset in (-2^23, -1]
    uint32_t gthr_x, gthr_y, gthr_z;
};

// Kernel-argument view of a map: what the fast tier reads every iteration.
struct DevMap {
    const int32_t* grid_occ;   // block table: voxel offset of the block or -1
    const uint16_t* codes;     // data_size palette codes, reference voxel order
    const float* palette;      // palette_size values
    const SubEntry* sub_tab;   // sub-voxel threshold buckets
    const uint64_t* neg_tab;   // index contribution of the wrap-band axes, by axis bit mask (8)
    const DevMapFull* full;    // device copy of the exact-tier fields
    uint32_t palette_size, sub_entries, sub_shift, sub_base;
    float min_x, min_y, min_z;
    float init;
    uint32_t gthr_x, gthr_y, gthr_z;  // u < gthr  <=>  trunc(fl(u/res)) < dim, for u < 2^23
    uint32_t gneg;                    // bit a: grid check passes for an axis-a offset in (-2^23, -1]
    uint32_t range_bits;              // bit pattern of R: |offset| < R keeps every block index in int32
    uint32_t up_dim_x, up_dim_2;      // 32-bit copies (block table < 2^31 entries)
    uint32_t occ, ds;                 // grid_occ_size, data_size (< 2^31)
    int32_t sub_dim, sub_dim_2;
    uint64_t data_size;               // == ds, for the index comparisons of the exact tiers
    // 4x4x4 brick layout of the codes (set when block offsets are aligned
    // multiples of sub_dim^3 and sub_dim % 4 == 0): physical address of a
    // voxel with reference index idx = slot*sd3 + lx + ly*sd + lz*sd^2 is
    // slot*sd3 + ((lx>>2) + (ly>>2)*B + (lz>>2)*B^2)*64 + (lx&3) + (ly&3)*4 + (lz&3)*16.
    uint32_t brick;                   // 1: bricked codes
    uint32_t sd3;                     // sub_dim^3
    uint32_t brick_cy, brick_cz;      // 64B - 16, 64B^2 - 64
    uint32_t mag_sd, sh_sd, mag_sd2, sh_sd2, mag_sd3, sh_sd3;  // x / d = umulhi(x, mag) >> sh, x < 2^31
};

// Smallest u in [0, 2^23] with trunc(fl(fl(u)/res)) >= dim; host IEEE division.
uint32_t grid_threshold(uint64_t dim, float res);

// tables.hip
int build_sub_table(float res, SubTable* out);
void magic_divisor(uint32_t d, uint32_t* mag, uint32_t* sh);
uint32_t sub_lookup_host(const SubTable& t, float a);
void negative_band_constants(const tsl_map_coef& c, DevMapFull& d);

}  // namespace tsl

// ---------------------------------------------------------------------------
// Handle types
// ---------------------------------------------------------------------------
struct tsl_map {
    int device = 0;
    tsl_map_coef coef{};
    tsl::DevBuf grid_occ, codes, palette, sub_tab, neg_tab, full;
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
    bool count_known = true;  // count_bound is the exact device count
    size_t snapshot_count = 0;
    tsl::DevBuf snapshot;    // tsl_particle[snapshot_count]
    bool timing = false;     // bracket the evaluation kernel with events
    void* ev_start = nullptr;
    void* ev_stop = nullptr;
    double eval_ms = 0.0;    // accumulated evaluation-kernel time
    uint64_t eval_launches = 0;
    uint64_t eval_pairs = 0;  // particle*point evaluations launched
    bool timing_pending = false;
    int tile_k = 0;            // particles per wavefront (0 = automatic)
    uint32_t tile_chunk = 0;   // scan points per wavefront (0 = automatic)
    size_t num_points = 0;
    float tf[16] = {};
    tsl::DevBuf particles;   // tsl_particle[capacity]
    tsl::DevBuf particles_alt;
    tsl::DevBuf L;           // float[12 * capacity]
    tsl::DevBuf raw;         // float[capacity]
    tsl::DevBuf partial;     // float[chunks * capacity]
    tsl::DevBuf points;      // float4[num_points]: x, y, z, range term
    tsl::DevBuf points_in;   // tsl_point staging (caller order)
    tsl::DevBuf perm;        // uint32[num_points]: caller index of sorted point j
    tsl::DevBuf sort_keys, sort_vals, sort_temp;
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
    uint32_t first_bad;    // resampling: first U candidate that failed its check
    uint32_t pad[2];
    double stats[10];      // sum, Σwx, Σwy, Σwz, Σw sin/cos roll, pitch, yaw
};

// Kernel launchers (eval.hip)
int launch_update(tsl_filter* f, hipStream_t stream);
int launch_raw(tsl_filter* f, hipStream_t stream, float* raw_out_device);
int launch_trace(tsl_filter* f, hipStream_t stream, void* out_device, int values);
int launch_restore(tsl_filter* f, hipStream_t stream);
int autotune_tile(tsl_filter* f, hipStream_t stream, int reps, int* best_k, uint32_t* best_chunk, float* best_ms);

struct Tile {
    int K;
    uint32_t chunk_len;
    uint32_t chunks;
};
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
