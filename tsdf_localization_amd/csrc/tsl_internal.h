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
constexpr uint32_t kMaxOccWords = 4096;    // occupancy-rank words kept in LDS (131072 blocks, 32 KB)

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
    bool owned = true;  // false: a view of memory someone else holds; never allocated or freed here
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (ptr && owned) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    // Point at n bytes of caller-held memory (no allocation, no free).
    void view(void* p, size_t n)
    {
        release();
        ptr = p;
        bytes = n;
        owned = false;
    }
    int reserve(size_t n)
    {
        if (n <= bytes && ptr) return TSL_OK;
        if (!owned)
            return set_error(TSL_ERR_CAPACITY, "a fixed buffer of %zu bytes cannot hold %zu bytes", bytes, n);
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

// One 64-block word of the occupancy bitmap with the number of allocated
// blocks before it.  When block offsets follow the reference's setData
// order (offset of the n-th allocated block in linear order = n * sub_dim^3,
// cuda_sub_voxel_map.tcc:202-211), grid_occ[up] == rank(up) * sub_dim^3 for
// allocated blocks, and rank(up) = prefix + popcount(bits below up).
// Sub-voxel bucket b of one axis: the in-block coordinate is count, or
// count + 1 when the fraction's bit pattern is >= thr; c_lo / c_hi are that
// axis's contribution to the voxel's physical byte offset inside its block
// for those two coordinates (a coordinate equal to sub_dim contributes
// kCarryPoison: such lanes recompute their address).
struct AxisEntry {
    uint32_t thr, c_lo, c_hi, count;
};
constexpr uint32_t kCarryPoison = 1u << 30;

struct OccWord {
    uint32_t bits;     // bit j: block 32w + j allocated
    uint32_t prefix;   // allocated blocks in words 0 .. w-1
};

// Every field the exact tiers of the lookup need (64-bit reference sizes,
// the wrap-band constants).  Lives in device memory; the kernels read it
// only when a wavefront leaves the fast tier.
// Wide maps (block offsets past the reference's int32, for maps of more than
// 2^31 voxels) keep, instead of the offset, each block's slot: offset /
// sub_dim^3, or kNoSlot for an empty block.
constexpr uint32_t kNoSlot = 0xffffffffu;

struct DevMapFull {
    const int32_t* grid_occ;
    const uint32_t* slots;            // wide maps: block slot or kNoSlot (grid_occ unused)
    uint64_t sd3;                     // sub_dim^3
    uint32_t wide, pad0;
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
    const uint32_t* slots;     // wide maps: block slot (offset / sub_dim^3) or kNoSlot
    const uint16_t* codes;     // per voxel: byte offset (4 * palette code) of its value in the palette;
                               // float form (f32): the float values themselves, at twice the byte offsets
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
    uint32_t sd3b;                    // 2 sub_dim^3: bytes of one block's codes
    uint32_t brick_cy, brick_cz;      // 64B - 16, 64B^2 - 64
    // per-axis sub-voxel buckets with the physical-address contribution of
    // both in-block coordinates a bucket can give: axis_tab[a * sub_entries + b]
    const AxisEntry* axis_tab;
    uint32_t axis_n;                  // 3 * sub_entries
    uint32_t carry_limit;             // fast-tier byte address from which a lane recomputes (kCarryPoison; 0 = every lane)
    uint32_t mag_sd, sh_sd, mag_sd2, sh_sd2, mag_sd3, sh_sd3;  // x / d = umulhi(x, mag) >> sh, x < 2^31
    // occupancy-rank form of the block table (set when the offsets follow
    // setData order and the bitmap fits the LDS budget)
    const OccWord* occ_words;
    uint32_t occ_nwords;              // 0: use grid_occ; else a power of two (zero words past the blocks)
    uint32_t occ_log2;                // log2(occ_nwords): block up reads word bits [5, 5 + occ_log2) of up
    uint32_t occ_used;                // words staged in LDS: every block index inside the grid lies in
                                      // them; the zero words past them are left out (a lane that forms a
                                      // word index there fails the grid check and is discarded)
    uint64_t sentinel;                // element of codes holding the code of init_value
    uint32_t sentinel_code;           // that byte offset
    uint32_t wide;                    // 1: slots instead of grid_occ, 64-bit voxel and byte offsets
    uint32_t f32;                     // 1: float form (more distinct values than the palette holds)
};

// Smallest u in [0, 2^23] with trunc(fl(fl(u)/res)) >= dim; host IEEE division.
uint32_t grid_threshold(uint64_t dim, float res);

// tables.hip
int build_sub_table(float res, SubTable* out);
void magic_divisor(uint32_t d, uint32_t* mag, uint32_t* sh);
uint32_t sub_lookup_host(const SubTable& t, float a);
void negative_band_constants(const tsl_map_coef& c, DevMapFull& d);

// Device-side bookkeeping of a filter (one per filter, in device memory).
// status values
constexpr uint32_t kStOk = 0, kStNoValid = 1, kStCapacity = 2, kStEmpty = 3, kStNoAdvance = 4;

struct FilterState {
    uint32_t count;        // particles in the resident set
    uint32_t status;       // kSt*
    uint32_t resample_count;
    uint32_t first_bad;    // resampling: first U candidate that failed its check
    uint32_t pad[4];
    double stats[10];      // sum, Σwx, Σwy, Σwz, Σw sin/cos roll, pitch, yaw
};

// Systematic-draw plan of one filter, built on the device from its count
// and u0: U_j run table (see resample.hip) and the number of candidates.
constexpr int kMaxSeg = 128;
struct USeg {
    uint32_t j0;   // first index of the run
    float u;       // U_{j0}
    double step;   // constant difference inside the run (exact)
};
constexpr int kMaxRsBlocks = 256;  // workgroups of one member's running-sum scan
struct RsPlan {
    uint32_t M;     // particles resampled
    uint32_t cap;   // candidates U_0 .. U_{cap-1}; 0 = skip this resampling
    uint32_t nseg;
    float u0;
    double inv_m;
    uint32_t s_bad;  // first running-sum element that failed its check
    uint32_t pad;
    USeg seg[kMaxSeg];
    double blk[kMaxRsBlocks];  // weight total of each scan workgroup's span
};

// Device descriptor of one filter for the member-indexed kernels (a single
// filter is a batch of one).
struct Member {
    tsl_particle* particles;   // resident set (in/out)
    tsl_particle* alt;         // resampling output before it is copied back
    float* L;                  // pose matrices
    const float4* points;      // prepared scan
    float* partial;            // chunk partial weights [chunks][stride]
    float* raw;                // raw weights
    const uint32_t* inv;       // slot of particle i in the evaluation order (NULL: identity)
    double* blocks;            // reduction block partials
    FilterState* st;
    double* cumsum;            // resampling running sum
    float* ubuf;               // resampling U candidates
    int32_t* idx_out;          // resampled source indices (optional)
    RsPlan* plan;
    const float* u0_dev;       // uniform draw in device memory (or NULL: kernel argument)
    const tsl_particle* snapshot;  // restore source (batched benchmarks)
    float tf[16];              // scanner -> base
    uint32_t N, chunk_len, chunks, stride;
    uint32_t cap_limit;        // room of particles/alt/ubuf for resampling output
    uint32_t snapshot_count;
    int32_t trig;              // TrigMode of the pose matrices and pose sums
    uint32_t pad;
};

constexpr uint32_t kMaxMembers = 1024;  // filters in one batch

// Doubles of reduction partials per evaluation workgroup of a member: the
// weight sum and the nine pose sums.
constexpr size_t kReducePartials = 10;

// Work plan of a batched evaluation, built on the device each step from the
// members' device counts: member f owns items [prefix[f], prefix[f+1]), one
// item = kWaves*K particles x one chunk of its scan.  The workgroups of XCD
// x claim items x, x + 8, x + 16, ... from counter next[x * 32] (counters
// 128 B apart; reset by the plan kernel).
constexpr uint32_t kPlanXcds = 8;
struct BatchPlan {
    uint32_t n, total, pad[30];
    uint32_t next[kPlanXcds * 32];
    uint32_t prefix[kMaxMembers + 1];
};

// Upper bound on the resampling output for M particles whose weights sum
// to at most 1 + 1e-5: U_j >= j (1/M - 2^-24) for every u0 (each float step
// loses at most half an ulp below 2), so at most (1 + 1e-5) / (1/M - 2^-24)
// + 2 candidates are needed.  0 if M >= 2^24 (the draw may not advance).
inline uint64_t resample_bound(uint64_t M)
{
    if (M == 0) return 0;
    const double d = 1.0 / (double)M - 5.9604644775390625e-08;
    if (!(d > 0.0)) return 0;
    return (uint64_t)((1.0 + 1e-5) / d) + 3;  // one more for the rounding of this quotient
}

}  // namespace tsl

// ---------------------------------------------------------------------------
// Handle types
// ---------------------------------------------------------------------------
struct tsl_map {
    int device = 0;
    tsl_map_coef coef{};
    tsl::DevBuf grid_occ, codes, palette, sub_tab, neg_tab, full, occ_words, axis_tab;
    tsl::DevBuf slots;       // wide maps: uint32 block slots (grid_occ unused)
    tsl::DevMap dev{};
};

struct tsl_evaluator {
    tsl_map* map = nullptr;
    int device = 0;
    int per_point = 0;
    float a_hit = 0.9f, a_range = 0.1f, a_max = 0.0f, max_range = 100.0f;
    float inv_max_range = 0.01f, max_range_squared = 10000.0f, range_term = 0.0f;
    int trig = tsl::kTrigGlibcFma;  // sin/cos of the particle angles (TrigMode)
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
    int tile_p = 0;            // points per lane per iteration (0 = automatic)
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
    tsl::DevBuf pkeys;       // particle-order buckets: counts/offsets, then bucket of particle i
    tsl::DevBuf red_scratch;          // scan reduction
    tsl::DevBuf raw_in, raw_ring;     // raw scan staging
    tsl::DevBuf porder;      // uint32: particle evaluated in slot s
    tsl::DevBuf pinv;        // uint32: slot of particle i
    int order_mode = -1;     // particle evaluation order: 0 caller, 1 spatial, -1 automatic
    int last_order = 0;      // order of the last evaluation (partial weights by slot)
    int rank_mode = -1;      // block offsets: 0 global table, 1 LDS rank table, -1 automatic
    tsl::DevBuf blocks;      // double block partials
    tsl::DevBuf state;       // FilterState
    tsl::DevBuf cumsum;      // double[capacity]
    tsl::DevBuf ubuf;        // float U table
    tsl::DevBuf plan;        // RsPlan
    tsl::DevBuf motion;      // tsl_motion of the pending motion update
    void* motion_host = nullptr;    // pinned staging
    void* motion_copied = nullptr;  // hipEvent_t: staging free again
    bool motion_pending = false;
    tsl::DevBuf member;      // Member (device descriptor of this filter)
    tsl::Member member_host{};  // last uploaded descriptor
    bool member_dirty = true;
    bool member_on_stream = false;  // write the descriptor with a kernel on the launch stream, every time
    void* host_state = nullptr;  // pinned FilterState mirror
    void* weights_host = nullptr;  // pinned read-back of the normalised weights (drop-in evaluation)
    size_t weights_host_cap = 0;   // floats it holds
};

namespace tsl {

// Kernel launchers (eval.hip)
int launch_update(tsl_filter* f, hipStream_t stream);
int launch_raw(tsl_filter* f, hipStream_t stream);
// Set the device particle count of a filter with a kernel on `stream`.
int launch_set_count(tsl_filter* f, uint32_t n, hipStream_t stream);
int launch_trace(tsl_filter* f, hipStream_t stream, void* out_device, int values);
int launch_restore(tsl_filter* f, hipStream_t stream);
int autotune_tile(tsl_filter* f, hipStream_t stream, int reps, int* best_order, int* best_k, int* best_p,
                  uint32_t* best_chunk, float* best_ms);

struct Tile {
    int K;               // particles per wavefront
    int P;               // points per lane per iteration
    int order;           // 1: particles evaluated in spatial order
    int rank;            // 1: block offsets from the LDS occupancy-rank table (when the map has one)
    uint32_t chunk_len;
    uint32_t chunks;
};
int launch_pose_matrices(const tsl_particle* d_particles, uint32_t n, const float tf[16], int trig,
                         float* d_out, hipStream_t stream);
// sin and cos of n angles on the device in a TrigMode.
int launch_trig(const float* d_x, uint32_t n, int mode, float* d_s, float* d_c, hipStream_t stream);
int launch_prep_points(tsl_filter* f, hipStream_t stream);
// Temporary storage the point sort of launch_prep_points needs for n points.
size_t prep_points_temp_bytes(uint32_t n);
// Buffers of the spatial particle order for f->count_bound particles.
int reserve_particle_order(tsl_filter* f);
// Spatial particle order of a filter and its pose matrices in that order
// (f->L slot s = particle f->porder[s], f->pinv the inverse), for a filter
// evaluated as a batch member whose inv points at f->pinv.
int launch_pose_spatial(tsl_filter* f, hipStream_t stream);

// Scan preparation of many scans at once (a serving tick): scan r is
// points [off[r], off[r+1]) of pts; each is Morton-sorted within its own
// segment exactly as launch_prep_points sorts a filter's scan (same keys,
// stable order) and written to dst[r].
struct ScanDest {
    float4* points;  // sorted points with their score constant
    uint32_t* perm;  // sorted position -> index in the scan
};
size_t prep_points_batch_temp(uint32_t total, uint32_t n);
int launch_prep_points_batch(const tsl_evaluator* ev, const tsl_point* pts, const uint32_t* off, const ScanDest* dst,
                             uint32_t n, uint32_t total, uint32_t max_count, uint64_t* keys, uint32_t* vals,
                             void* temp, size_t temp_bytes, hipStream_t stream);
int launch_particle_order(tsl_filter* f, hipStream_t stream);
int filter_reserve_scan(tsl_filter* f, size_t n);
// Scan reductions: kReduceEvaluator is TSDFEvaluator::evaluateParticles
// (src/evaluation/tsdf_evaluator.cpp:304-376; ring int16), kReduceCloud the
// centre reduction of CudaEvaluator::evaluate(PointCloud2)
// (src/cuda/cuda_evaluator.cu:78-116; ring int32, cell fixed at 0.064 m).
constexpr int kReduceEvaluator = 0;
constexpr int kReduceCloud = 1;
int launch_reduce_scan(tsl_filter* f, int mode, const tsl_point* raw, const void* ring, uint32_t n, float cell,
                       tsl_point* out, uint32_t* count_dev, uint32_t* bad_dev, hipStream_t stream);

// Resampling (resample.hip): members[0..n) resampled in one sequence of
// launches; grid_cap = max cap_limit over the members.
int resample_capacity(size_t M, float u0, size_t* cap);
int launch_resample_members(Member* d_members, uint32_t n, uint32_t grid_cap, float u0_arg, bool copy_back,
                            hipStream_t stream);
// Pose estimate from the ten weighted sums of a FilterState (api.hip).
void pose_from_stats(const double* st, tsl_pose* pose);
// Map a FilterState status to the library's error (TSL_OK if none).
int status_error(uint32_t status);

// Batched update (eval.hip)
struct BatchLaunch {
    const tsl_map* map = nullptr;
    float a_hit = 0.9f;
    Member* members = nullptr;   // device array [n]
    BatchPlan* plan = nullptr;   // device
    uint32_t n = 0;
    int K = 1, P = 1;
    uint64_t max_items = 0;      // upper bound of the evaluation items over all members
    uint32_t max_cap = 0;        // largest member capacity (grid bound of the per-particle kernels)
    uint32_t max_rs = 0;         // largest resampling output room
    uint32_t grid = 0;           // persistent evaluation grid (multiple of the XCD count)
    bool rank = false;           // block offsets from the LDS occupancy-rank table
    uint64_t slice_items = 0;    // evaluation items per launch (0: one launch for all)
};
int launch_batch_update(const BatchLaunch& b, hipStream_t stream);
int batch_grid(const BatchLaunch& b, uint32_t* grid);
// Automatic chunk length of a batch member with an N-point scan.
uint32_t batch_chunk(size_t N);
int launch_batch_restore(const BatchLaunch& b, hipStream_t stream);
Tile choose_tile(const tsl_filter* f, size_t P, size_t N);

// Upload the filter's device descriptor if it changed; chunk_len/chunks
// describe the partial-weight layout of the last evaluation (0 = keep).
int filter_member(tsl_filter* f, Member** d_member, uint32_t chunk_len = 0, uint32_t chunks = 0,
                  hipStream_t stream = nullptr);
// Write a Member descriptor to device memory with a one-thread kernel on
// `stream` (ordered with the stream's other work, no host copy).
int launch_store_member(const Member& m, Member* dst, hipStream_t stream);

// Motion update of members[0..n) with motions[0..n) (motion.hip).
int launch_motion_members(const Member* d_members, const tsl_motion* d_motions, uint32_t n, uint32_t grid_cap,
                          hipStream_t stream);

// Filter buffer management (api.hip)
int filter_reserve_particles(tsl_filter* f, size_t n);
int filter_reserve_points(tsl_filter* f, size_t n);

}  // namespace tsl
