// C ABI of the MI355X sensor update: handles, map upload, host-buffer
// drop-in entry points (synchronous) and the device-resident filter API.
#include "tsl_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

static_assert(sizeof(tsl_particle) == 28, "Particle is 7 floats (particle.h:26-55)");
static_assert(sizeof(tsl_point) == 12, "CudaPoint is 3 floats (cuda_evaluator.h:22-39)");
static_assert(sizeof(tsl_map_coef) == 120, "MapCoef layout (cuda_sub_voxel_map.h:22-50)");
static_assert(offsetof(tsl_map_coef, min_x) == 24 && offsetof(tsl_map_coef, up_dim_x) == 56,
              "MapCoef field offsets");

namespace tsl {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

uint32_t grid_threshold(uint64_t dim, float res)
{
    // predicate is monotone in u: trunc(fl(fl(u)/res)) >= dim
    auto pass = [&](uint32_t u) { return cvt_f32_u64(cvt_u64_f32(u) / res) >= dim; };
    uint32_t lo = 0, hi = kFastLimit;  // answer in [lo, hi]; hi means "none below 2^23"
    if (!pass(hi - 1)) return hi;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (pass(mid)) hi = mid; else lo = mid + 1;
    }
    return lo;
}

int launch_set_count(tsl_filter* f, uint32_t n, hipStream_t stream);

int filter_reserve_particles(tsl_filter* f, size_t n)
{
    if (n <= f->capacity && f->particles.ptr) return TSL_OK;
    // keep the resident set when growing
    DevBuf grown;
    TSL_TRY(grown.reserve(sizeof(tsl_particle) * n));
    if (f->particles.ptr && f->count_bound)
        TSL_HIP_CHECK(hipMemcpy(grown.ptr, f->particles.ptr, sizeof(tsl_particle) * f->count_bound,
                                hipMemcpyDeviceToDevice));
    std::swap(grown.ptr, f->particles.ptr);
    std::swap(grown.bytes, f->particles.bytes);
    TSL_TRY(f->particles_alt.reserve(sizeof(tsl_particle) * n));
    TSL_TRY(f->L.reserve(sizeof(float) * 12 * n));
    TSL_TRY(f->raw.reserve(sizeof(float) * n));
    TSL_TRY(f->cumsum.reserve(sizeof(double) * n));
    f->partial.release();
    f->capacity = n;
    return TSL_OK;
}

int filter_reserve_points(tsl_filter* f, size_t n)
{
    TSL_TRY(f->points.reserve(16 * (n ? n : 1)));
    TSL_TRY(f->points_in.reserve(sizeof(tsl_point) * (n ? n : 1)));
    return TSL_OK;
}

static FilterState* host_state(tsl_filter* f) { return static_cast<FilterState*>(f->host_state); }

// Device descriptor of the filter, uploaded when any of its buffers or its
// scan changed since the last upload.
int filter_member(tsl_filter* f, Member** d_member, uint32_t chunk_len, uint32_t chunks, hipStream_t stream)
{
    TSL_TRY(f->plan.reserve(sizeof(RsPlan)));
    TSL_TRY(f->member.reserve(sizeof(Member)));
    Member h;
    std::memset(&h, 0, sizeof(h));
    h.particles = f->particles.as<tsl_particle>();
    h.alt = f->particles_alt.as<tsl_particle>();
    h.L = f->L.as<float>();
    h.points = f->points.as<float4>();
    h.partial = f->partial.as<float>();
    h.raw = f->raw.as<float>();
    h.inv = f->last_order ? f->pinv.as<uint32_t>() : nullptr;
    h.blocks = f->blocks.as<double>();
    h.st = f->state.as<FilterState>();
    h.cumsum = f->cumsum.as<double>();
    h.ubuf = f->ubuf.as<float>();
    h.idx_out = nullptr;
    h.plan = f->plan.as<RsPlan>();
    h.u0_dev = nullptr;
    std::memcpy(h.tf, f->tf, sizeof(h.tf));
    h.trig = f->ev->trig;
    h.N = (uint32_t)f->num_points;
    h.chunk_len = chunks ? chunk_len : f->member_host.chunk_len;
    h.chunks = chunks ? chunks : f->member_host.chunks;
    h.stride = (uint32_t)f->capacity;
    h.snapshot = f->snapshot.as<tsl_particle>();
    h.snapshot_count = (uint32_t)f->snapshot_count;
    const size_t ucap = f->ubuf.bytes / sizeof(float);
    h.cap_limit = (uint32_t)std::min(f->capacity, ucap);
    if (f->member_on_stream) {
        TSL_TRY(launch_store_member(h, f->member.as<Member>(), stream));
        f->member_host = h;
    } else if (f->member_dirty || std::memcmp(&h, &f->member_host, sizeof(h)) != 0) {
        TSL_HIP_CHECK(hipMemcpy(f->member.ptr, &h, sizeof(h), hipMemcpyHostToDevice));
        f->member_host = h;
        f->member_dirty = false;
    }
    *d_member = f->member.as<Member>();
    return TSL_OK;
}

int status_error(uint32_t status)
{
    switch (status) {
    case kStOk: return TSL_OK;
    case kStNoValid: return set_error(TSL_ERR_NO_VALID_PARTICLE, "No particle is valid!");
    case kStCapacity:
        return set_error(TSL_ERR_CAPACITY, "resampling: weights sum beyond 1 + 1e-5 or output beyond capacity");
    case kStEmpty: return set_error(TSL_ERR_INVALID_ARGUMENT, "filter holds no particles");
    case kStNoAdvance: return set_error(TSL_ERR_UNSUPPORTED, "systematic draw does not advance");
    default: return set_error(TSL_ERR_HIP, "unknown filter status %u", status);
    }
}

// pose from the ten sums: weighted mean position, circular mean of each
// angle (tsdf_evaluator.cpp:219-221), quaternion as tf2::Quaternion::setRPY
// (roll about x, pitch about y, yaw about z, q = qz(yaw) qy(pitch) qx(roll)).
void pose_from_stats(const double* st, tsl_pose* pose)
{
    std::memset(pose, 0, sizeof(*pose));
    pose->position[0] = (float)st[1];
    pose->position[1] = (float)st[2];
    pose->position[2] = (float)st[3];
    const double roll = (float)std::atan2(st[4], st[5]);
    const double pitch = (float)std::atan2(st[6], st[7]);
    const double yaw = (float)std::atan2(st[8], st[9]);
    const double hr = roll * 0.5, hp = pitch * 0.5, hy = yaw * 0.5;
    const double cr = std::cos(hr), sr = std::sin(hr);
    const double cp = std::cos(hp), sp = std::sin(hp);
    const double cy = std::cos(hy), sy = std::sin(hy);
    pose->orientation[0] = sr * cp * cy - cr * sp * sy;
    pose->orientation[1] = cr * sp * cy + sr * cp * sy;
    pose->orientation[2] = cr * cp * sy - sr * sp * cy;
    pose->orientation[3] = cr * cp * cy + sr * sp * sy;
}

}  // namespace tsl

using namespace tsl;

extern "C" {

const char* tsl_last_error(void) { return g_last_error.c_str(); }

int tsl_version(void) { return 10000; }

int tsl_device_count(int* count)
{
    if (!count) return set_error(TSL_ERR_INVALID_ARGUMENT, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return TSL_OK;
}

int tsl_set_device(int device)
{
    TSL_HIP_CHECK(hipSetDevice(device));
    return TSL_OK;
}

int tsl_map_coef_init(float min_x, float min_y, float min_z, float max_x, float max_y, float max_z,
                      float resolution, float init_value, tsl_map_coef* c)
{
    if (!c) return set_error(TSL_ERR_INVALID_ARGUMENT, "coef is NULL");
    // cuda_sub_voxel_map.tcc:16-34 (float arithmetic, std::ceil, size_t conversion)
    const float sub_voxel_size = 1.0f;
    c->dim_x = cvt_f32_u64(std::ceil(std::fabs(max_x - min_x) / resolution));
    c->dim_y = cvt_f32_u64(std::ceil(std::fabs(max_y - min_y) / resolution));
    c->dim_z = cvt_f32_u64(std::ceil(std::fabs(max_z - min_z) / resolution));
    c->min_x = min_x; c->min_y = min_y; c->min_z = min_z;
    c->max_x = max_x; c->max_y = max_y; c->max_z = max_z;
    c->resolution = resolution;
    c->init_value = init_value;
    c->up_dim_x = cvt_f32_u64(std::ceil(std::fabs(max_x - min_x) / sub_voxel_size));
    c->up_dim_y = cvt_f32_u64(std::ceil(std::fabs(max_y - min_y) / sub_voxel_size));
    c->up_dim_z = cvt_f32_u64(std::ceil(std::fabs(max_z - min_z) / sub_voxel_size));
    c->up_dim_2 = c->up_dim_x * c->up_dim_y;
    c->sub_dim = cvt_f32_u64(std::ceil(sub_voxel_size / resolution));
    c->sub_dim_2 = c->sub_dim * c->sub_dim;
    c->grid_occ_size = c->up_dim_x * c->up_dim_y * c->up_dim_z;
    c->data_size = 0;
    return TSL_OK;
}

static int validate_coef(const tsl_map_coef* c)
{
    if (!(c->resolution > 0.0f) || !std::isfinite(c->resolution))
        return set_error(TSL_ERR_INVALID_ARGUMENT, "resolution must be positive and finite");
    if (c->grid_occ_size != c->up_dim_x * c->up_dim_y * c->up_dim_z || c->grid_occ_size == 0)
        return set_error(TSL_ERR_INVALID_ARGUMENT, "grid_occ_size does not match the block dims");
    if (c->up_dim_2 != c->up_dim_x * c->up_dim_y || c->sub_dim_2 != c->sub_dim * c->sub_dim)
        return set_error(TSL_ERR_INVALID_ARGUMENT, "inconsistent derived dims");
    if (c->data_size > 0x7fffffffull || c->grid_occ_size > 0x7fffffffull || c->sub_dim > 1024)
        return set_error(TSL_ERR_UNSUPPORTED, "map exceeds the int32 block table of the reference (%llu voxels)",
                         (unsigned long long)c->data_size);
    return TSL_OK;
}

static int map_finish(tsl_map* m, const int32_t* grid_occ, const std::vector<float>& palette, uint16_t init_code,
                      bool brick, uint64_t phys_size, bool f32 = false);

// Brick layout when every block offset is an aligned multiple of sd^3.
static bool brick_layout(const tsl_map_coef& c, const int32_t* grid_occ)
{
    const uint64_t sd = c.sub_dim, sd3 = sd * sd * sd;
    bool brick = sd >= 4 && sd % 4 == 0 && sd3 < (1ull << 30);
    for (uint64_t b = 0; brick && b < c.grid_occ_size; ++b)
        if (grid_occ[b] >= 0 && (uint64_t)grid_occ[b] % sd3 != 0) brick = false;
    return brick;
}

// The voxels of the reference's linear order (slot*sd^3 + lx + ly*sd +
// lz*sd^2) in 4x4x4 bricks: element slot*sd^3 + brick*64 + (lx&3) +
// (ly&3)*4 + (lz&3)*16, the last block padded to a whole sd^3.
extern "C++" template <typename T>
static std::vector<T> to_bricks(const T* in, const tsl_map_coef& c)
{
    const uint64_t sd = c.sub_dim, sd3 = sd * sd * sd, B = sd / 4;
    std::vector<T> phys((c.data_size + sd3 - 1) / sd3 * sd3, T{});
    for (uint64_t idx = 0; idx < c.data_size; ++idx) {
        const uint64_t slot = idx / sd3, rem = idx % sd3;
        const uint64_t lz = rem / (sd * sd), ly = (rem / sd) % sd, lx = rem % sd;
        const uint64_t off = ((lx >> 2) + (ly >> 2) * B + (lz >> 2) * B * B) * 64 + (lx & 3) + (ly & 3) * 4 +
                             (lz & 3) * 16;
        phys[slot * sd3 + off] = in[idx];
    }
    return phys;
}

static int map_upload(tsl_map* m, const int32_t* grid_occ, const std::vector<uint16_t>& codes,
                      const std::vector<float>& palette, uint16_t init_code)
{
    const tsl_map_coef& c = m->coef;
    const bool brick = brick_layout(c, grid_occ);
    // device codes are byte offsets into the palette (code * 4 < 2^14), so a
    // lookup addresses the LDS palette without scaling
    std::vector<uint16_t> scaled(codes.size());
    for (size_t i = 0; i < codes.size(); ++i) scaled[i] = (uint16_t)(codes[i] * 4u);
    const std::vector<uint16_t>* upload = &scaled;
    std::vector<uint16_t> phys;
    uint64_t phys_size = c.data_size;
    if (brick && c.data_size) {
        phys = to_bricks(scaled.data(), c);
        phys_size = phys.size();
        upload = &phys;
    }
    // one sentinel element after the voxels holds the code of init_value:
    // rejected lookups read it instead of branching
    TSL_TRY(m->grid_occ.reserve(sizeof(int32_t) * c.grid_occ_size));
    TSL_TRY(m->codes.reserve(sizeof(uint16_t) * (phys_size + 1)));
    TSL_TRY(m->palette.reserve(sizeof(float) * (palette.empty() ? 1 : palette.size())));
    TSL_HIP_CHECK(hipMemcpy(m->grid_occ.ptr, grid_occ, sizeof(int32_t) * c.grid_occ_size,
                            hipMemcpyHostToDevice));
    if (phys_size)
        TSL_HIP_CHECK(hipMemcpy(m->codes.ptr, upload->data(), sizeof(uint16_t) * phys_size,
                                hipMemcpyHostToDevice));
    const uint16_t init_offset = (uint16_t)(init_code * 4u);
    TSL_HIP_CHECK(hipMemcpy(m->codes.as<uint16_t>() + phys_size, &init_offset, sizeof(uint16_t),
                            hipMemcpyHostToDevice));
    if (!palette.empty())
        TSL_HIP_CHECK(hipMemcpy(m->palette.ptr, palette.data(), sizeof(float) * palette.size(),
                                hipMemcpyHostToDevice));
    return map_finish(m, grid_occ, palette, init_code, brick, phys_size);
}

// Float form of a map with more distinct values than the palette holds:
// the values themselves, 4 bytes per voxel, in the element order the codes
// would have (bricked when the block offsets allow), init_value after them.
static int map_upload_f32(tsl_map* m, const int32_t* grid_occ, const float* data)
{
    const tsl_map_coef& c = m->coef;
    const bool brick = brick_layout(c, grid_occ);
    std::vector<float> phys;
    const float* upload = data;
    uint64_t phys_size = c.data_size;
    if (brick && c.data_size) {
        phys = to_bricks(data, c);
        phys_size = phys.size();
        upload = phys.data();
    }
    TSL_TRY(m->grid_occ.reserve(sizeof(int32_t) * c.grid_occ_size));
    TSL_TRY(m->codes.reserve(sizeof(float) * (phys_size + 1)));
    TSL_TRY(m->palette.reserve(sizeof(float)));
    TSL_HIP_CHECK(hipMemcpy(m->grid_occ.ptr, grid_occ, sizeof(int32_t) * c.grid_occ_size,
                            hipMemcpyHostToDevice));
    if (phys_size)
        TSL_HIP_CHECK(hipMemcpy(m->codes.ptr, upload, sizeof(float) * phys_size, hipMemcpyHostToDevice));
    TSL_HIP_CHECK(hipMemcpy(m->codes.as<float>() + phys_size, &c.init_value, sizeof(float),
                            hipMemcpyHostToDevice));
    return map_finish(m, grid_occ, std::vector<float>(), 0, brick, phys_size, true);
}

// The tables and the device views of an uploaded map (codes, palette and
// block table already on the device).  grid_occ: the int32 block table on
// the host, or NULL for a wide map (slot table on the device).  f32: the
// codes array holds the float values themselves.
static int map_finish(tsl_map* m, const int32_t* grid_occ, const std::vector<float>& palette, uint16_t init_code,
                      bool brick, uint64_t phys_size, bool f32)
{
    const tsl_map_coef& c = m->coef;
    const bool wide = grid_occ == nullptr;
    const uint64_t sd = c.sub_dim, sd3 = sd * sd * sd;
    SubTable st;
    TSL_TRY(build_sub_table(c.resolution, &st));
    TSL_TRY(m->sub_tab.reserve(sizeof(SubEntry) * st.entries.size()));
    TSL_HIP_CHECK(hipMemcpy(m->sub_tab.ptr, st.entries.data(), sizeof(SubEntry) * st.entries.size(),
                            hipMemcpyHostToDevice));
    DevMapFull fm{};
    fm.grid_occ = m->grid_occ.as<int32_t>();
    fm.slots = m->slots.as<uint32_t>();
    fm.sd3 = sd3;
    fm.wide = wide ? 1u : 0u;
    fm.res = c.resolution;
    fm.dim_x = c.dim_x; fm.dim_y = c.dim_y; fm.dim_z = c.dim_z;
    fm.up_dim_x = c.up_dim_x;
    fm.up_dim_2 = c.up_dim_2;
    fm.sub_dim = c.sub_dim;
    fm.sub_dim_2 = c.sub_dim_2;
    fm.grid_occ_size = c.grid_occ_size;
    fm.data_size = c.data_size;
    fm.gthr_x = grid_threshold(c.dim_x, c.resolution);
    fm.gthr_y = grid_threshold(c.dim_y, c.resolution);
    fm.gthr_z = grid_threshold(c.dim_z, c.resolution);
    negative_band_constants(c, fm);
    TSL_TRY(m->full.reserve(sizeof(DevMapFull)));
    TSL_HIP_CHECK(hipMemcpy(m->full.ptr, &fm, sizeof(DevMapFull), hipMemcpyHostToDevice));
    DevMap& d = m->dev;
    d.grid_occ = m->grid_occ.as<int32_t>();
    d.slots = m->slots.as<uint32_t>();
    d.wide = wide ? 1u : 0u;
    d.codes = m->codes.as<uint16_t>();
    d.palette = m->palette.as<float>();
    d.sub_tab = m->sub_tab.as<SubEntry>();
    d.full = m->full.as<DevMapFull>();
    d.palette_size = (uint32_t)palette.size();
    d.sub_entries = (uint32_t)st.entries.size();
    d.sub_shift = st.shift;
    d.sub_base = st.base_bits;
    d.min_x = c.min_x; d.min_y = c.min_y; d.min_z = c.min_z;
    d.init = c.init_value;
    d.gthr_x = fm.gthr_x; d.gthr_y = fm.gthr_y; d.gthr_z = fm.gthr_z;
    d.gneg = fm.gneg_x | (fm.gneg_y << 1) | (fm.gneg_z << 2);
    {
        // |offset| < R  =>  |trunc| * (1 + up_dim_x + up_dim_2) < 2^31
        const double lim = 2147483647.0 / (1.0 + (double)c.up_dim_x + (double)c.up_dim_2);
        float R = (float)std::min(8388608.0, std::floor(lim));
        if (R < 1.0f) R = 1.0f;
        std::memcpy(&d.range_bits, &R, 4);
        uint64_t neg[8];
        for (int mask = 0; mask < 8; ++mask)
            neg[mask] = ((mask & 1) ? fm.neg_x : 0ull) + ((mask & 2) ? fm.neg_y : 0ull) + ((mask & 4) ? fm.neg_z : 0ull);
        TSL_TRY(m->neg_tab.reserve(sizeof(neg)));
        TSL_HIP_CHECK(hipMemcpy(m->neg_tab.ptr, neg, sizeof(neg), hipMemcpyHostToDevice));
        d.neg_tab = m->neg_tab.as<uint64_t>();
    }
    d.up_dim_x = (uint32_t)c.up_dim_x;
    d.up_dim_2 = (uint32_t)c.up_dim_2;
    d.occ = (uint32_t)c.grid_occ_size;
    d.ds = (uint32_t)std::min<uint64_t>(c.data_size, 0xffffffffull);
    d.sub_dim = (int32_t)c.sub_dim;
    d.sub_dim_2 = (int32_t)c.sub_dim_2;
    d.data_size = c.data_size;
    d.brick = brick ? 1u : 0u;
    d.sd3 = (uint32_t)sd3;
    d.sd3b = (uint32_t)(2 * sd3);
    d.sentinel = phys_size;
    d.sentinel_code = init_code * 4u;
    d.f32 = f32 ? 1u : 0u;
    {
        // occupancy-rank form of the block table
        // (rank < 2^17 and 2 sub_dim^3 < 2^24 keep the device product, a
        // byte offset, a 24-bit multiply; rank * sub_dim^3 < 2^31 as a
        // block offset)
        // (the word count is padded to a power of two, so any block index's
        // word bits address a word inside the table)
        // The table also covers every block index a lane inside the grid
        // check can form (gthr may exceed a block dim by rounding), so the
        // lookup needs no separate up < grid_occ_size test: such indices
        // read zero padding (unallocated).
        const uint64_t max_up = (uint64_t)(fm.gthr_x ? fm.gthr_x - 1 : 0) +
                                (uint64_t)(fm.gthr_y ? fm.gthr_y - 1 : 0) * c.up_dim_x +
                                (uint64_t)(fm.gthr_z ? fm.gthr_z - 1 : 0) * c.up_dim_2;
        const uint64_t span = std::max<uint64_t>(c.grid_occ_size, max_up + 1);
        uint64_t nwords = 1;
        uint32_t log2w = 0;
        while (nwords * 32 < span) {
            nwords *= 2;
            ++log2w;
        }
        bool rank = !wide && !f32 && nwords <= kMaxOccWords && 2 * sd3 < (1ull << 24);
        std::vector<OccWord> words;
        if (rank) {
            words.assign(nwords, OccWord{0u, 0u});
            uint64_t n = 0;
            for (uint64_t b = 0; rank && b < c.grid_occ_size; ++b) {
                if (b % 32 == 0) words[b / 32].prefix = (uint32_t)n;
                if (grid_occ[b] < 0) continue;
                if ((uint64_t)grid_occ[b] != n * sd3 || n * sd3 >= (1ull << 31)) rank = false;
                words[b / 32].bits |= 1u << (b % 32);
                ++n;
            }
        }
        d.occ_nwords = 0;
        d.occ_used = 0;
        d.occ_words = nullptr;
        if (rank) {
            TSL_TRY(m->occ_words.reserve(sizeof(OccWord) * nwords));
            TSL_HIP_CHECK(hipMemcpy(m->occ_words.ptr, words.data(), sizeof(OccWord) * nwords, hipMemcpyHostToDevice));
            d.occ_words = m->occ_words.as<OccWord>();
            d.occ_nwords = (uint32_t)nwords;
            d.occ_used = (uint32_t)((span + 31) / 32);
            d.occ_log2 = log2w;
        }
    }
    {
        const uint32_t B = (uint32_t)(sd / 4);
        d.brick_cy = 64u * B - 16u;
        d.brick_cz = 64u * B * B - 64u;
        if (brick) {
            magic_divisor((uint32_t)sd, &d.mag_sd, &d.sh_sd);
            magic_divisor((uint32_t)(sd * sd), &d.mag_sd2, &d.sh_sd2);
            magic_divisor((uint32_t)sd3, &d.mag_sd3, &d.sh_sd3);
        }
        // in-block byte-offset contribution of each axis coordinate k in
        // [0, sd]; k = sd (a lane whose index leaves the block row, which
        // recomputes its address) contributes kCarryPoison
        std::vector<uint32_t> ctab(3 * (sd + 1), kCarryPoison);
        for (uint64_t k = 0; k < sd; ++k) {
            if (brick) {
                ctab[k] = (uint32_t)(2 * ((k & 3) + 64 * (k >> 2)));
                ctab[(sd + 1) + k] = (uint32_t)(2 * (4 * (k & 3) + 64 * B * (k >> 2)));
                ctab[2 * (sd + 1) + k] = (uint32_t)(2 * (16 * (k & 3) + 64 * B * B * (k >> 2)));
            } else {
                ctab[k] = (uint32_t)(2 * k);
                ctab[(sd + 1) + k] = (uint32_t)(2 * k * sd);
                ctab[2 * (sd + 1) + k] = (uint32_t)(2 * k * sd * sd);
            }
        }
        const size_t ne = st.entries.size();
        std::vector<AxisEntry> atab(3 * ne);
        for (size_t a = 0; a < 3; ++a)
            for (size_t b = 0; b < ne; ++b) {
                const SubEntry& e = st.entries[b];
                const uint64_t k0 = std::min<uint64_t>(e.count, sd), k1 = std::min<uint64_t>(e.count + 1ull, sd);
                atab[a * ne + b] = AxisEntry{e.thr, ctab[a * (sd + 1) + k0], ctab[a * (sd + 1) + k1], e.count};
            }
        TSL_TRY(m->axis_tab.reserve(sizeof(AxisEntry) * atab.size()));
        TSL_HIP_CHECK(hipMemcpy(m->axis_tab.ptr, atab.data(), sizeof(AxisEntry) * atab.size(), hipMemcpyHostToDevice));
        d.axis_tab = m->axis_tab.as<AxisEntry>();
        d.axis_n = (uint32_t)atab.size();
        // a lane without a poisoned axis has s + pc < 2 (phys_size + sd^3)
        // bytes (s of an unallocated or outside lane included); one with
        // one to three poisoned axes stays below 2^32.  Larger maps send
        // every lane through the recomputation.
        d.carry_limit = 2 * (phys_size + sd3) <= kCarryPoison ? kCarryPoison : 0u;
    }
    return TSL_OK;
}

int tsl_map_create(const tsl_map_coef* coef, const int32_t* grid_occ, const float* data, tsl_map** out)
{
    if (!coef || !grid_occ || !out || (coef->data_size && !data))
        return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    TSL_TRY(validate_coef(coef));
    // value palette: distinct bit patterns in first-seen order; a map with
    // more than kMaxPalette of them is kept in the float form
    std::vector<float> palette;
    std::vector<uint16_t> codes(coef->data_size);
    std::unordered_map<uint32_t, uint16_t> seen;
    bool f32 = false;
    for (uint64_t i = 0; i < coef->data_size && !f32; ++i) {
        uint32_t bits;
        std::memcpy(&bits, &data[i], 4);
        auto it = seen.find(bits);
        if (it == seen.end()) {
            if (palette.size() >= kMaxPalette) {
                f32 = true;
                break;
            }
            it = seen.emplace(bits, (uint16_t)palette.size()).first;
            palette.push_back(data[i]);
        }
        codes[i] = it->second;
    }
    if (f32) {
        std::vector<uint16_t>().swap(codes);
        tsl_map* m = new (std::nothrow) tsl_map();
        if (!m) return set_error(TSL_ERR_OUT_OF_MEMORY, "out of host memory");
        (void)hipGetDevice(&m->device);
        m->coef = *coef;
        const int rc = map_upload_f32(m, grid_occ, data);
        if (rc != TSL_OK) {
            delete m;
            return rc;
        }
        *out = m;
        return TSL_OK;
    }
    uint16_t init_code;
    {
        uint32_t bits;
        std::memcpy(&bits, &coef->init_value, 4);
        auto it = seen.find(bits);
        if (it != seen.end()) {
            init_code = it->second;
        } else {
            init_code = (uint16_t)palette.size();  // at most kMaxPalette + 1 entries
            palette.push_back(coef->init_value);
        }
    }
    tsl_map* m = new (std::nothrow) tsl_map();
    if (!m) return set_error(TSL_ERR_OUT_OF_MEMORY, "out of host memory");
    (void)hipGetDevice(&m->device);
    m->coef = *coef;
    int rc = map_upload(m, grid_occ, codes, palette, init_code);
    if (rc != TSL_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return TSL_OK;
}

float tsl_likelihood_init(float sigma)
{
    // map_util.h:68-71
    const float sigma_quad = sigma * sigma;
    const float e = std::exp((float)(-(10.0 * 10.0) / (double)sigma_quad / 2));
    const float d = std::sqrt((float)((double)(2 * sigma_quad) * M_PI));
    float init = e / d;
    init = init * init * init;
    return init;
}

float tsl_likelihood_value(int16_t tsdf_mm, float sigma)
{
    // map_util.h:108,124-126: value is a double, expf/sqrtf are float
    const float sigma_quad = sigma * sigma;
    const float tsdf_value = (float)tsdf_mm;
    double value = (double)tsdf_value * 0.001;
    const float e = expf((float)(-(value * value) / (double)sigma_quad / 2));
    const float d = sqrtf((float)((double)(2 * sigma_quad) * M_PI));
    value = (double)(e / d);
    value = value * value * value;
    return (float)value;
}

int tsl_map_create_tsdf(const tsl_map_coef* coef, const int32_t* grid_occ, const int16_t* tsdf_mm,
                        int truncation_mm, float sigma, tsl_map** out)
{
    if (!coef || !grid_occ || !out || (coef->data_size && !tsdf_mm))
        return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    TSL_TRY(validate_coef(coef));
    if (truncation_mm < 1 || 2 * (uint32_t)truncation_mm > kMaxPalette)
        return set_error(TSL_ERR_UNSUPPORTED, "truncation_mm must be in [1, %u]", kMaxPalette / 2);
    const int t = truncation_mm;
    std::vector<float> palette(2 * t);
    for (int mm = -(t - 1); mm <= t - 1; ++mm)
        palette[mm + t - 1] = tsl_likelihood_value((int16_t)mm, sigma);
    const uint16_t init_code = (uint16_t)(2 * t - 1);
    palette[init_code] = coef->init_value;
    std::vector<uint16_t> codes(coef->data_size);
    for (uint64_t i = 0; i < coef->data_size; ++i) {
        const int v = tsdf_mm[i];
        codes[i] = (v != INT16_MIN && v > -t && v < t) ? (uint16_t)(v + t - 1) : init_code;
    }
    tsl_map* m = new (std::nothrow) tsl_map();
    if (!m) return set_error(TSL_ERR_OUT_OF_MEMORY, "out of host memory");
    (void)hipGetDevice(&m->device);
    m->coef = *coef;
    int rc = map_upload(m, grid_occ, codes, palette, init_code);
    if (rc != TSL_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return TSL_OK;
}

namespace {

// Codes of a wide TSDF map from its millimetres: workgroup b converts block
// slot s0 + b, whose voxels are mm[(b * sd3) ..] of the staged chunk in the
// reference's linear layout (x fastest), writing each voxel's palette byte
// offset at its bricked position; elements at or past `limit` (the map's
// data size, relative to the chunk) are left alone.
__global__ void k_wide_codes(const int16_t* __restrict__ mm, uint64_t s0, uint64_t limit, int t, uint16_t init_off,
                             uint32_t sd, uint32_t sd3, uint32_t bdim, uint16_t* __restrict__ codes)
{
    const uint64_t slot = s0 + blockIdx.x;
    const uint64_t first = (uint64_t)blockIdx.x * sd3;
    for (uint32_t rem = threadIdx.x; rem < sd3; rem += blockDim.x) {
        if (first + rem >= limit) break;
        const uint32_t lz = rem / (sd * sd), ly = (rem / sd) % sd, lx = rem % sd;
        const uint32_t off = ((lx >> 2) + (ly >> 2) * bdim + (lz >> 2) * bdim * bdim) * 64u + (lx & 3u) +
                             (ly & 3u) * 4u + (lz & 3u) * 16u;
        const int v = mm[first + rem];
        codes[slot * sd3 + off] = (v != INT16_MIN && v > -t && v < t) ? (uint16_t)((v + t - 1) * 4) : init_off;
    }
}

__global__ void k_fill_u16(uint16_t* __restrict__ p, uint64_t n, uint16_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

}  // namespace

int tsl_map_create_tsdf_wide(const tsl_map_coef* coef, const int64_t* grid_occ, const int16_t* tsdf_mm,
                             int truncation_mm, float sigma, tsl_map** out)
{
    if (!coef || !grid_occ || !out || (coef->data_size && !tsdf_mm))
        return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    const tsl_map_coef& c = *coef;
    if (!(c.resolution > 0.0f) || !std::isfinite(c.resolution))
        return set_error(TSL_ERR_INVALID_ARGUMENT, "resolution must be positive and finite");
    if (c.grid_occ_size != c.up_dim_x * c.up_dim_y * c.up_dim_z || c.grid_occ_size == 0 ||
        c.up_dim_2 != c.up_dim_x * c.up_dim_y || c.sub_dim_2 != c.sub_dim * c.sub_dim)
        return set_error(TSL_ERR_INVALID_ARGUMENT, "inconsistent map dims");
    if (c.grid_occ_size > 0x7fffffffull || c.sub_dim < 4 || c.sub_dim > 256 || c.sub_dim % 4 != 0 ||
        c.data_size >= (1ull << 46))
        return set_error(TSL_ERR_UNSUPPORTED,
                         "the 64-bit map form needs sub_dim a multiple of 4 in [4, 256], fewer than 2^31 blocks "
                         "and fewer than 2^46 voxels");
    if (truncation_mm < 1 || 2 * (uint32_t)truncation_mm > kMaxPalette)
        return set_error(TSL_ERR_UNSUPPORTED, "truncation_mm must be in [1, %u]", kMaxPalette / 2);
    const uint64_t sd = c.sub_dim, sd3 = sd * sd * sd;
    // slot table: a block's offset must be a whole multiple of sub_dim^3
    // with the whole block inside the voxel array
    std::vector<uint32_t> slots(c.grid_occ_size);
    for (uint64_t b = 0; b < c.grid_occ_size; ++b) {
        const int64_t off = grid_occ[b];
        if (off < 0) {
            slots[b] = kNoSlot;
            continue;
        }
        if ((uint64_t)off % sd3 != 0 || (uint64_t)off + sd3 > c.data_size || (uint64_t)off / sd3 >= kNoSlot)
            return set_error(TSL_ERR_UNSUPPORTED,
                             "block %llu: offset %lld is not a multiple of sub_dim^3 inside the voxel array",
                             (unsigned long long)b, (long long)off);
        slots[b] = (uint32_t)((uint64_t)off / sd3);
    }
    const int t = truncation_mm;
    std::vector<float> palette(2 * t);
    for (int mm = -(t - 1); mm <= t - 1; ++mm) palette[mm + t - 1] = tsl_likelihood_value((int16_t)mm, sigma);
    const uint16_t init_code = (uint16_t)(2 * t - 1);
    palette[init_code] = c.init_value;
    tsl_map* m = new (std::nothrow) tsl_map();
    if (!m) return set_error(TSL_ERR_OUT_OF_MEMORY, "out of host memory");
    (void)hipGetDevice(&m->device);
    m->coef = c;
    int rc = [&]() -> int {
        const uint64_t nslots = (c.data_size + sd3 - 1) / sd3;
        const uint64_t phys_size = nslots * sd3;
        const uint16_t init_off = (uint16_t)(init_code * 4u);
        TSL_TRY(m->slots.reserve(sizeof(uint32_t) * c.grid_occ_size));
        TSL_HIP_CHECK(hipMemcpy(m->slots.ptr, slots.data(), sizeof(uint32_t) * c.grid_occ_size,
                                hipMemcpyHostToDevice));
        TSL_TRY(m->codes.reserve(sizeof(uint16_t) * (phys_size + 1)));
        // every element, padding and the sentinel after the voxels included,
        // starts as the code of init_value
        hipLaunchKernelGGL(k_fill_u16, dim3(8192), dim3(256), 0, 0, m->codes.as<uint16_t>(), phys_size + 1, init_off);
        TSL_HIP_CHECK(hipGetLastError());
        // millimetres staged a whole number of blocks at a time (about 256 MB)
        const uint64_t per = std::max<uint64_t>(1, (128ull << 20) / sd3);
        DevBuf stage;
        TSL_TRY(stage.reserve(sizeof(int16_t) * per * sd3));
        for (uint64_t s0 = 0; s0 < nslots; s0 += per) {
            const uint64_t s1 = std::min(nslots, s0 + per);
            const uint64_t e0 = s0 * sd3, e1 = std::min(s1 * sd3, c.data_size);
            TSL_HIP_CHECK(hipMemcpy(stage.ptr, tsdf_mm + e0, sizeof(int16_t) * (e1 - e0), hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_wide_codes, dim3((unsigned)(s1 - s0)), dim3(256), 0, 0, stage.as<int16_t>(), s0,
                               e1 - e0, t, init_off, (uint32_t)sd, (uint32_t)sd3, (uint32_t)(sd / 4),
                               m->codes.as<uint16_t>());
            TSL_HIP_CHECK(hipGetLastError());
        }
        TSL_HIP_CHECK(hipDeviceSynchronize());
        TSL_TRY(m->palette.reserve(sizeof(float) * palette.size()));
        TSL_HIP_CHECK(hipMemcpy(m->palette.ptr, palette.data(), sizeof(float) * palette.size(),
                                hipMemcpyHostToDevice));
        return map_finish(m, nullptr, palette, init_code, true, phys_size);
    }();
    if (rc != TSL_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return TSL_OK;
}

// CudaSubVoxelMap::getIndex on the host, used by the setData replay.
static uint64_t host_get_index(const tsl_map_coef& c, const int32_t* grid_occ, float x, float y, float z)
{
    uint64_t gx = cvt_f32_u64(x - c.min_x), gy = cvt_f32_u64(y - c.min_y), gz = cvt_f32_u64(z - c.min_z);
    gx = cvt_f32_u64(cvt_u64_f32(gx) / c.resolution);
    gy = cvt_f32_u64(cvt_u64_f32(gy) / c.resolution);
    gz = cvt_f32_u64(cvt_u64_f32(gz) / c.resolution);
    if (gx >= c.dim_x || gy >= c.dim_y || gz >= c.dim_z) return c.data_size;
    const float ox = x - c.min_x, oy = y - c.min_y, oz = z - c.min_z;
    const uint64_t ux = cvt_f32_u64(ox / 1.0f), uy = cvt_f32_u64(oy / 1.0f), uz = cvt_f32_u64(oz / 1.0f);
    const float spx = ox - cvt_u64_f32(ux) * 1.0f, spy = oy - cvt_u64_f32(uy) * 1.0f,
                spz = oz - cvt_u64_f32(uz) * 1.0f;
    const uint64_t up = ux + uy * c.up_dim_x + uz * c.up_dim_2;
    if (up >= c.grid_occ_size) return c.data_size;
    const int32_t s = grid_occ[up];
    if (s < 0) return c.data_size;
    return (uint64_t)(int64_t)s + cvt_f32_u64(spx / c.resolution) +
           cvt_f32_u64(spy / c.resolution) * c.sub_dim + cvt_f32_u64(spz / c.resolution) * c.sub_dim_2;
}

int tsl_map_set_data(tsl_map_coef* c, const float* cell_xyz, const float* cell_value, size_t num_cells,
                     int32_t* grid_occ, float* data, size_t data_capacity)
{
    if (!c || !grid_occ || (num_cells && (!cell_xyz || !cell_value)))
        return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    // cuda_sub_voxel_map.tcc:171-231
    for (uint64_t i = 0; i < c->grid_occ_size; ++i) grid_occ[i] = -1;
    if (num_cells == 0) return TSL_OK;  // :173-176 leaves data untouched
    for (size_t i = 0; i < num_cells; ++i) {
        const float ox = cell_xyz[3 * i] - c->min_x, oy = cell_xyz[3 * i + 1] - c->min_y,
                    oz = cell_xyz[3 * i + 2] - c->min_z;
        const uint64_t ux = cvt_f32_u64(ox / 1.0f), uy = cvt_f32_u64(oy / 1.0f), uz = cvt_f32_u64(oz / 1.0f);
        const uint64_t up = ux + uy * c->up_dim_x + uz * c->up_dim_x * c->up_dim_y;
        if (up >= c->grid_occ_size) return set_error(TSL_ERR_INVALID_ARGUMENT, "Upper voxel index overflow!");
        grid_occ[up] = 0;
    }
    // the reference's offsets are int (OCC_T, cuda_sub_voxel_map.h:13) and so
    // is its running offset (.tcc:202-211): a map whose voxels do not fit
    // in int32 has no representation there and is refused before any
    // offset is written
    const uint64_t sub_size = c->sub_dim * c->sub_dim * c->sub_dim;
    uint64_t blocks = 0;
    for (uint64_t i = 0; i < c->grid_occ_size; ++i) blocks += grid_occ[i] >= 0;
    if (sub_size != 0 && blocks > (uint64_t)INT32_MAX / sub_size)
        return set_error(TSL_ERR_UNSUPPORTED,
                         "%llu blocks of %llu voxels exceed the int32 block offsets of the reference's map",
                         (unsigned long long)blocks, (unsigned long long)sub_size);
    int64_t offset = 0;
    for (uint64_t i = 0; i < c->grid_occ_size; ++i) {
        if (grid_occ[i] >= 0) {
            grid_occ[i] = (int32_t)offset;
            offset += (int64_t)sub_size;
        }
    }
    c->data_size = (uint64_t)offset;
    if (!data || data_capacity < c->data_size)
        return set_error(TSL_ERR_CAPACITY, "data needs %llu floats", (unsigned long long)c->data_size);
    for (uint64_t i = 0; i < c->data_size; ++i) data[i] = c->init_value;
    for (size_t i = 0; i < num_cells; ++i) {
        const uint64_t idx = host_get_index(*c, grid_occ, cell_xyz[3 * i], cell_xyz[3 * i + 1], cell_xyz[3 * i + 2]);
        if (idx < c->data_size) data[idx] = cell_value[i];
    }
    return TSL_OK;
}

int tsl_map_device_bytes(const tsl_map* m, uint64_t* bytes)
{
    if (!m || !bytes) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    *bytes = m->grid_occ.bytes + m->codes.bytes + m->palette.bytes + m->sub_tab.bytes + m->full.bytes +
             m->neg_tab.bytes + m->occ_words.bytes + m->axis_tab.bytes + m->slots.bytes;
    return TSL_OK;
}

int tsl_map_destroy(tsl_map* m)
{
    if (m) {
        (void)hipSetDevice(m->device);
        delete m;
    }
    return TSL_OK;
}

int tsl_filter_create(tsl_evaluator* ev, size_t capacity, tsl_filter** out)
{
    if (!ev || !out) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    TSL_HIP_CHECK(hipSetDevice(ev->device));
    tsl_filter* f = new (std::nothrow) tsl_filter();
    if (!f) return set_error(TSL_ERR_OUT_OF_MEMORY, "out of host memory");
    f->ev = ev;
    f->device = ev->device;
    int rc = f->state.reserve(sizeof(FilterState));
    if (rc == TSL_OK && hipHostMalloc(&f->host_state, sizeof(FilterState), hipHostMallocDefault) != hipSuccess)
        rc = set_error(TSL_ERR_OUT_OF_MEMORY, "hipHostMalloc failed");
    if (rc == TSL_OK) {
        FilterState init{};
        init.first_bad = 0xffffffffu;
        if (hipMemcpy(f->state.ptr, &init, sizeof(FilterState), hipMemcpyHostToDevice) != hipSuccess)
            rc = set_error(TSL_ERR_HIP, "hipMemcpy failed");
    }
    if (rc == TSL_OK) rc = filter_reserve_particles(f, capacity ? capacity : 1);
    if (rc == TSL_OK) rc = filter_reserve_points(f, 1);
    if (rc != TSL_OK) {
        tsl_filter_destroy(f);
        return rc;
    }
    std::memset(f->host_state, 0, sizeof(FilterState));
    *out = f;
    return TSL_OK;
}

int tsl_filter_destroy(tsl_filter* f)
{
    if (!f) return TSL_OK;
    (void)hipSetDevice(f->device);
    (void)hipDeviceSynchronize();
    if (f->host_state) (void)hipHostFree(f->host_state);
    if (f->weights_host) (void)hipHostFree(f->weights_host);
    if (f->motion_host) (void)hipHostFree(f->motion_host);
    if (f->motion_copied) (void)hipEventDestroy(static_cast<hipEvent_t>(f->motion_copied));
    if (f->ev_start) (void)hipEventDestroy(static_cast<hipEvent_t>(f->ev_start));
    if (f->ev_stop) (void)hipEventDestroy(static_cast<hipEvent_t>(f->ev_stop));
    delete f;
    return TSL_OK;
}

int tsl_filter_set_particles(tsl_filter* f, const tsl_particle* p, size_t n)
{
    if (!f || (n && !p)) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n > 0x7fffffffull) return set_error(TSL_ERR_UNSUPPORTED, "too many particles");
    TSL_HIP_CHECK(hipSetDevice(f->device));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    f->count_bound = 0;
    TSL_TRY(filter_reserve_particles(f, n ? n : 1));
    if (n) TSL_HIP_CHECK(hipMemcpy(f->particles.ptr, p, sizeof(tsl_particle) * n, hipMemcpyHostToDevice));
    f->count_bound = n;
    f->count_known = true;
    TSL_TRY(launch_set_count(f, (uint32_t)n, 0));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    host_state(f)->count = (uint32_t)n;
    return TSL_OK;
}

int tsl_filter_get_particles(tsl_filter* f, tsl_particle* p, size_t capacity, size_t* n)
{
    if (!f || !n) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    TSL_HIP_CHECK(hipSetDevice(f->device));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    TSL_HIP_CHECK(hipMemcpy(f->host_state, f->state.ptr, sizeof(FilterState), hipMemcpyDeviceToHost));
    const size_t cnt = host_state(f)->count;
    *n = cnt;
    if (cnt > capacity) return set_error(TSL_ERR_CAPACITY, "filter holds %zu particles", cnt);
    if (cnt) TSL_HIP_CHECK(hipMemcpy(p, f->particles.ptr, sizeof(tsl_particle) * cnt, hipMemcpyDeviceToHost));
    return TSL_OK;
}

int tsl_filter_set_scan(tsl_filter* f, const tsl_point* pts, size_t n, const float tf[16])
{
    if (!f || !tf || (n && !pts)) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n > 0x7fffffffull) return set_error(TSL_ERR_UNSUPPORTED, "too many points");
    TSL_HIP_CHECK(hipSetDevice(f->device));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    TSL_TRY(filter_reserve_points(f, n));
    if (n) TSL_HIP_CHECK(hipMemcpy(f->points_in.ptr, pts, sizeof(tsl_point) * n, hipMemcpyHostToDevice));
    f->num_points = n;
    std::memcpy(f->tf, tf, sizeof(float) * 16);
    TSL_TRY(launch_prep_points(f, 0));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    return TSL_OK;
}

int tsl_filter_update_async(tsl_filter* f, void* stream)
{
    if (!f) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL filter");
    if (f->num_points == 0) return TSL_OK;  // empty scan: nothing to do (cuda_evaluator.cu:122-125)
    TSL_HIP_CHECK(hipSetDevice(f->device));
    return launch_update(f, static_cast<hipStream_t>(stream));
}

int tsl_filter_resample_async(tsl_filter* f, float u0, void* stream)
{
    if (!f) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL filter");
    TSL_HIP_CHECK(hipSetDevice(f->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the device plans the draw from its own count; the host only sizes the
    // buffers for the largest output the current count bound allows
    const size_t M = f->count_bound;
    if (M == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "filter holds no particles");
    const uint64_t need = resample_bound(M);
    if (need == 0 || need > 0x7fffffffull)
        return set_error(TSL_ERR_UNSUPPORTED, "systematic draw does not advance for %zu particles", M);
    if (need > f->capacity || f->ubuf.bytes < sizeof(float) * need) {
        TSL_HIP_CHECK(hipStreamSynchronize(s));  // buffers are reallocated below
        if (need > f->capacity) TSL_TRY(filter_reserve_particles(f, need));
        TSL_TRY(f->ubuf.reserve(sizeof(float) * need));
    }
    Member* dm;
    TSL_TRY(filter_member(f, &dm));
    TSL_TRY(launch_resample_members(dm, 1, (uint32_t)need, u0, true, s));
    f->count_bound = need;
    f->count_known = false;
    return TSL_OK;
}

int tsl_filter_snapshot(tsl_filter* f)
{
    if (!f) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL filter");
    TSL_HIP_CHECK(hipSetDevice(f->device));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    TSL_HIP_CHECK(hipMemcpy(f->host_state, f->state.ptr, sizeof(FilterState), hipMemcpyDeviceToHost));
    const size_t n = host_state(f)->count;
    TSL_TRY(f->snapshot.reserve(sizeof(tsl_particle) * (n ? n : 1)));
    if (n) TSL_HIP_CHECK(hipMemcpy(f->snapshot.ptr, f->particles.ptr, sizeof(tsl_particle) * n,
                                   hipMemcpyDeviceToDevice));
    f->snapshot_count = n;
    f->count_bound = n;
    f->count_known = true;
    return TSL_OK;
}

int tsl_filter_restore_async(tsl_filter* f, void* stream)
{
    if (!f) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL filter");
    if (!f->snapshot.ptr) return set_error(TSL_ERR_INVALID_ARGUMENT, "filter has no snapshot");
    TSL_HIP_CHECK(hipSetDevice(f->device));
    if (f->snapshot_count > f->capacity) TSL_TRY(filter_reserve_particles(f, f->snapshot_count));
    TSL_TRY(launch_restore(f, static_cast<hipStream_t>(stream)));
    f->count_bound = f->snapshot_count;
    f->count_known = true;
    return TSL_OK;
}

int tsl_filter_set_tile(tsl_filter* f, int k, int points_per_lane, uint32_t chunk_len)
{
    if (!f) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL filter");
    if (k != 0 && k != 1 && k != 2) return set_error(TSL_ERR_INVALID_ARGUMENT, "K must be 0, 1 or 2");
    if (points_per_lane != 0 && points_per_lane != 1 && points_per_lane != 2 && points_per_lane != 4)
        return set_error(TSL_ERR_INVALID_ARGUMENT, "points per lane must be 0, 1, 2 or 4");
    f->tile_k = k;
    f->tile_p = points_per_lane;
    f->tile_chunk = chunk_len;
    return TSL_OK;
}

int tsl_filter_set_block_lookup(tsl_filter* f, int mode)
{
    if (!f) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL filter");
    if (mode < -1 || mode > 1) return set_error(TSL_ERR_INVALID_ARGUMENT, "block lookup mode must be -1, 0 or 1");
    f->rank_mode = mode;
    return TSL_OK;
}

int tsl_filter_set_order(tsl_filter* f, int mode)
{
    if (!f) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL filter");
    if (mode < -1 || mode > 1) return set_error(TSL_ERR_INVALID_ARGUMENT, "order mode must be -1, 0 or 1");
    f->order_mode = mode;
    return TSL_OK;
}

int tsl_filter_autotune(tsl_filter* f, void* stream, int reps, int* best_order, int* best_k,
                        int* best_points_per_lane, uint32_t* best_chunk, float* best_ms)
{
    if (!f) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL filter");
    TSL_HIP_CHECK(hipSetDevice(f->device));
    if (!f->count_known) {
        FilterState* hs = host_state(f);
        TSL_HIP_CHECK(hipMemcpy(hs, f->state.ptr, sizeof(FilterState), hipMemcpyDeviceToHost));
        f->count_bound = hs->count;
        f->count_known = true;
    }
    return autotune_tile(f, static_cast<hipStream_t>(stream), reps > 0 ? reps : 3, best_order, best_k,
                         best_points_per_lane, best_chunk, best_ms);
}

int tsl_filter_set_timing(tsl_filter* f, int enable)
{
    if (!f) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL filter");
    TSL_HIP_CHECK(hipSetDevice(f->device));
    if (enable && !f->ev_start) {
        hipEvent_t a, b;
        TSL_HIP_CHECK(hipEventCreate(&a));
        TSL_HIP_CHECK(hipEventCreate(&b));
        f->ev_start = a;
        f->ev_stop = b;
    }
    f->timing = enable != 0;
    f->timing_pending = false;
    f->eval_ms = 0.0;
    f->eval_launches = 0;
    f->eval_pairs = 0;
    return TSL_OK;
}

int tsl_filter_eval_timing(tsl_filter* f, double* ms, uint64_t* launches, uint64_t* pairs)
{
    if (!f || !ms || !launches || !pairs) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    TSL_HIP_CHECK(hipSetDevice(f->device));
    if (f->timing_pending) {
        float t = 0.0f;
        TSL_HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(f->ev_stop)));
        TSL_HIP_CHECK(hipEventElapsedTime(&t, static_cast<hipEvent_t>(f->ev_start),
                                          static_cast<hipEvent_t>(f->ev_stop)));
        f->eval_ms += t;
        f->timing_pending = false;
    }
    *ms = f->eval_ms;
    *launches = f->eval_launches;
    *pairs = f->eval_pairs;
    return TSL_OK;
}

int tsl_filter_sync(tsl_filter* f, void* stream, tsl_pose* pose, size_t* count)
{
    if (!f) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL filter");
    TSL_HIP_CHECK(hipSetDevice(f->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    FilterState* hs = host_state(f);
    TSL_HIP_CHECK(hipMemcpyAsync(hs, f->state.ptr, sizeof(FilterState), hipMemcpyDeviceToHost, s));
    TSL_HIP_CHECK(hipStreamSynchronize(s));
    if (count) *count = hs->count;
    f->count_bound = hs->count;
    f->count_known = true;
    if (pose) pose_from_stats(hs->stats, pose);
    return status_error(hs->status);
}


int tsl_evaluator_create(tsl_map* map, int per_point, float a_hit, float a_range, float a_max,
                         float max_range, tsl_evaluator** out)
{
    if (!map || !out) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    tsl_evaluator* ev = new (std::nothrow) tsl_evaluator();
    if (!ev) return set_error(TSL_ERR_OUT_OF_MEMORY, "out of host memory");
    ev->map = map;
    ev->device = map->device;
    ev->per_point = per_point;
    ev->a_hit = a_hit;
    ev->a_range = a_range;
    ev->a_max = a_max;
    ev->max_range = max_range;
    // tsdf_evaluator.h:75: inv_max_range_(1.0 / max_range), max_range_squared_(max_range * max_range)
    ev->inv_max_range = (float)(1.0 / (double)max_range);
    ev->max_range_squared = max_range * max_range;
    ev->range_term = a_range * ev->inv_max_range;
    *out = ev;
    return TSL_OK;
}

int tsl_evaluator_set_trig(tsl_evaluator* ev, int mode)
{
    if (!ev) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL evaluator");
    if (mode < TSL_TRIG_GLIBC_FMA || mode > TSL_TRIG_DOUBLE) return set_error(TSL_ERR_INVALID_ARGUMENT, "bad trig mode");
    ev->trig = mode;
    return TSL_OK;
}

int tsl_evaluator_get_trig(const tsl_evaluator* ev, int* mode)
{
    if (!ev || !mode) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    *mode = ev->trig;
    return TSL_OK;
}

int tsl_trig(int mode, const float* x, size_t n, float* s, float* c, int on_device)
{
    if (n && (!x || !s || !c)) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (mode < TSL_TRIG_GLIBC_FMA || mode > TSL_TRIG_DOUBLE) return set_error(TSL_ERR_INVALID_ARGUMENT, "bad trig mode");
    if (n > 0x7fffffffull) return set_error(TSL_ERR_UNSUPPORTED, "too many angles");
    if (n == 0) return TSL_OK;
    if (!on_device) {
        for (size_t i = 0; i < n; ++i) sincosf_mode(mode, x[i], &s[i], &c[i]);
        return TSL_OK;
    }
    DevBuf d;
    TSL_TRY(d.reserve(sizeof(float) * 3 * n));
    float* dx = d.as<float>();
    TSL_HIP_CHECK(hipMemcpy(dx, x, sizeof(float) * n, hipMemcpyHostToDevice));
    TSL_TRY(launch_trig(dx, (uint32_t)n, mode, dx + n, dx + 2 * n, 0));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    TSL_HIP_CHECK(hipMemcpy(s, dx + n, sizeof(float) * n, hipMemcpyDeviceToHost));
    TSL_HIP_CHECK(hipMemcpy(c, dx + 2 * n, sizeof(float) * n, hipMemcpyDeviceToHost));
    return TSL_OK;
}

int tsl_check_trig_libm(int mode, uint32_t stride, uint64_t* mismatches, uint64_t* checked)
{
    if (!mismatches || !checked || stride == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "bad argument");
    if (mode < TSL_TRIG_GLIBC_FMA || mode > TSL_TRIG_DOUBLE) return set_error(TSL_ERR_INVALID_ARGUMENT, "bad trig mode");
    const unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    std::vector<uint64_t> bad(nt, 0), seen(nt, 0);
    std::vector<std::thread> th;
    const uint64_t span = 0x100000000ull;
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t]() {
            const uint64_t per = (span / stride + nt - 1) / nt;
            const uint64_t k0 = per * t, k1 = std::min(per * (t + 1), (span + stride - 1) / stride);
            uint64_t b = 0, n = 0;
            for (uint64_t k = k0; k < k1; ++k) {
                const uint32_t u = (uint32_t)(k * stride);
                float x;
                std::memcpy(&x, &u, 4);
                if (x != x) continue;
                float s, c;
                sincosf_mode(mode, x, &s, &c);
                volatile float xv = x;
                const float ls = ::sinf(xv), lc = ::cosf(xv);
                if (std::memcmp(&s, &ls, 4) != 0 || std::memcmp(&c, &lc, 4) != 0) ++b;
                ++n;
            }
            bad[t] = b;
            seen[t] = n;
        });
    }
    for (auto& x : th) x.join();
    uint64_t b = 0, n = 0;
    for (unsigned t = 0; t < nt; ++t) {
        b += bad[t];
        n += seen[t];
    }
    *mismatches = b;
    *checked = n;
    return TSL_OK;
}

int tsl_check_invtrig_libm(uint32_t stride, uint64_t atan2_samples, uint64_t* mismatches, uint64_t* checked)
{
    if (!mismatches || !checked || stride == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "bad argument");
    const unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    std::vector<uint64_t> bad(nt, 0), seen(nt, 0);
    std::vector<std::thread> th;
    const uint64_t span = 0x100000000ull, steps = (span + stride - 1) / stride;
    auto same = [](float a, float b) { return std::memcmp(&a, &b, 4) == 0 || (a != a && b != b); };
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t]() {
            const uint64_t per = (steps + nt - 1) / nt;
            const uint64_t k0 = per * t, k1 = std::min(per * (t + 1), steps);
            uint64_t b = 0, n = 0;
            for (uint64_t k = k0; k < k1; ++k) {
                const uint32_t u = (uint32_t)(k * stride);
                float x;
                std::memcpy(&x, &u, 4);
                if (x != x) continue;
                volatile float xv = x;
                if (!same(asinf_glibc(x), ::asinf(xv))) ++b;
                if (!same(atanf_glibc(x), ::atanf(xv))) ++b;
                n += 2;
            }
            // atan2f: random bit patterns, every other pair with exponents
            // drawn from [2^-27, 2^28]
            const uint64_t per2 = (atan2_samples + nt - 1) / nt;
            const uint64_t j0 = per2 * t, j1 = std::min(per2 * (t + 1), atan2_samples);
            for (uint64_t j = j0; j < j1; ++j) {
                const uint64_t h = splitmix64(0x5EEDA7A2ull ^ splitmix64(j));
                uint32_t a = (uint32_t)h, c = (uint32_t)(h >> 32);
                if (j & 1) {
                    a = (a & 0x807fffffu) | ((100u + (a >> 23) % 56u) << 23);
                    c = (c & 0x807fffffu) | ((100u + (c >> 23) % 56u) << 23);
                }
                float y, x;
                std::memcpy(&y, &a, 4);
                std::memcpy(&x, &c, 4);
                volatile float yv = y, xv = x;
                if (!same(atan2f_glibc(y, x), ::atan2f(yv, xv))) ++b;
                ++n;
            }
            bad[t] = b;
            seen[t] = n;
        });
    }
    for (auto& x : th) x.join();
    uint64_t b = 0, n = 0;
    for (unsigned t = 0; t < nt; ++t) {
        b += bad[t];
        n += seen[t];
    }
    *mismatches = b;
    *checked = n;
    return TSL_OK;
}

int tsl_evaluator_destroy(tsl_evaluator* ev)
{
    if (!ev) return TSL_OK;
    if (ev->scratch) tsl_filter_destroy(ev->scratch);
    delete ev;
    return TSL_OK;
}

static int scratch_filter(tsl_evaluator* ev, size_t P, tsl_filter** f)
{
    if (!ev->scratch) TSL_TRY(tsl_filter_create(ev, P ? P : 1, &ev->scratch));
    *f = ev->scratch;
    return TSL_OK;
}

// The evaluation of a host particle set on the evaluator's scratch filter,
// against a host scan (points != nullptr) or the scan already on it:
// weights overwritten with the normalised weights (cuda_evaluator.cu:
// 405-408), pose returned.  The copies in, the scan preparation, the update
// and the read-back of the state and of the normalised weights (kept
// contiguous by the normalisation) are queued back to back on the null
// stream with one synchronisation at the end; the calls before it leave no
// work in flight on this filter, and a buffer that grows is freed by
// hipFree, which waits for the device.
static int evaluate_filter_scan(tsl_evaluator* ev, tsl_filter* f, tsl_particle* particles, size_t P,
                                tsl_pose* out_pose, const tsl_point* points = nullptr, size_t N = 0,
                                const float* tf = nullptr)
{
    (void)ev;
    if (P == 0) return set_error(TSL_ERR_NO_VALID_PARTICLE, "No particle is valid!");
    if (P > 0x7fffffffull) return set_error(TSL_ERR_UNSUPPORTED, "too many particles");
    if (N > 0x7fffffffull) return set_error(TSL_ERR_UNSUPPORTED, "too many points");
    TSL_HIP_CHECK(hipSetDevice(f->device));
    hipStream_t s = nullptr;
    f->member_on_stream = true;  // the descriptor is written in stream order, by a kernel
    if (f->weights_host_cap < P) {
        if (f->weights_host) (void)hipHostFree(f->weights_host);
        f->weights_host = nullptr;
        f->weights_host_cap = 0;
        if (hipHostMalloc(&f->weights_host, sizeof(float) * P, hipHostMallocDefault) != hipSuccess)
            return set_error(TSL_ERR_OUT_OF_MEMORY, "hipHostMalloc failed");
        f->weights_host_cap = P;
    }
    if (points) {
        TSL_TRY(filter_reserve_points(f, N));
        TSL_HIP_CHECK(hipMemcpyAsync(f->points_in.ptr, points, sizeof(tsl_point) * N, hipMemcpyHostToDevice, s));
        f->num_points = N;
        std::memcpy(f->tf, tf, sizeof(float) * 16);
        TSL_TRY(launch_prep_points(f, s));
    }
    f->count_bound = 0;
    TSL_TRY(filter_reserve_particles(f, P));
    TSL_HIP_CHECK(hipMemcpyAsync(f->particles.ptr, particles, sizeof(tsl_particle) * P, hipMemcpyHostToDevice, s));
    f->count_bound = P;
    f->count_known = true;
    TSL_TRY(launch_set_count(f, (uint32_t)P, s));
    TSL_TRY(launch_update(f, s));
    FilterState* hs = host_state(f);
    TSL_HIP_CHECK(hipMemcpyAsync(hs, f->state.ptr, sizeof(FilterState), hipMemcpyDeviceToHost, s));
    TSL_HIP_CHECK(hipMemcpyAsync(f->weights_host, f->raw.ptr, sizeof(float) * P, hipMemcpyDeviceToHost, s));
    TSL_HIP_CHECK(hipStreamSynchronize(s));
    f->count_bound = hs->count;
    pose_from_stats(hs->stats, out_pose);
    TSL_TRY(status_error(hs->status));
    const float* w = static_cast<const float*>(f->weights_host);
    for (size_t i = 0; i < P; ++i) particles[i].second = w[i];
    return TSL_OK;
}

int tsl_evaluate(tsl_evaluator* ev, tsl_particle* particles, size_t P, const tsl_point* points, size_t N,
                 const float tf[16], tsl_pose* out_pose)
{
    if (!ev || !out_pose || !tf || (P && !particles) || (N && !points))
        return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    clear_error();
    if (N == 0) {  // cuda_evaluator.cu:122-125
        std::memset(out_pose, 0, sizeof(*out_pose));
        return TSL_OK;
    }
    tsl_filter* f;
    TSL_TRY(scratch_filter(ev, P, &f));
    return evaluate_filter_scan(ev, f, particles, P, out_pose, points, N, tf);
}

int tsl_evaluate_raw(tsl_evaluator* ev, const tsl_particle* particles, size_t P, const tsl_point* points,
                     size_t N, const float tf[16], float* weights_out)
{
    if (!ev || !tf || (P && (!particles || !weights_out)) || (N && !points))
        return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (P == 0) return TSL_OK;
    if (N == 0) {
        std::memset(weights_out, 0, sizeof(float) * P);
        return TSL_OK;
    }
    tsl_filter* f;
    TSL_TRY(scratch_filter(ev, P, &f));
    TSL_TRY(tsl_filter_set_particles(f, particles, P));
    TSL_TRY(tsl_filter_set_scan(f, points, N, tf));
    TSL_TRY(launch_raw(f, 0));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    TSL_HIP_CHECK(hipMemcpy(weights_out, f->raw.ptr, sizeof(float) * P, hipMemcpyDeviceToHost));
    return TSL_OK;
}

static int trace_common(tsl_evaluator* ev, const tsl_particle* particles, size_t P, const tsl_point* points,
                        size_t N, const float tf[16], void* out, size_t elem, int values)
{
    if (!ev || !tf || (P && !particles) || (N && !points) || (P * N && !out))
        return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (P == 0 || N == 0) return TSL_OK;
    tsl_filter* f;
    TSL_TRY(scratch_filter(ev, P, &f));
    TSL_TRY(tsl_filter_set_particles(f, particles, P));
    TSL_TRY(tsl_filter_set_scan(f, points, N, tf));
    DevBuf d_out;
    TSL_TRY(d_out.reserve(elem * P * N));
    TSL_TRY(launch_trace(f, 0, d_out.ptr, values));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    TSL_HIP_CHECK(hipMemcpy(out, d_out.ptr, elem * P * N, hipMemcpyDeviceToHost));
    return TSL_OK;
}

int tsl_trace_indices(tsl_evaluator* ev, const tsl_particle* particles, size_t P, const tsl_point* points,
                      size_t N, const float tf[16], uint64_t* out)
{
    return trace_common(ev, particles, P, points, N, tf, out, sizeof(uint64_t), 0);
}

int tsl_trace_values(tsl_evaluator* ev, const tsl_particle* particles, size_t P, const tsl_point* points,
                     size_t N, const float tf[16], float* out)
{
    return trace_common(ev, particles, P, points, N, tf, out, sizeof(float), 1);
}

int tsl_pose_matrices(const tsl_particle* particles, size_t P, const float tf[16], int trig, float* out)
{
    if (!tf || (P && (!particles || !out))) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (trig < TSL_TRIG_GLIBC_FMA || trig > TSL_TRIG_DOUBLE) return set_error(TSL_ERR_INVALID_ARGUMENT, "bad trig mode");
    if (P == 0) return TSL_OK;
    DevBuf d_p, d_l;
    TSL_TRY(d_p.reserve(sizeof(tsl_particle) * P));
    TSL_TRY(d_l.reserve(sizeof(float) * 12 * P));
    TSL_HIP_CHECK(hipMemcpy(d_p.ptr, particles, sizeof(tsl_particle) * P, hipMemcpyHostToDevice));
    TSL_TRY(launch_pose_matrices(d_p.as<tsl_particle>(), (uint32_t)P, tf, trig, d_l.as<float>(), 0));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    TSL_HIP_CHECK(hipMemcpy(out, d_l.ptr, sizeof(float) * 12 * P, hipMemcpyDeviceToHost));
    return TSL_OK;
}

int tsl_resample_capacity(size_t M, float u0, size_t* capacity)
{
    if (!capacity) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    return resample_capacity(M, u0, capacity);
}

static int set_scan_reduced(tsl_filter* f, int mode, const tsl_point* raw, const void* ring, size_t ring_bytes,
                            size_t n, float cell, const float tf[16])
{
    if (!f || !tf || (n && (!raw || !ring))) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n > 0x7fffffffull) return set_error(TSL_ERR_UNSUPPORTED, "too many points");
    if (!(cell > 0.0f)) return set_error(TSL_ERR_INVALID_ARGUMENT, "reduction cell size must be positive");
    TSL_HIP_CHECK(hipSetDevice(f->device));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    TSL_TRY(filter_reserve_points(f, n));
    TSL_TRY(f->raw_in.reserve(sizeof(tsl_point) * (n ? n : 1)));
    const size_t ring_room = (ring_bytes * (n ? n : 1) + 3) & ~(size_t)3;
    TSL_TRY(f->raw_ring.reserve(ring_room + 16));  // rings, then two aligned counters
    uint32_t* counters = reinterpret_cast<uint32_t*>(f->raw_ring.as<char>() + ring_room);
    const uint32_t zero[2] = {0u, 0u};
    TSL_HIP_CHECK(hipMemcpy(counters, zero, sizeof(zero), hipMemcpyHostToDevice));
    uint32_t host[2] = {0u, 0u};
    if (n) {
        TSL_HIP_CHECK(hipMemcpy(f->raw_in.ptr, raw, sizeof(tsl_point) * n, hipMemcpyHostToDevice));
        TSL_HIP_CHECK(hipMemcpy(f->raw_ring.ptr, ring, ring_bytes * n, hipMemcpyHostToDevice));
        TSL_TRY(launch_reduce_scan(f, mode, f->raw_in.as<tsl_point>(), f->raw_ring.ptr, (uint32_t)n, cell,
                                   f->points_in.as<tsl_point>(), counters, counters + 1, 0));
        TSL_HIP_CHECK(hipDeviceSynchronize());
        TSL_HIP_CHECK(hipMemcpy(host, counters, sizeof(host), hipMemcpyDeviceToHost));
    }
    if (host[1]) return set_error(TSL_ERR_INVALID_ARGUMENT, "a kept point's ring is outside [0, 64)");
    f->num_points = host[0];
    std::memcpy(f->tf, tf, sizeof(float) * 16);
    TSL_TRY(launch_prep_points(f, 0));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    return TSL_OK;
}

int tsl_filter_set_scan_raw(tsl_filter* f, const tsl_point* raw, const int16_t* ring, size_t n, float cell,
                            const float tf[16])
{
    return set_scan_reduced(f, kReduceEvaluator, raw, ring, sizeof(int16_t), n, cell, tf);
}

int tsl_filter_set_scan_cloud(tsl_filter* f, const tsl_point* raw, const int32_t* ring, size_t n,
                              const float tf[16])
{
    return set_scan_reduced(f, kReduceCloud, raw, ring, sizeof(int32_t), n, 0.064f, tf);
}

int tsl_evaluate_cloud(tsl_evaluator* ev, tsl_particle* particles, size_t P, const tsl_point* raw,
                       const int32_t* ring, size_t n, const float tf[16], tsl_pose* out_pose)
{
    if (!ev || !out_pose || !tf || (P && !particles) || (n && (!raw || !ring)))
        return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    clear_error();
    tsl_filter* f;
    TSL_TRY(scratch_filter(ev, P, &f));
    TSL_TRY(tsl_filter_set_scan_cloud(f, raw, ring, n, tf));
    if (f->num_points == 0) {  // evaluate(particles, {}, tf): cuda_evaluator.cu:122-125
        std::memset(out_pose, 0, sizeof(*out_pose));
        return TSL_OK;
    }
    return evaluate_filter_scan(ev, f, particles, P, out_pose);
}

int tsl_filter_get_scan(tsl_filter* f, tsl_point* out, size_t capacity, size_t* n)
{
    if (!f || !n) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    TSL_HIP_CHECK(hipSetDevice(f->device));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    *n = f->num_points;
    if (f->num_points > capacity) return set_error(TSL_ERR_CAPACITY, "scan holds %zu points", f->num_points);
    if (f->num_points && out)
        TSL_HIP_CHECK(hipMemcpy(out, f->points_in.ptr, sizeof(tsl_point) * f->num_points, hipMemcpyDeviceToHost));
    return TSL_OK;
}

int tsl_resample_systematic(const tsl_particle* in, size_t M, float u0, tsl_particle* out, size_t out_capacity,
                            int32_t* out_indices, size_t* out_count)
{
    if (!in || !out_count || M == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument or M == 0");
    if (M > 0x7fffffffull) return set_error(TSL_ERR_UNSUPPORTED, "too many particles");
    size_t cap = 0;
    TSL_TRY(resample_capacity(M, u0, &cap));
    if (cap > 0xffffffffull) return set_error(TSL_ERR_UNSUPPORTED, "resampling output too large");
    DevBuf d_in, d_out, d_s, d_u, d_idx, d_st, d_plan, d_mem;
    TSL_TRY(d_in.reserve(sizeof(tsl_particle) * M));
    TSL_TRY(d_out.reserve(sizeof(tsl_particle) * cap));
    TSL_TRY(d_s.reserve(sizeof(double) * M));
    TSL_TRY(d_u.reserve(sizeof(float) * cap));
    TSL_TRY(d_idx.reserve(sizeof(int32_t) * cap));
    TSL_TRY(d_st.reserve(sizeof(FilterState)));
    TSL_TRY(d_plan.reserve(sizeof(RsPlan)));
    TSL_TRY(d_mem.reserve(sizeof(Member)));
    TSL_HIP_CHECK(hipMemcpy(d_in.ptr, in, sizeof(tsl_particle) * M, hipMemcpyHostToDevice));
    FilterState hs{};
    hs.count = (uint32_t)M;
    hs.first_bad = 0xffffffffu;
    TSL_HIP_CHECK(hipMemcpy(d_st.ptr, &hs, sizeof(hs), hipMemcpyHostToDevice));
    Member h;
    std::memset(&h, 0, sizeof(h));
    h.particles = d_in.as<tsl_particle>();
    h.alt = d_out.as<tsl_particle>();
    h.st = d_st.as<FilterState>();
    h.cumsum = d_s.as<double>();
    h.ubuf = d_u.as<float>();
    h.idx_out = d_idx.as<int32_t>();
    h.plan = d_plan.as<RsPlan>();
    h.cap_limit = (uint32_t)cap;
    TSL_HIP_CHECK(hipMemcpy(d_mem.ptr, &h, sizeof(h), hipMemcpyHostToDevice));
    TSL_TRY(launch_resample_members(d_mem.as<Member>(), 1, (uint32_t)cap, u0, false, 0));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    TSL_HIP_CHECK(hipMemcpy(&hs, d_st.ptr, sizeof(hs), hipMemcpyDeviceToHost));
    if (hs.status == kStCapacity) return set_error(TSL_ERR_CAPACITY, "weights sum beyond 1 + 1e-5: normalise them first");
    if (hs.status != kStOk) return set_error(TSL_ERR_UNSUPPORTED, "resampling failed (status %u)", hs.status);
    const uint32_t J = hs.resample_count;
    *out_count = J;
    if (J > out_capacity) return set_error(TSL_ERR_CAPACITY, "resampling emits %u particles", J);
    if (out && J) TSL_HIP_CHECK(hipMemcpy(out, d_out.ptr, sizeof(tsl_particle) * J, hipMemcpyDeviceToHost));
    if (out_indices && J) TSL_HIP_CHECK(hipMemcpy(out_indices, d_idx.ptr, sizeof(int32_t) * J, hipMemcpyDeviceToHost));
    return TSL_OK;
}

int tsl_stream_create(void** stream)
{
    if (!stream) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    hipStream_t s;
    TSL_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return TSL_OK;
}

int tsl_stream_destroy(void* stream)
{
    if (stream) TSL_HIP_CHECK(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    return TSL_OK;
}

int tsl_stream_synchronize(void* stream)
{
    TSL_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return TSL_OK;
}

int tsl_event_create(void** event)
{
    if (!event) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    hipEvent_t e;
    TSL_HIP_CHECK(hipEventCreate(&e));
    *event = e;
    return TSL_OK;
}

int tsl_event_destroy(void* event)
{
    if (event) TSL_HIP_CHECK(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return TSL_OK;
}

int tsl_event_record(void* event, void* stream)
{
    TSL_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
    return TSL_OK;
}

int tsl_event_elapsed_ms(void* start, void* stop, float* ms)
{
    if (!ms) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    TSL_HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    TSL_HIP_CHECK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return TSL_OK;
}

}  // extern "C"
