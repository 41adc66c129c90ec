// Batched sensor update: many device-resident filters that share one map
// and sensor model are updated and resampled by one fixed sequence of
// launches (a persistent evaluation kernel walks every member's work
// items), optionally captured once into a HIP graph and replayed per step.
// No counterpart in the reference, which updates one filter per process
// (src/mcl_3d.cpp:378-425).
#include "tsl_internal.h"

#include <algorithm>
#include <cstring>
#include <vector>

using namespace tsl;

struct tsl_batch {
    tsl_evaluator* ev = nullptr;
    int device = 0;
    std::vector<tsl_filter*> filters;
    int order = 0, K = 1, P = 1;
    bool rank = true;        // block offsets from the LDS occupancy-rank table when the map has one
    uint32_t chunk_len = 0;  // 0: batch_chunk of each member's scan
    DevBuf members, plan, u0_dev, states, motions;
    std::vector<Member> host_members;
    BatchLaunch launch;
    float* u0_host = nullptr;        // pinned staging of the per-member draws
    tsl_motion* motion_host = nullptr;  // pinned staging of the per-member motion steps
    hipEvent_t motion_copied = nullptr;
    bool motion_pending = false;
    FilterState* states_host = nullptr;  // pinned
    hipEvent_t u0_copied = nullptr;
    bool u0_pending = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool prepared = false;
};

namespace {

__global__ void k_gather_states(const Member* __restrict__ members, uint32_t n, FilterState* __restrict__ out)
{
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < n) out[f] = *members[f].st;
}

void drop_graph(tsl_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// (Re)build every member's descriptor and reserve its buffers for the
// largest state one step can reach.
int prepare(tsl_batch* b)
{
    const uint32_t n = (uint32_t)b->filters.size();
    TSL_HIP_CHECK(hipSetDevice(b->device));
    TSL_HIP_CHECK(hipDeviceSynchronize());
    drop_graph(b);
    TSL_TRY(b->members.reserve(sizeof(Member) * n));
    TSL_TRY(b->plan.reserve(sizeof(BatchPlan)));
    TSL_TRY(b->u0_dev.reserve(sizeof(float) * n));
    TSL_TRY(b->states.reserve(sizeof(FilterState) * n));
    TSL_TRY(b->motions.reserve(sizeof(tsl_motion) * n));
    uint64_t max_cap = 0, max_items = 0;
    for (tsl_filter* f : b->filters) {
        const uint64_t need = resample_bound(std::max<size_t>(f->count_bound, f->snapshot_count));
        if (need == 0 || need > 0x7fffffffull)
            return set_error(TSL_ERR_UNSUPPORTED, "systematic draw does not advance for %zu particles", f->count_bound);
        if (need > f->capacity) TSL_TRY(filter_reserve_particles(f, need));
        max_cap = std::max<uint64_t>(max_cap, f->capacity);
    }
    const uint64_t nb = (max_cap + 255) / 256;
    b->host_members.assign(n, Member{});
    for (uint32_t i = 0; i < n; ++i) {
        tsl_filter* f = b->filters[i];
        TSL_TRY(f->ubuf.reserve(sizeof(float) * f->capacity));
        TSL_TRY(f->plan.reserve(sizeof(RsPlan)));
        TSL_TRY(f->blocks.reserve(sizeof(double) * kReducePartials * nb));
        const int saved_k = f->tile_k, saved_p = f->tile_p;
        const uint32_t saved_chunk = f->tile_chunk;
        f->tile_k = b->K;
        f->tile_p = b->P;
        f->tile_chunk = b->chunk_len ? b->chunk_len : batch_chunk(f->num_points);
        const Tile t = choose_tile(f, f->capacity, f->num_points);
        f->tile_k = saved_k;
        f->tile_p = saved_p;
        f->tile_chunk = saved_chunk;
        TSL_TRY(f->partial.reserve(sizeof(float) * (size_t)t.chunks * f->capacity));
        Member& h = b->host_members[i];
        std::memset(&h, 0, sizeof(h));
        h.particles = f->particles.as<tsl_particle>();
        h.alt = f->particles_alt.as<tsl_particle>();
        h.L = f->L.as<float>();
        h.points = f->points.as<float4>();
        h.partial = f->partial.as<float>();
        h.raw = f->raw.as<float>();
        h.inv = nullptr;
        h.blocks = f->blocks.as<double>();
        h.st = f->state.as<FilterState>();
        h.cumsum = f->cumsum.as<double>();
        h.ubuf = f->ubuf.as<float>();
        h.idx_out = nullptr;
        h.plan = f->plan.as<RsPlan>();
        h.u0_dev = b->u0_dev.as<float>() + i;
        h.snapshot = f->snapshot.as<tsl_particle>();
        std::memcpy(h.tf, f->tf, sizeof(h.tf));
        h.trig = f->ev->trig;
        h.N = (uint32_t)f->num_points;
        h.chunk_len = t.chunk_len;
        h.chunks = t.chunks;
        h.stride = (uint32_t)f->capacity;
        h.cap_limit = (uint32_t)f->capacity;
        h.snapshot_count = (uint32_t)f->snapshot_count;
        const uint64_t per = 4ull * (uint64_t)b->K;  // waves per workgroup x K
        max_items += (f->capacity + per - 1) / per * t.chunks;
    }
    TSL_HIP_CHECK(hipMemcpy(b->members.ptr, b->host_members.data(), sizeof(Member) * n, hipMemcpyHostToDevice));
    BatchLaunch& l = b->launch;
    l.map = b->ev->map;
    l.a_hit = b->ev->a_hit;
    l.members = b->members.as<Member>();
    l.plan = b->plan.as<BatchPlan>();
    l.n = n;
    l.K = b->K;
    l.P = b->P;
    l.max_items = max_items;
    l.max_cap = (uint32_t)max_cap;
    l.max_rs = (uint32_t)max_cap;
    l.rank = b->rank;
    TSL_TRY(batch_grid(l, &l.grid));
    for (tsl_filter* f : b->filters) {
        f->count_bound = f->capacity;  // the device count is authoritative from here on
        f->count_known = false;
    }
    b->prepared = true;
    return TSL_OK;
}

int stage_u0(tsl_batch* b, const float* u0, hipStream_t s)
{
    const uint32_t n = (uint32_t)b->filters.size();
    if (b->u0_pending) {
        TSL_HIP_CHECK(hipEventSynchronize(b->u0_copied));  // staging buffer free again
        b->u0_pending = false;
    }
    std::memcpy(b->u0_host, u0, sizeof(float) * n);
    TSL_HIP_CHECK(hipMemcpyAsync(b->u0_dev.ptr, b->u0_host, sizeof(float) * n, hipMemcpyHostToDevice, s));
    TSL_HIP_CHECK(hipEventRecord(b->u0_copied, s));
    b->u0_pending = true;
    return TSL_OK;
}

int enqueue_step(tsl_batch* b, bool restore, bool resample, hipStream_t s)
{
    if (restore) TSL_TRY(launch_batch_restore(b->launch, s));
    TSL_TRY(launch_batch_update(b->launch, s));
    if (resample) TSL_TRY(launch_resample_members(b->launch.members, b->launch.n, b->launch.max_rs, 0.0f, true, s));
    return TSL_OK;
}

}  // namespace

extern "C" {

int tsl_batch_create(tsl_evaluator* ev, tsl_filter* const* filters, size_t n, tsl_batch** out)
{
    if (!ev || !out || (n && !filters)) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n == 0 || n > kMaxMembers) return set_error(TSL_ERR_INVALID_ARGUMENT, "a batch holds 1 to %u filters", kMaxMembers);
    *out = nullptr;
    for (size_t i = 0; i < n; ++i) {
        if (!filters[i]) return set_error(TSL_ERR_INVALID_ARGUMENT, "filter %zu is NULL", i);
        if (filters[i]->ev != ev) return set_error(TSL_ERR_INVALID_ARGUMENT, "filter %zu uses another evaluator", i);
        if (filters[i]->num_points == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "filter %zu has no scan", i);
        for (size_t j = 0; j < i; ++j)
            if (filters[j] == filters[i]) return set_error(TSL_ERR_INVALID_ARGUMENT, "filter %zu listed twice", i);
    }
    TSL_HIP_CHECK(hipSetDevice(ev->device));
    tsl_batch* b = new (std::nothrow) tsl_batch();
    if (!b) return set_error(TSL_ERR_OUT_OF_MEMORY, "out of host memory");
    b->ev = ev;
    b->device = ev->device;
    b->filters.assign(filters, filters + n);
    int rc = TSL_OK;
    if (hipHostMalloc(reinterpret_cast<void**>(&b->u0_host), sizeof(float) * n, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&b->states_host), sizeof(FilterState) * n, hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&b->motion_host), sizeof(tsl_motion) * n, hipHostMallocDefault) !=
            hipSuccess ||
        hipEventCreateWithFlags(&b->u0_copied, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->motion_copied, hipEventDisableTiming) != hipSuccess)
        rc = set_error(TSL_ERR_OUT_OF_MEMORY, "pinned host buffers");
    if (rc == TSL_OK) rc = prepare(b);
    if (rc != TSL_OK) {
        tsl_batch_destroy(b);
        return rc;
    }
    *out = b;
    return TSL_OK;
}

int tsl_batch_destroy(tsl_batch* b)
{
    if (!b) return TSL_OK;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    drop_graph(b);
    if (b->u0_copied) (void)hipEventDestroy(b->u0_copied);
    if (b->u0_host) (void)hipHostFree(b->u0_host);
    if (b->states_host) (void)hipHostFree(b->states_host);
    if (b->motion_host) (void)hipHostFree(b->motion_host);
    if (b->motion_copied) (void)hipEventDestroy(b->motion_copied);
    delete b;
    return TSL_OK;
}

int tsl_batch_set_tile(tsl_batch* b, int k, int points_per_lane, uint32_t chunk_len)
{
    if (!b) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL batch");
    if (k != 1 && k != 2) return set_error(TSL_ERR_INVALID_ARGUMENT, "K must be 1 or 2");
    if (points_per_lane != 1 && points_per_lane != 2 && points_per_lane != 4)
        return set_error(TSL_ERR_INVALID_ARGUMENT, "points per lane must be 1, 2 or 4");
    b->K = k;
    b->P = points_per_lane;
    b->chunk_len = chunk_len;
    return prepare(b);
}

int tsl_batch_prepare(tsl_batch* b)
{
    if (!b) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL batch");
    return prepare(b);
}

int tsl_batch_autotune(tsl_batch* b, void* stream, int reps, int* best_k, int* best_points_per_lane,
                       uint32_t* best_chunk, float* best_ms)
{
    if (!b || !b->prepared) return set_error(TSL_ERR_INVALID_ARGUMENT, "batch not prepared");
    TSL_HIP_CHECK(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (reps < 1) reps = 1;
    const int ks[2] = {1, 2};
    const int ps[3] = {1, 2, 4};
    const uint32_t chunks[4] = {0u, 1024u, 4096u, 8192u};
    const int saved_k = b->K, saved_p = b->P;
    const uint32_t saved_chunk = b->chunk_len;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TSL_HIP_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return set_error(TSL_ERR_HIP, "hipEventCreate failed");
    }
    float best = 1e30f;
    int bk = saved_k, bp = saved_p;
    uint32_t bc = saved_chunk;
    int rc = TSL_OK;
    for (int ki = 0; ki < 2 && rc == TSL_OK; ++ki)
        for (int pi = 0; pi < 3 && rc == TSL_OK; ++pi)
            for (int ci = 0; ci < 4 && rc == TSL_OK; ++ci) {
                b->K = ks[ki];
                b->P = ps[pi];
                b->chunk_len = chunks[ci];
                rc = prepare(b);
                if (rc == TSL_OK) rc = enqueue_step(b, false, false, s);  // warm-up
                if (rc != TSL_OK) break;
                (void)hipEventRecord(e0, s);
                for (int r = 0; r < reps && rc == TSL_OK; ++r) rc = enqueue_step(b, false, false, s);
                (void)hipEventRecord(e1, s);
                if (hipEventSynchronize(e1) != hipSuccess) rc = set_error(TSL_ERR_HIP, "autotune launch failed");
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                ms /= (float)reps;
                if (rc == TSL_OK && ms < best) {
                    best = ms;
                    bk = b->K;
                    bp = b->P;
                    bc = b->chunk_len;
                }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->K = rc == TSL_OK ? bk : saved_k;
    b->P = rc == TSL_OK ? bp : saved_p;
    b->chunk_len = rc == TSL_OK ? bc : saved_chunk;
    const int rc2 = prepare(b);
    if (rc != TSL_OK) return rc;
    TSL_TRY(rc2);
    if (best_k) *best_k = bk;
    if (best_points_per_lane) *best_points_per_lane = bp;
    if (best_chunk) *best_chunk = bc;
    if (best_ms) *best_ms = best;
    return TSL_OK;
}

int tsl_batch_restore_async(tsl_batch* b, void* stream)
{
    if (!b || !b->prepared) return set_error(TSL_ERR_INVALID_ARGUMENT, "batch not prepared");
    TSL_HIP_CHECK(hipSetDevice(b->device));
    return launch_batch_restore(b->launch, static_cast<hipStream_t>(stream));
}

int tsl_batch_update_async(tsl_batch* b, void* stream)
{
    if (!b || !b->prepared) return set_error(TSL_ERR_INVALID_ARGUMENT, "batch not prepared");
    TSL_HIP_CHECK(hipSetDevice(b->device));
    return enqueue_step(b, false, false, static_cast<hipStream_t>(stream));
}

int tsl_batch_resample_async(tsl_batch* b, const float* u0, void* stream)
{
    if (!b || !u0 || !b->prepared) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument or batch not prepared");
    TSL_HIP_CHECK(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    TSL_TRY(stage_u0(b, u0, s));
    return launch_resample_members(b->launch.members, b->launch.n, b->launch.max_rs, 0.0f, true, s);
}

int tsl_batch_step_async(tsl_batch* b, const float* u0, int restore, void* stream)
{
    if (!b || !u0 || !b->prepared) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument or batch not prepared");
    TSL_HIP_CHECK(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    TSL_TRY(stage_u0(b, u0, s));
    return enqueue_step(b, restore != 0, true, s);
}

int tsl_batch_capture(tsl_batch* b, int restore, void* stream)
{
    if (!b || !b->prepared) return set_error(TSL_ERR_INVALID_ARGUMENT, "batch not prepared");
    if (!stream) return set_error(TSL_ERR_INVALID_ARGUMENT, "graph capture needs a created stream");
    TSL_HIP_CHECK(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    drop_graph(b);
    TSL_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_step(b, restore != 0, true, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != TSL_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return set_error(TSL_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = g;
    TSL_HIP_CHECK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    return TSL_OK;
}

int tsl_batch_launch(tsl_batch* b, const float* u0, void* stream)
{
    if (!b || !u0 || !b->exec) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument or no captured step");
    TSL_HIP_CHECK(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    TSL_TRY(stage_u0(b, u0, s));
    TSL_HIP_CHECK(hipGraphLaunch(b->exec, s));
    return TSL_OK;
}

int tsl_batch_motion_async(tsl_batch* b, const tsl_motion* motions, void* stream)
{
    if (!b || !motions || !b->prepared) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument or batch not prepared");
    TSL_HIP_CHECK(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t n = b->launch.n;
    if (b->motion_pending) {
        TSL_HIP_CHECK(hipEventSynchronize(b->motion_copied));
        b->motion_pending = false;
    }
    std::memcpy(b->motion_host, motions, sizeof(tsl_motion) * n);
    TSL_HIP_CHECK(hipMemcpyAsync(b->motions.ptr, b->motion_host, sizeof(tsl_motion) * n, hipMemcpyHostToDevice, s));
    TSL_HIP_CHECK(hipEventRecord(b->motion_copied, s));
    b->motion_pending = true;
    return launch_motion_members(b->launch.members, b->motions.as<tsl_motion>(), n, b->launch.max_cap, s);
}

int tsl_batch_sync(tsl_batch* b, void* stream, tsl_pose* poses, uint64_t* counts, int* status)
{
    if (!b || !b->prepared) return set_error(TSL_ERR_INVALID_ARGUMENT, "batch not prepared");
    TSL_HIP_CHECK(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t n = b->launch.n;
    hipLaunchKernelGGL(k_gather_states, dim3((n + 63) / 64), dim3(64), 0, s, b->launch.members, n,
                       b->states.as<FilterState>());
    TSL_HIP_CHECK(hipGetLastError());
    TSL_HIP_CHECK(hipMemcpyAsync(b->states_host, b->states.ptr, sizeof(FilterState) * n, hipMemcpyDeviceToHost, s));
    TSL_HIP_CHECK(hipStreamSynchronize(s));
    int first_error = TSL_OK;
    for (uint32_t i = 0; i < n; ++i) {
        const FilterState& st = b->states_host[i];
        if (counts) counts[i] = st.count;
        if (poses) pose_from_stats(st.stats, &poses[i]);
        if (status) status[i] = (int)st.status;
        if (st.status != kStOk && first_error == TSL_OK) first_error = status_error(st.status);
    }
    return first_error;
}

}  // extern "C"
