// Serving loop: many robots' particle sets stay resident on the device;
// requests (odometry step + scan + resampling draw) arrive from any thread;
// a worker thread keeps the GPU busy with continuous batching -- every
// "tick" takes the pending requests of robots that are not in flight (at
// most max_batch, oldest first), enqueues motion update, scan preparation
// (one upload and one key-then-scan sort of all scans), batched sensor update
// and resampling for all of them, and completes
// earlier ticks as their events fire, so requests that arrive while others
// run are admitted into the next tick at once.  Robots are admitted against
// an HBM budget.  Each robot's result is the same whether it runs alone or
// in a mix (member-indexed kernels, counter-based motion noise).
//
// Robots come in two classes: "bulk" robots (particle sets of 65,536 slots
// and more, e.g. global localisation over a whole map) and the rest
// ("tracking").  Each class has its own stream and its own ticks, so a
// tracking request never queues behind a bulk one; a bulk tick's evaluation
// is cut into slices of a bounded number of particle-point pairs, launched
// one after the other on a grid that leaves two workgroup slots per compute
// unit free, so tracking kernels (of this server or of another server on
// the same GPU) start within about a slice.  Slicing changes no result bit.
// The two class streams are shared by every server of a device (one per
// map, say): a process then uses two hardware queues for serving, so a
// tracking stream never shares a queue with a bulk one.
// No counterpart in the reference (one robot per ROS node, src/mcl_3d.cpp).
#include "tsl_internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

using namespace tsl;

namespace {

constexpr int kMaxInflight = 3;    // ticks in flight per class
constexpr int kClasses = 2;        // 0 tracking, 1 bulk
constexpr uint64_t kDefaultSlicePairs = 1ull << 27;  // ~1 ms of evaluation per bulk slice on an HBM-resident map
constexpr uint32_t kSegmentedScans = 2;  // ticks with this many scans prepare them together (two device-wide sorts)
// Robots with at least this many particle slots are evaluated in spatial
// particle order (a global-localisation set spread over the whole map: the
// order keeps one XCD's particles in the same part of the map); smaller sets
// stay in their own order.  The order changes no result bit.
constexpr size_t kSpatialOrderMin = 65536;

struct Robot {
    tsl_filter* f = nullptr;
    uint32_t max_points = 0;
    uint64_t bytes = 0;
    bool active = false;    // accepts requests
    bool busy = false;      // in a tick, or its particles being read
    bool removing = false;  // being removed: neither active nor free for a new robot
    bool bulk = false;      // evaluated in the bulk class
};

struct Pending {
    tsl_request req;
    std::vector<tsl_point> points;
    double submit_ms;
};

struct Slot {  // one tick in flight
    std::vector<uint32_t> robots;
    std::vector<Pending> reqs;
    DevBuf members, plan, u0, motions, states;
    Member* members_host = nullptr;      // pinned
    float* u0_host = nullptr;            // pinned
    tsl_motion* motions_host = nullptr;  // pinned
    FilterState* states_host = nullptr;  // pinned
    // the tick's scans, prepared together: concatenated points (pinned
    // staging and device copy), segment offsets, destinations, sort buffers
    DevBuf scan_pts, scan_off, scan_dst, scan_keys, scan_vals, scan_temp;
    tsl_point* scan_host = nullptr;      // pinned, scan_cap points
    uint32_t* off_host = nullptr;        // pinned, max_batch + 1
    ScanDest* dst_host = nullptr;        // pinned, max_batch
    size_t scan_cap = 0;
    hipEvent_t done = nullptr;
    double start_ms = 0.0;
    bool used = false;
    int cls = 0;                         // robot class of the tick
};

// The class streams of a device, shared by its servers (reference counted).
constexpr int kMaxDevices = 64;
struct ClassStreams {
    int refs = 0;
    hipStream_t s[kClasses] = {nullptr, nullptr};
};
std::mutex g_streams_mu;
ClassStreams g_streams[kMaxDevices];

int acquire_streams(int device, hipStream_t out[kClasses])
{
    if (device < 0 || device >= kMaxDevices) return set_error(TSL_ERR_UNSUPPORTED, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_streams_mu);
    ClassStreams& cs = g_streams[device];
    if (cs.refs == 0) {
        for (int c = 0; c < kClasses; ++c)
            if (hipStreamCreateWithFlags(&cs.s[c], hipStreamNonBlocking) != hipSuccess) {
                for (int k = 0; k < c; ++k) (void)hipStreamDestroy(cs.s[k]);
                return set_error(TSL_ERR_HIP, "hipStreamCreate failed");
            }
    }
    ++cs.refs;
    for (int c = 0; c < kClasses; ++c) out[c] = cs.s[c];
    return TSL_OK;
}

void release_streams(int device)
{
    std::lock_guard<std::mutex> lk(g_streams_mu);
    ClassStreams& cs = g_streams[device];
    if (cs.refs <= 0 || --cs.refs > 0) return;
    for (int c = 0; c < kClasses; ++c) {
        (void)hipStreamSynchronize(cs.s[c]);
        (void)hipStreamDestroy(cs.s[c]);
        cs.s[c] = nullptr;
    }
}

__global__ void k_gather_states_s(const Member* __restrict__ members, uint32_t n, FilterState* __restrict__ out)
{
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < n) out[f] = *members[f].st;
}

}  // namespace

struct tsl_server {
    tsl_evaluator* ev = nullptr;
    int device = 0;
    uint64_t budget = 0, used = 0;
    uint32_t max_batch = 64;
    uint32_t chunk_len = 0;        // 0: each request's whole scan is one chunk
    int K[kClasses] = {2, 1};      // particles per wavefront of the batched evaluation, per class
    int P[kClasses] = {1, 2};      // points per lane, per class
    uint64_t max_particles = 0;    // per robot; sizes the shared reduction grid
    std::atomic<uint32_t> max_points{0};  // largest scan any admitted robot may send
    std::vector<Robot> robots;     // fixed size (max_robots): entries never move
    std::mutex mu;
    std::condition_variable cv_req, cv_done;
    std::deque<Pending> pending;
    std::deque<tsl_result> done;
    std::deque<Slot*> inflight[kClasses];
    std::vector<Slot*> free_slots[kClasses];
    Slot slots[kClasses * kMaxInflight];
    hipStream_t stream[kClasses] = {nullptr, nullptr};
    uint64_t slice_pairs = kDefaultSlicePairs;  // bulk evaluation pairs per launch (0: no slicing)
    uint32_t reserved_per_cu = 2;  // workgroup slots per CU a sliced bulk evaluation leaves to the other class
    std::thread worker;
    std::atomic<bool> stop{false};
    bool stopped = false;          // worker: stop requested and nothing left (worker thread only)
    std::chrono::steady_clock::time_point t0;
    int per_cu[kClasses] = {0, 0}, cus = 0;
    int worker_error = TSL_OK;
    uint64_t ticks = 0, batched_requests = 0;

    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

void complete(tsl_server* s, Slot* sl)
{
    const double t = s->now_ms();
    std::lock_guard<std::mutex> lk(s->mu);
    for (size_t i = 0; i < sl->robots.size(); ++i) {
        const FilterState& st = sl->states_host[i];
        tsl_result r;
        std::memset(&r, 0, sizeof(r));
        r.id = sl->reqs[i].req.id;
        r.robot = sl->robots[i];
        r.status = (int)st.status;
        r.count = st.count;
        pose_from_stats(st.stats, &r.pose);
        r.submit_ms = sl->reqs[i].submit_ms;
        r.start_ms = sl->start_ms;
        r.done_ms = t;
        r.batch = (uint32_t)sl->robots.size();
        Robot& rb = s->robots[sl->robots[i]];
        rb.busy = false;
        rb.f->count_bound = rb.f->capacity;
        rb.f->count_known = false;
        s->done.push_back(r);
    }
    sl->reqs.clear();
    sl->robots.clear();
    sl->used = false;
    s->free_slots[sl->cls].push_back(sl);
    s->cv_done.notify_all();
}

// Grow a free slot's scan buffers to hold `total` points of up to
// max_batch scans (worker thread only, so no tick in flight uses them);
// sized for max_batch full scans of the largest robot, so they grow at most
// once per admission of a robot with a larger scan.
int slot_reserve_scans(tsl_server* s, Slot& sl, size_t total)
{
    const size_t want = std::max<size_t>(total, (size_t)s->max_batch * s->max_points.load());
    if (want <= sl.scan_cap && sl.scan_host) return TSL_OK;
    const size_t cap = std::max<size_t>(want, 1);
    if (sl.scan_host) (void)hipHostFree(sl.scan_host);
    sl.scan_host = nullptr;
    sl.scan_cap = 0;
    if (hipHostMalloc(reinterpret_cast<void**>(&sl.scan_host), sizeof(tsl_point) * cap, hipHostMallocDefault) !=
        hipSuccess)
        return set_error(TSL_ERR_OUT_OF_MEMORY, "pinned scan staging");
    TSL_TRY(sl.scan_pts.reserve(sizeof(tsl_point) * cap));
    TSL_TRY(sl.scan_keys.reserve(sizeof(uint64_t) * 2 * cap));
    TSL_TRY(sl.scan_vals.reserve(sizeof(uint32_t) * 2 * cap));
    const size_t temp = prep_points_batch_temp((uint32_t)cap, s->max_batch);
    if (temp == 0) return set_error(TSL_ERR_HIP, "scan sort temporary size");
    TSL_TRY(sl.scan_temp.reserve(temp));
    sl.scan_cap = cap;
    return TSL_OK;
}

// Enqueue one tick for the selected requests (robots marked busy already).
int dispatch(tsl_server* s, Slot* sl)
{
    hipStream_t st = s->stream[sl->cls];
    const uint32_t n = (uint32_t)sl->reqs.size();
    uint64_t max_items = 0;
    uint32_t max_cap = 0;
    // every scan of the tick: one upload, one preparation
    uint32_t total = 0, max_np = 0;
    for (uint32_t i = 0; i < n; ++i) {
        sl->off_host[i] = total;
        const uint32_t np = (uint32_t)sl->reqs[i].points.size();
        total += np;
        max_np = std::max(max_np, np);
    }
    sl->off_host[n] = total;
    TSL_TRY(slot_reserve_scans(s, *sl, total));
    if (total) TSL_TRY(sl->scan_temp.reserve(prep_points_batch_temp(total, n)));
    for (uint32_t i = 0; i < n; ++i) {
        const Pending& pd = sl->reqs[i];
        if (!pd.points.empty())
            std::memcpy(sl->scan_host + sl->off_host[i], pd.points.data(), sizeof(tsl_point) * pd.points.size());
        tsl_filter* f = s->robots[sl->robots[i]].f;
        sl->dst_host[i] = ScanDest{f->points.as<float4>(), f->perm.as<uint32_t>()};
    }
    if (n >= kSegmentedScans) {
        // one upload; all points sorted by key, then by scan (both stable)
        if (total)
            TSL_HIP_CHECK(hipMemcpyAsync(sl->scan_pts.ptr, sl->scan_host, sizeof(tsl_point) * total,
                                         hipMemcpyHostToDevice, st));
        TSL_HIP_CHECK(
            hipMemcpyAsync(sl->scan_off.ptr, sl->off_host, sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice, st));
        TSL_HIP_CHECK(hipMemcpyAsync(sl->scan_dst.ptr, sl->dst_host, sizeof(ScanDest) * n, hipMemcpyHostToDevice, st));
        TSL_TRY(launch_prep_points_batch(s->ev, sl->scan_pts.as<tsl_point>(), sl->scan_off.as<uint32_t>(),
                                         sl->scan_dst.as<ScanDest>(), n, total, max_np, sl->scan_keys.as<uint64_t>(),
                                         sl->scan_vals.as<uint32_t>(), sl->scan_temp.ptr, sl->scan_temp.bytes, st));
    } else {
        // few scans: the device-wide sort of each (same keys, same stable
        // order, so the same points in the same order)
        for (uint32_t i = 0; i < n; ++i) {
            tsl_filter* f = s->robots[sl->robots[i]].f;
            const uint32_t np = sl->off_host[i + 1] - sl->off_host[i];
            f->num_points = np;
            if (np)
                TSL_HIP_CHECK(hipMemcpyAsync(f->points_in.ptr, sl->scan_host + sl->off_host[i], sizeof(tsl_point) * np,
                                             hipMemcpyHostToDevice, st));
            TSL_TRY(launch_prep_points(f, st));
        }
    }
    for (uint32_t i = 0; i < n; ++i) {
        Robot& rb = s->robots[sl->robots[i]];
        tsl_filter* f = rb.f;
        const Pending& pd = sl->reqs[i];
        const size_t np = pd.points.size();
        f->num_points = np;
        // tracking robots: the server's chunk length; bulk robots: the whole
        // scan per item (their particles alone give enough items, and one
        // chunk stages each pose matrix and scan tile once)
        const uint32_t chunk = (s->chunk_len && !rb.bulk) ? s->chunk_len
                                                          : (uint32_t)std::max<size_t>(64, (np + 63) / 64 * 64);
        const uint32_t chunks = (uint32_t)std::max<size_t>(1, (np + chunk - 1) / chunk);
        Member& h = sl->members_host[i];
        std::memset(&h, 0, sizeof(h));
        h.particles = f->particles.as<tsl_particle>();
        h.alt = f->particles_alt.as<tsl_particle>();
        h.L = f->L.as<float>();
        h.points = f->points.as<float4>();
        h.partial = f->partial.as<float>();
        h.raw = f->raw.as<float>();
        h.blocks = f->blocks.as<double>();
        h.st = f->state.as<FilterState>();
        h.cumsum = f->cumsum.as<double>();
        h.ubuf = f->ubuf.as<float>();
        h.plan = f->plan.as<RsPlan>();
        h.u0_dev = sl->u0.as<float>() + i;
        if (f->capacity >= kSpatialOrderMin) h.inv = f->pinv.as<uint32_t>();
        std::memcpy(h.tf, f->tf, sizeof(h.tf));
        h.trig = f->ev->trig;
        h.N = (uint32_t)np;
        h.chunk_len = chunk;
        h.chunks = chunks;
        h.stride = (uint32_t)f->capacity;
        h.cap_limit = (uint32_t)f->capacity;
        sl->u0_host[i] = pd.req.u0;
        sl->motions_host[i] = pd.req.motion;
        max_items += (f->capacity + 4 * s->K[sl->cls] - 1) / (4 * s->K[sl->cls]) * chunks;
        max_cap = std::max<uint32_t>(max_cap, (uint32_t)f->capacity);
    }
    TSL_HIP_CHECK(hipMemcpyAsync(sl->members.ptr, sl->members_host, sizeof(Member) * n, hipMemcpyHostToDevice, st));
    TSL_HIP_CHECK(hipMemcpyAsync(sl->u0.ptr, sl->u0_host, sizeof(float) * n, hipMemcpyHostToDevice, st));
    TSL_HIP_CHECK(hipMemcpyAsync(sl->motions.ptr, sl->motions_host, sizeof(tsl_motion) * n, hipMemcpyHostToDevice, st));
    BatchLaunch l;
    l.map = s->ev->map;
    l.a_hit = s->ev->a_hit;
    l.members = sl->members.as<Member>();
    l.plan = sl->plan.as<BatchPlan>();
    l.n = n;
    l.K = s->K[sl->cls];
    l.P = s->P[sl->cls];
    l.rank = true;
    l.max_items = max_items;
    l.max_cap = max_cap;
    l.max_rs = max_cap;
    // a bulk tick leaves one workgroup slot per compute unit to the other
    // stream and evaluates in slices of about slice_pairs pairs
    const int pc = s->per_cu[sl->cls];
    const int per_cu = (sl->cls == 1 && s->slice_pairs)
                           ? std::max(1, pc - (int)std::min<uint32_t>(s->reserved_per_cu, (uint32_t)pc))
                           : pc;
    uint64_t grid = (uint64_t)per_cu * (uint64_t)s->cus;
    const uint64_t need = (max_items + 7) / 8 * 8;
    if (grid > need) grid = need;
    grid = std::max<uint64_t>(8, (grid + 7) / 8 * 8);
    l.grid = (uint32_t)grid;
    if (sl->cls == 1 && s->slice_pairs) {
        uint64_t pairs_per_item = 0;
        for (uint32_t i = 0; i < n; ++i)
            pairs_per_item =
                std::max<uint64_t>(pairs_per_item, (uint64_t)4 * s->K[1] * sl->members_host[i].chunk_len);
        l.slice_items = std::max<uint64_t>(grid, s->slice_pairs / std::max<uint64_t>(pairs_per_item, 1));
    }
    TSL_TRY(launch_motion_members(l.members, sl->motions.as<tsl_motion>(), n, max_cap, st));
    for (uint32_t i = 0; i < n; ++i) {
        tsl_filter* f = s->robots[sl->robots[i]].f;
        if (f->capacity < kSpatialOrderMin) continue;
        // the device count is at most the capacity; the order's buffers
        // were reserved for it at admission
        f->count_bound = f->capacity;
        TSL_TRY(launch_pose_spatial(f, st));
    }
    TSL_TRY(launch_batch_update(l, st));
    TSL_TRY(launch_resample_members(l.members, n, max_cap, 0.0f, true, st));
    hipLaunchKernelGGL(k_gather_states_s, dim3((n + 63) / 64), dim3(64), 0, st, l.members, n,
                       sl->states.as<FilterState>());
    TSL_HIP_CHECK(hipGetLastError());
    TSL_HIP_CHECK(hipMemcpyAsync(sl->states_host, sl->states.ptr, sizeof(FilterState) * n, hipMemcpyDeviceToHost, st));
    TSL_HIP_CHECK(hipEventRecord(sl->done, st));
    return TSL_OK;
}

void worker_main(tsl_server* s)
{
    (void)hipSetDevice(s->device);
    while (true) {
        bool progressed = false;
        // completions, oldest first within each class's stream
        for (int c = 0; c < kClasses; ++c) {
            while (!s->inflight[c].empty()) {
                Slot* sl = s->inflight[c].front();
                const hipError_t q = hipEventQuery(sl->done);
                if (q == hipErrorNotReady) break;
                s->inflight[c].pop_front();
                if (q != hipSuccess) {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->worker_error = set_error(TSL_ERR_HIP, "tick failed: %s", hipGetErrorString(q));
                }
                complete(s, sl);
                progressed = true;
            }
        }
        // admit the next tick of each class
        for (int c = 0; c < kClasses; ++c) {
        Slot* sl = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (c == 0 && s->stop && s->pending.empty() && s->inflight[0].empty() && s->inflight[1].empty()) {
                s->stopped = true;
                break;
            }
            if (!s->free_slots[c].empty() && !s->pending.empty()) {
                Slot* cand = s->free_slots[c].back();
                std::vector<uint8_t> taken(s->robots.size(), 0);
                for (auto it = s->pending.begin(); it != s->pending.end() && cand->reqs.size() < s->max_batch;) {
                    const uint32_t r = it->req.robot;
                    Robot& rb = s->robots[r];
                    if (rb.active && !rb.busy && !taken[r] && (rb.bulk ? 1 : 0) == c) {
                        taken[r] = 1;
                        rb.busy = true;
                        cand->robots.push_back(r);
                        cand->reqs.push_back(std::move(*it));
                        it = s->pending.erase(it);
                    } else {
                        ++it;
                    }
                }
                if (!cand->reqs.empty()) {
                    s->free_slots[c].pop_back();
                    sl = cand;
                }
            }
        }
        if (sl) {
            sl->start_ms = s->now_ms();
            const int rc = dispatch(s, sl);
            if (rc != TSL_OK) {
                std::lock_guard<std::mutex> lk(s->mu);
                s->worker_error = rc;
            }
            sl->used = true;
            s->inflight[c].push_back(sl);
            s->ticks += 1;
            s->batched_requests += sl->reqs.size();
            progressed = true;
        }
        }
        if (s->stopped) break;
        if (!progressed) {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->inflight[0].empty() && s->inflight[1].empty()) {
                s->cv_req.wait_for(lk, std::chrono::milliseconds(2));
            } else {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
        }
    }
}

int slot_init(tsl_server* s, Slot& sl)
{
    const size_t n = s->max_batch;
    TSL_TRY(sl.members.reserve(sizeof(Member) * n));
    TSL_TRY(sl.plan.reserve(sizeof(BatchPlan)));
    TSL_TRY(sl.u0.reserve(sizeof(float) * n));
    TSL_TRY(sl.motions.reserve(sizeof(tsl_motion) * n));
    TSL_TRY(sl.states.reserve(sizeof(FilterState) * n));
    if (hipHostMalloc(reinterpret_cast<void**>(&sl.members_host), sizeof(Member) * n, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&sl.u0_host), sizeof(float) * n, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&sl.motions_host), sizeof(tsl_motion) * n, hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&sl.states_host), sizeof(FilterState) * n, hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&sl.off_host), sizeof(uint32_t) * (n + 1), hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&sl.dst_host), sizeof(ScanDest) * n, hipHostMallocDefault) != hipSuccess)
        return set_error(TSL_ERR_OUT_OF_MEMORY, "pinned host buffers");
    TSL_TRY(sl.scan_off.reserve(sizeof(uint32_t) * (n + 1)));
    TSL_TRY(sl.scan_dst.reserve(sizeof(ScanDest) * n));
    TSL_HIP_CHECK(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    return TSL_OK;
}

// Workgroups per compute unit of the batched evaluation kernel a class's
// tile selects on this map (its persistent grid is that times the CUs).
int class_occupancy(tsl_server* s, int c)
{
    BatchLaunch probe;
    probe.map = s->ev->map;
    probe.K = s->K[c];
    probe.P = s->P[c];
    probe.rank = true;
    probe.max_items = ~0ull >> 8;
    uint32_t g = 0;
    TSL_TRY(batch_grid(probe, &g));
    s->per_cu[c] = (int)std::max<uint32_t>(1, g / (uint32_t)s->cus);
    return TSL_OK;
}

uint64_t filter_bytes(const tsl_filter* f)
{
    const DevBuf* bufs[] = {&f->particles, &f->particles_alt, &f->L, &f->raw, &f->partial, &f->points, &f->points_in,
                            &f->perm, &f->sort_keys, &f->sort_vals, &f->sort_temp, &f->blocks, &f->state, &f->cumsum,
                            &f->ubuf, &f->plan, &f->member, &f->pkeys, &f->porder, &f->pinv};
    uint64_t b = 0;
    for (const DevBuf* d : bufs) b += d->bytes;
    return b;
}

}  // namespace

extern "C" {

int tsl_server_create(tsl_evaluator* ev, uint64_t hbm_budget_bytes, uint32_t max_robots, uint64_t max_particles,
                      uint32_t max_batch, uint32_t chunk_len, tsl_server** out)
{
    if (!ev || !out) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (max_robots == 0 || max_particles == 0)
        return set_error(TSL_ERR_INVALID_ARGUMENT, "max_robots and max_particles must be positive");
    if (resample_bound(max_particles) == 0 || resample_bound(max_particles) > 0x7fffffffull)
        return set_error(TSL_ERR_UNSUPPORTED, "max_particles too large");
    if (max_batch == 0 || max_batch > kMaxMembers)
        return set_error(TSL_ERR_INVALID_ARGUMENT, "max_batch must be in [1, %u]", kMaxMembers);
    *out = nullptr;
    TSL_HIP_CHECK(hipSetDevice(ev->device));
    tsl_server* s = new (std::nothrow) tsl_server();
    if (!s) return set_error(TSL_ERR_OUT_OF_MEMORY, "out of host memory");
    s->ev = ev;
    s->device = ev->device;
    s->budget = hbm_budget_bytes;
    s->max_batch = max_batch;
    s->chunk_len = chunk_len ? (chunk_len + 63) / 64 * 64 : 0;
    s->max_particles = max_particles;
    s->robots.resize(max_robots);
    s->t0 = std::chrono::steady_clock::now();
    uint64_t map_bytes = 0;
    (void)tsl_map_device_bytes(ev->map, &map_bytes);
    s->used = map_bytes;
    int rc = acquire_streams(s->device, s->stream);
    if (rc != TSL_OK) s->stream[0] = s->stream[1] = nullptr;
    for (int i = 0; i < kClasses * kMaxInflight && rc == TSL_OK; ++i) {
        rc = slot_init(s, s->slots[i]);
        s->slots[i].cls = i / kMaxInflight;
        s->free_slots[i / kMaxInflight].push_back(&s->slots[i]);
    }
    if (rc == TSL_OK) {
        int cus = 0;
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ev->device);
        s->cus = cus > 0 ? cus : 1;
        for (int c = 0; c < kClasses && rc == TSL_OK; ++c) rc = class_occupancy(s, c);
    }
    if (rc != TSL_OK) {
        tsl_server_destroy(s);
        return rc;
    }
    s->worker = std::thread(worker_main, s);
    *out = s;
    return TSL_OK;
}

int tsl_server_destroy(tsl_server* s)
{
    if (!s) return TSL_OK;
    if (s->worker.joinable()) {
        s->stop = true;
        s->cv_req.notify_all();
        s->worker.join();
    }
    (void)hipSetDevice(s->device);
    // the worker finished this server's ticks; the streams are shared
    for (Slot& sl : s->slots)
        if (sl.done) (void)hipEventSynchronize(sl.done);
    for (Robot& r : s->robots)
        if (r.f) tsl_filter_destroy(r.f);
    for (Slot& sl : s->slots) {
        if (sl.members_host) (void)hipHostFree(sl.members_host);
        if (sl.u0_host) (void)hipHostFree(sl.u0_host);
        if (sl.motions_host) (void)hipHostFree(sl.motions_host);
        if (sl.states_host) (void)hipHostFree(sl.states_host);
        if (sl.off_host) (void)hipHostFree(sl.off_host);
        if (sl.dst_host) (void)hipHostFree(sl.dst_host);
        if (sl.scan_host) (void)hipHostFree(sl.scan_host);
        if (sl.done) (void)hipEventDestroy(sl.done);
    }
    if (s->stream[0]) release_streams(s->device);
    delete s;
    return TSL_OK;
}

int tsl_server_add_robot(tsl_server* s, const tsl_particle* particles, size_t n, uint32_t max_points,
                         const float tf[16], uint32_t* robot)
{
    if (!s || !robot || !tf || (n && !particles)) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n == 0 || max_points == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "a robot needs particles and a scan size");
    if (n > s->max_particles)
        return set_error(TSL_ERR_INVALID_ARGUMENT, "%zu particles exceed the server's %llu per robot", n,
                         (unsigned long long)s->max_particles);
    TSL_HIP_CHECK(hipSetDevice(s->device));
    tsl_filter* f = nullptr;
    TSL_TRY(tsl_filter_create(s->ev, n, &f));
    int rc = tsl_filter_set_particles(f, particles, n);
    const uint64_t cap = resample_bound(n);
    if (rc == TSL_OK && (cap == 0 || cap > 0x7fffffffull))
        rc = set_error(TSL_ERR_UNSUPPORTED, "systematic draw does not advance for %zu particles", n);
    // everything a tick can touch, reserved now: ticks never allocate
    if (rc == TSL_OK) rc = filter_reserve_particles(f, cap);
    if (rc == TSL_OK) rc = filter_reserve_scan(f, max_points);
    const bool bulk = f->capacity >= kSpatialOrderMin;
    const size_t chunks = (s->chunk_len && !bulk) ? ((size_t)max_points + s->chunk_len - 1) / s->chunk_len : 1;
    if (rc == TSL_OK) rc = f->partial.reserve(sizeof(float) * std::max<size_t>(chunks, 1) * f->capacity);
    if (rc == TSL_OK) rc = f->ubuf.reserve(sizeof(float) * f->capacity);
    if (rc == TSL_OK) rc = f->plan.reserve(sizeof(RsPlan));
    // the reduction grid of a tick spans its largest member
    const uint64_t nb_max = (resample_bound(s->max_particles) + 255) / 256;
    if (rc == TSL_OK) rc = f->blocks.reserve(sizeof(double) * kReducePartials * nb_max);
    if (rc == TSL_OK && f->capacity >= kSpatialOrderMin) {
        f->count_bound = f->capacity;
        rc = reserve_particle_order(f);
    }
    if (rc == TSL_OK) std::memcpy(f->tf, tf, sizeof(float) * 16);
    if (rc != TSL_OK) {
        tsl_filter_destroy(f);
        return rc;
    }
    const uint64_t bytes = filter_bytes(f);
    std::lock_guard<std::mutex> lk(s->mu);
    uint32_t slot = 0;
    while (slot < s->robots.size() && (s->robots[slot].active || s->robots[slot].busy || s->robots[slot].removing))
        ++slot;
    if (slot == s->robots.size() || (s->budget && s->used + bytes > s->budget)) {
        tsl_filter_destroy(f);
        if (slot == s->robots.size())
            return set_error(TSL_ERR_ADMISSION, "server holds its maximum of %zu robots", s->robots.size());
        return set_error(TSL_ERR_ADMISSION, "robot needs %llu bytes; %llu of %llu in use",
                         (unsigned long long)bytes, (unsigned long long)s->used, (unsigned long long)s->budget);
    }
    s->used += bytes;
    Robot& r = s->robots[slot];
    r.f = f;
    r.max_points = max_points;
    r.bytes = bytes;
    r.active = true;
    r.busy = false;
    r.removing = false;
    r.bulk = f->capacity >= kSpatialOrderMin;
    if (max_points > s->max_points.load()) s->max_points.store(max_points);
    f->count_bound = f->capacity;
    f->count_known = false;
    *robot = slot;
    return TSL_OK;
}

int tsl_server_remove_robot(tsl_server* s, uint32_t robot)
{
    if (!s) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (robot >= s->robots.size() || !s->robots[robot].active)
        return set_error(TSL_ERR_INVALID_ARGUMENT, "no robot %u", robot);
    for (const Pending& p : s->pending)
        if (p.req.robot == robot) return set_error(TSL_ERR_INVALID_ARGUMENT, "robot %u has pending requests", robot);
    // no longer accepts requests (submit checks active under this lock), and
    // its slot is not reused while the removal waits for its tick
    Robot& r = s->robots[robot];
    r.active = false;
    r.removing = true;
    s->cv_done.wait(lk, [&] { return !r.busy; });
    s->used -= r.bytes;
    tsl_filter* f = r.f;
    r.f = nullptr;
    r.removing = false;
    lk.unlock();
    tsl_filter_destroy(f);
    return TSL_OK;
}

int tsl_server_submit(tsl_server* s, const tsl_request* req)
{
    if (!s || !req || (req->num_points && !req->points)) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    Pending p;
    p.req = *req;
    p.points.assign(req->points, req->points + req->num_points);
    p.req.points = nullptr;
    p.submit_ms = s->now_ms();
    std::lock_guard<std::mutex> lk(s->mu);
    if (req->robot >= s->robots.size() || !s->robots[req->robot].active)
        return set_error(TSL_ERR_INVALID_ARGUMENT, "no robot %u", req->robot);
    if (req->num_points > s->robots[req->robot].max_points)
        return set_error(TSL_ERR_CAPACITY, "scan of %zu points exceeds the robot's %u", req->num_points,
                         s->robots[req->robot].max_points);
    if (req->num_points == 0) return set_error(TSL_ERR_INVALID_ARGUMENT, "empty scan");
    s->pending.push_back(std::move(p));
    s->cv_req.notify_one();
    return TSL_OK;
}

int tsl_server_poll(tsl_server* s, tsl_result* out, size_t capacity, size_t* n, int wait_ms)
{
    if (!s || !n || (capacity && !out)) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty(); });
    size_t k = 0;
    while (k < capacity && !s->done.empty()) {
        out[k++] = s->done.front();
        s->done.pop_front();
    }
    *n = k;
    if (s->worker_error != TSL_OK) return s->worker_error;
    return TSL_OK;
}

int tsl_server_get_particles(tsl_server* s, uint32_t robot, tsl_particle* out, size_t capacity, size_t* n)
{
    if (!s || !n) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (robot >= s->robots.size() || !s->robots[robot].active)
        return set_error(TSL_ERR_INVALID_ARGUMENT, "no robot %u", robot);
    // held busy for the copy: no tick takes the robot and a removal waits
    Robot& r = s->robots[robot];
    s->cv_done.wait(lk, [&] { return !r.busy; });
    if (!r.active) return set_error(TSL_ERR_INVALID_ARGUMENT, "no robot %u", robot);
    r.busy = true;
    tsl_filter* f = r.f;
    lk.unlock();
    const int rc = tsl_filter_get_particles(f, out, capacity, n);
    lk.lock();
    r.busy = false;
    s->cv_done.notify_all();
    return rc;
}

int tsl_server_set_tile(tsl_server* s, int bulk, int k, int points_per_lane)
{
    if (!s) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL server");
    if (k != 1 && k != 2) return set_error(TSL_ERR_INVALID_ARGUMENT, "K must be 1 or 2");
    if (points_per_lane != 1 && points_per_lane != 2 && points_per_lane != 4)
        return set_error(TSL_ERR_INVALID_ARGUMENT, "points per lane must be 1, 2 or 4");
    std::lock_guard<std::mutex> lk(s->mu);
    const int c = bulk ? 1 : 0;
    const int k0 = s->K[c], p0 = s->P[c];
    s->K[c] = k;
    s->P[c] = points_per_lane;
    const int rc = class_occupancy(s, c);
    if (rc != TSL_OK) {
        s->K[c] = k0;
        s->P[c] = p0;
    }
    return rc;
}

int tsl_server_set_slice(tsl_server* s, uint64_t pairs, uint32_t reserved_per_cu)
{
    if (!s) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->slice_pairs = pairs;
    s->reserved_per_cu = reserved_per_cu;
    return TSL_OK;
}

int tsl_server_stats(tsl_server* s, uint64_t* used_bytes, uint64_t* ticks, uint64_t* requests)
{
    if (!s) return set_error(TSL_ERR_INVALID_ARGUMENT, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (used_bytes) *used_bytes = s->used;
    if (ticks) *ticks = s->ticks;
    if (requests) *requests = s->batched_requests;
    return TSL_OK;
}

}  // extern "C"
