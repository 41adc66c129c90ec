"""Benchmark of the MI355X MCL sensor update.

One step = one full update of a resident particle set against a resident
scan: restore the particle set from its device snapshot (so every step
evaluates the same particles), pose matrices, particle x point evaluation,
normalisation + pose sums, systematic resampling.  Inputs are resident in
HBM before the timed region.

Default workload: global localisation at scale -- sparse 1024x1024x128 TSDF
at 0.05 m (int16, ~40% of 1 m blocks allocated) of a seeded multi-room
building, 128x2048 ray-cast scan reduced to 20k points, 100k particles
uniform over free space.  Metric: particle*point evaluations per second.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

BYTES_PER_EVAL = 6  # 2 B int16 voxel + 4 B int32 block-table entry, both data-dependent gathers
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(w, sample_particles, repeats, threads):
    """The oracle (a C restatement of the reference CPU evaluator) on this
    host's cores, on the first `sample_particles` particles and the full scan."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_binding as ob
    om = ob.OracleMap.from_host_map(w.map)
    parts = w.particles[:sample_particles]
    evals = parts.shape[0] * w.points.shape[0]
    om.evaluate(parts[: max(1, sample_particles // 10)], w.points, w.tf, threads=threads)  # warm-up
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        om.evaluate(parts, w.points, w.tf, threads=threads)
        times.append(time.perf_counter() - t0)
    tuned = evals / statistics.median(times)
    # reference-faithful form: 8 OpenMP threads, per-particle copy of the scan
    # (the reference's num_particles_eval method: warm-up, then the median)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        om.evaluate(parts, w.points, w.tf, threads=8, faithful_copy=1)
        times.append(time.perf_counter() - t0)
    faithful = evals / statistics.median(times)
    return {"value": tuned, "unit": "particle*point evals/s", "cores": threads, "kind": "port",
            "sample": f"{parts.shape[0]} of {w.particles.shape[0]} particles x {w.points.shape[0]} points, "
                      f"median of {repeats} after warm-up (both forms); oracle (C/OpenMP restatement of "
                      f"TSDFEvaluator::evaluate), {threads} threads",
            "faithful_8_threads_value": faithful,
            "host_cpu": _cpu_model(), "host_cores": host_cores()}


def host_cores():
    """The host cores this process may use: its CPU affinity, the OpenMP
    thread setting of the environment, and the machine's logical CPUs."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    omp = int(omp) if omp and omp.isdigit() else None
    return {"affinity": aff, "omp_num_threads": omp, "logical_cpus": os.cpu_count(),
            "used": min(aff, omp) if omp else aff}


def _cpu_model():
    """Model name as lscpu reports it (or /proc/cpuinfo), with the CPU
    family, model and stepping numbers, which identify the part where a
    virtual machine shortens the name."""
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        fields = {}
        for line in out.splitlines():
            key, sep, val = line.partition(":")
            if sep and key.strip() not in fields:
                fields[key.strip()] = val.strip()
        if "Model name" in fields:
            ids = [f"{k.lower()} {fields[k]}" for k in ("CPU family", "Model", "Stepping") if fields.get(k)]
            return fields["Model name"] + (f" ({', '.join(ids)})" if ids else "")
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_eval_global_localisation_traffic.md")
VALU_FILE = os.path.join(ROOT, "profiles", "k_eval_global_localisation_valu.md")
# VALU issue peak: 256 CUs x 4 SIMDs x 16 lanes per cycle (a wave64 instruction
# takes 4 cycles) at the 2.4 GHz peak engine clock, in lane-operations/s
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9


L2_FILE = os.path.join(ROOT, "profiles", "k_eval_global_localisation_l2.md")

# the evaluation tile the default run uses (the fastest measured on this
# workload: 2 particles x 1 point per lane, at eight wavefronts per SIMD);
# the counter profiles under profiles/ are taken with it
PINNED_TILE = {"order": 1, "k": 2, "p": 1, "chunk": 0, "lookup": -1}
# the campus workload's (64-bit map form, whose instances with two particles
# per wavefront keep seven wavefronts per SIMD)
CAMPUS_TILE = {"order": 1, "k": 1, "p": 2, "chunk": 0, "lookup": -1}


def tile_key(t):
    return f"order={t['order']} k={t['k']} p={t['p']} chunk={t['chunk']} lookup={t['lookup']}"


def measured_traffic(tile):
    """HBM bytes per evaluation-kernel launch from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (scaled by the plain-copy calibration),
    or None if that summary is absent or was taken with another tile."""
    return _profile_value(TRAFFIC_FILE, "traffic_bytes_per_launch", tile)


def _profile_value(path, key, tile):
    """A value recorded in a profile summary, if the summary's tile line
    names the tile of this run."""
    found, same_tile = None, False
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("tile:"):
                    same_tile = line.split(":", 1)[1].strip() == tile_key(tile)
                if line.startswith(key + ":"):
                    found = float(line.split(":", 1)[1])
    except OSError:
        return None
    return found if same_tile else None


def valu_issue(kern_avg_ms, tile):
    """Issue-rate view of the evaluation kernel: VALU instructions per
    launch from the committed rocprofv3 SQ_INSTS_VALU pass (same workload and
    tile), x 64 lanes, over the measured launch time, against the VALU peak."""
    instr = _profile_value(VALU_FILE, "valu_instructions_per_launch", tile)
    if instr is None or kern_avg_ms <= 0:
        return None
    achieved = instr * 64 / (kern_avg_ms / 1000.0)
    return {"achieved": achieved, "peak": VALU_PEAK_LANE_OPS, "unit": "VALU lane-ops/s",
            "frac": achieved / VALU_PEAK_LANE_OPS, "instructions_per_launch": instr,
            "source": "profiles/k_eval_global_localisation_valu.md"}


def gather_probe(sizes_mib=(128, 16384), reps=5):
    """Random-line gather ceiling (tsl_probe_gather) for buffers the size of
    a map's code array (L2 + Infinity Cache) and of many GB (HBM)."""
    import ctypes
    from tsdf_localization_amd._lib import check, lib
    out = {}
    for mib in sizes_mib:
        lps, ms = ctypes.c_double(), ctypes.c_double()
        check(lib.tsl_probe_gather(ctypes.c_uint64(mib << 20), reps, ctypes.byref(lps), ctypes.byref(ms)))
        out[f"{mib}MiB"] = lps.value
        log(f"gather probe {mib} MiB: {lps.value:.3e} lines/s ({ms.value:.2f} ms per launch)")
    return out


def l2_view(kern_avg_ms, tile, probe):
    """Gather view of the evaluation kernel: L1->L2 read requests per launch
    and per vector-memory read instruction (rocprofv3 TCP_TCC_READ_REQ_sum,
    SQ_INSTS_VMEM_RD), and the 128-byte lines it misses in L2 (calibrated
    FETCH_SIZE / 128) per second against the random-line probe on a buffer
    of the map's size (lines that miss L2 and are served by the Infinity
    Cache), same workload and tile."""
    req = _profile_value(L2_FILE, "l2_read_requests_per_launch", tile)
    instr = _profile_value(L2_FILE, "vmem_read_instructions_per_launch", tile)
    lines = _profile_value(TRAFFIC_FILE, "fetch_lines_per_launch", tile)
    if req is None or lines is None or kern_avg_ms <= 0:
        return None
    sec = kern_avg_ms / 1000.0
    peak = probe.get("128MiB") if probe else None
    return {"l2_read_requests_per_launch": req, "l2_requests_per_s": req / sec,
            "l2_requests_per_vmem_instruction": req / instr if instr else None,
            "l2_miss_lines_per_launch": lines, "l2_miss_lines_per_s": lines / sec,
            "probe_lines_per_s": peak, "frac_of_probe": (lines / sec) / peak if peak else None,
            "probe_lines_per_s_hbm": probe.get("16384MiB") if probe else None,
            "source": "profiles/k_eval_global_localisation_l2.md, profiles/k_eval_global_localisation_traffic.md"}


def copy_calibration(mib, reps):
    """Plain device-to-device copies of known size (tsl_probe_copy), for
    scaling the FETCH_SIZE / WRITE_SIZE counters: each copy reads and writes
    mib MiB."""
    import ctypes
    from tsdf_localization_amd._lib import check, lib
    ms = ctypes.c_double()
    check(lib.tsl_probe_copy(ctypes.c_uint64(mib << 20), reps, ctypes.byref(ms)))
    print(json.dumps({"copy_calibration_bytes": mib << 20, "reps": reps, "ms_per_copy": ms.value}), flush=True)


def bench_batched(args):
    """64 independent filters on one shared map (particle counts log-uniform
    in 500-50k, scans of 2k-30k points): one step = every filter restored,
    updated and resampled.  Timed four ways: plain per-filter launches of
    device-resident filters on one stream (one filter after the other), the
    batched launch sequence with the tile the batch autotuner picks, the
    batched step replayed from a HIP graph, and the drop-in as the
    reference's node calls it, filter after filter on host arrays
    (Evaluator::evaluate: particles and scan copied in, weights copied back;
    then the systematic resampler on the host particle array)."""
    import tsdf_localization_amd as tsl
    from tsdf_localization_amd import synth

    t0 = time.time()
    w = synth.batched()
    n = len(w.robots)
    evals = w.meta["evals_per_step"]
    log(f"workload batched: {n} filters, {w.meta['particles_total']} particles, {evals:.3e} evals/step "
        f"(generated in {time.time() - t0:.1f}s)")
    ev = tsl.TSDFEvaluator(w.map)
    stream = tsl.Stream()
    rng = np.random.default_rng(777)
    draws = [np.array([np.float32(rng.uniform(0, 1.0 / r.particles.shape[0])) for r in w.robots], np.float32)
             for _ in range(args.warmup + args.steps)]

    def make_filters():
        fs = []
        for r in w.robots:
            f = tsl.DeviceFilter(ev, r.particles.shape[0])
            f.set_particles(r.particles)
            f.set_scan(r.points, w.tf)
            f.snapshot()
            fs.append(f)
        return fs

    def timed(step):
        for i in range(args.warmup):
            step(i)
        stream.synchronize()
        e0, e1 = tsl.Event(), tsl.Event()
        e0.record(stream)
        for i in range(args.steps):
            step(args.warmup + i)
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1) / args.steps

    plain = make_filters()

    def plain_step(i):
        for f, u0 in zip(plain, draws[i]):
            f.restore_async(stream)
            f.update_async(stream)
            f.resample_async(u0, stream)
    plain_ms = timed(plain_step)
    for f in plain:
        f.sync(stream)
    log(f"plain per-filter launches: {plain_ms:.3f} ms/step, {evals / plain_ms * 1e3:.3e} evals/s")

    filters = make_filters()
    b = tsl.Batch(ev, filters)
    k, pp, chunk, tune_ms = b.autotune(stream)
    log(f"batch autotune: K={k} points/lane={pp} chunk={chunk or 'auto'} ({tune_ms:.3f} ms per batched update)")
    eager_ms = timed(lambda i: b.step_async(draws[i], restore=True, stream=stream))
    b.sync(stream)
    log(f"batched eager: {eager_ms:.3f} ms/step, {evals / eager_ms * 1e3:.3e} evals/s")
    b.capture(stream, restore=True)
    graph_ms = timed(lambda i: b.launch(draws[i], stream))
    _, counts, status = b.sync(stream)
    log(f"batched graph K={k} points/lane={pp} chunk={chunk or 'auto'}: {graph_ms:.3f} ms/step, "
        f"{evals / graph_ms * 1e3:.3e} evals/s ({plain_ms / graph_ms:.2f}x plain); "
        f"member status {sorted(set(status.tolist()))}")

    # the drop-in on host arrays, filter after filter; each filter keeps its
    # own copy of its particles (evaluate overwrites only the weights, the
    # resampled set is dropped, as the device step restores its snapshot)
    host_parts = [r.particles.copy() for r in w.robots]

    def dropin_step(i):
        for r, parts, u0 in zip(w.robots, host_parts, draws[i]):
            ev.evaluate(parts, r.points, w.tf)
            tsl.evaluator.resample_systematic(parts, u0)
    for i in range(args.warmup):
        dropin_step(i)
    t_d = time.perf_counter()
    for i in range(args.steps):
        dropin_step(args.warmup + i)
    dropin_ms = (time.perf_counter() - t_d) * 1e3 / args.steps
    log(f"drop-in per filter on host arrays: {dropin_ms:.3f} ms/step, {evals / dropin_ms * 1e3:.3e} evals/s "
        f"(batched graph {dropin_ms / graph_ms:.2f}x)")
    out = {
        "metric": "particle*point evals/s",
        "value": evals / graph_ms * 1e3,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": graph_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded multi-room floor, ray-cast scans, particles around each robot)",
        "config": {"workload": "batched filters", "filters": n, "particles_total": w.meta["particles_total"],
                   "evals_per_step": evals, "map": "sparse 512x512x96 @0.05 m int16",
                   "step": "restore + pose + evaluate + normalise + resample, every filter",
                   "tile": {"particles_per_wavefront": k, "points_per_lane": pp, "chunk_points": chunk or "auto",
                            "picked_by": "tsl_batch_autotune"}},
        "plain_ms_per_step": plain_ms,
        "plain_value": evals / plain_ms * 1e3,
        "batched_eager_ms_per_step": eager_ms,
        "speedup_vs_plain": plain_ms / graph_ms,
        "plain_dropin_ms_per_step": dropin_ms,
        "plain_dropin_value": evals / dropin_ms * 1e3,
        "speedup_vs_plain_dropin": dropin_ms / graph_ms,
    }
    return out


def bench_serving(args):
    """256 robots resident on the building map, requests (odometry step +
    5k-20k point scan + draw) from a Poisson process of 200 requests/s:
    p50/p99 latency (submission to result) under that load, and the
    saturation throughput of the served form (every request submitted at
    once) against plain per-request launches of the same requests, two
    ways: device-resident filters (motion, update and resampling on the
    device), and the drop-in as the reference's node calls it on host arrays
    (Evaluator::evaluate, then the systematic resampler; the motion step,
    host code in that node, is left out of this leg)."""
    import tsdf_localization_amd as tsl
    from tsdf_localization_amd import synth
    from tsdf_localization_amd.server import Server

    t0 = time.time()
    w = synth.serving()
    R = len(w.robots)
    log(f"workload serving: {R} robots, mean {w.meta['mean_evals_per_request']:.3e} evals/request "
        f"(generated in {time.time() - t0:.1f}s)")
    ev = tsl.TSDFEvaluator(w.map)
    rng = np.random.default_rng(4242)
    rng_h = np.random.default_rng(4343)
    n_req = args.requests
    rate = args.rate
    robots_of = rng.integers(0, R, n_req)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    motions = [tsl.evaluator.motion_from_delta(0.05, 0.0, 0.0, 0.0, 0.0, 0.01, 0.01, 0.005, seed=i)
               for i in range(n_req)]
    u0 = [np.float32(rng.uniform(0, 0.9 / w.robots[r].particles.shape[0])) for r in robots_of]
    evals = [w.robots[r].particles.shape[0] * w.robots[r].points.shape[0] for r in robots_of]

    def fresh_server():
        srv = Server(ev, hbm_budget=args.hbm_budget_gb * (1 << 30), max_robots=R, max_particles=20_000,
                     max_batch=args.max_batch, chunk_len=args.chunk or 2048)
        if args.served_tile:
            srv.set_tile(0, args.served_tile[0], args.served_tile[1])
        ids = [srv.add_robot(r.particles, 20_000, w.tf) for r in w.robots]
        return srv, ids

    # 1. latency under Poisson arrivals
    srv, ids = fresh_server()
    log(f"server: {srv.stats()['used_bytes'] / 2**30:.2f} GiB resident for {R} robots")
    for i in range(min(16, n_req)):  # warm-up
        srv.submit(ids[robots_of[i]], motions[i], w.robots[robots_of[i]].points, u0[i], request_id=i)
    srv.wait(min(16, n_req))
    start = time.perf_counter()
    for i in range(n_req):
        delay = arrivals[i] - (time.perf_counter() - start)
        if delay > 0:
            time.sleep(delay)
        srv.submit(ids[robots_of[i]], motions[i], w.robots[robots_of[i]].points, u0[i], request_id=i)
    res = srv.wait(n_req, timeout_ms=600_000)
    wall = time.perf_counter() - start
    lat = np.array([r["done_ms"] - r["submit_ms"] for r in res])
    bad = sum(r["status"] != 0 for r in res)
    batch = np.array([r["batch"] for r in res])
    p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
    log(f"poisson {rate}/s: {n_req} requests in {wall:.2f} s, latency p50 {p50:.2f} ms p99 {p99:.2f} ms, "
        f"mean batch {batch.mean():.1f}, failed {bad}")

    # 1b. the same requests under heavier Poisson loads (latency headroom)
    by_rate = {str(rate): {"p50": p50, "p99": p99, "mean_batch": float(batch.mean())}}
    for hr in args.extra_rates:
        if hr <= 0:
            continue
        arr = np.cumsum(rng_h.exponential(1.0 / hr, n_req))
        t_h = time.perf_counter()
        for i in range(n_req):
            delay = arr[i] - (time.perf_counter() - t_h)
            if delay > 0:
                time.sleep(delay)
            srv.submit(ids[robots_of[i]], motions[i], w.robots[robots_of[i]].points, u0[i], request_id=i)
        rh = srv.wait(n_req, timeout_ms=600_000)
        lh = np.array([r["done_ms"] - r["submit_ms"] for r in rh])
        bh = np.array([r["batch"] for r in rh])
        by_rate[str(hr)] = {"p50": float(np.percentile(lh, 50)), "p99": float(np.percentile(lh, 99)),
                            "mean_batch": float(bh.mean())}
        log(f"poisson {hr}/s: latency p50 {by_rate[str(hr)]['p50']:.2f} ms p99 {by_rate[str(hr)]['p99']:.2f} ms, "
            f"mean batch {bh.mean():.1f}")
    srv.close()

    # 2. saturation: every request at once
    srv, ids = fresh_server()
    t1 = time.perf_counter()
    for i in range(n_req):
        srv.submit(ids[robots_of[i]], motions[i], w.robots[robots_of[i]].points, u0[i], request_id=i)
    res = srv.wait(n_req, timeout_ms=600_000)
    served_s = time.perf_counter() - t1
    served_rate = sum(evals) / served_s
    sat_batch = float(np.mean([r["batch"] for r in res]))
    log(f"served saturation: {served_s * 1e3:.1f} ms for {n_req} requests, {served_rate:.3e} evals/s, "
        f"mean batch {sat_batch:.1f}")
    srv.close()

    # 3. plain per-request launches, back to back
    filters = {}
    stream = tsl.Stream()
    for r in set(robots_of.tolist()):
        f = tsl.DeviceFilter(ev, w.robots[r].particles.shape[0])
        f.set_particles(w.robots[r].particles)
        filters[r] = f
    t2 = time.perf_counter()
    for i in range(n_req):
        f = filters[robots_of[i]]
        f.motion_async(motions[i], stream)
        f.set_scan(w.robots[robots_of[i]].points, w.tf)
        f.update_async(stream)
        f.resample_async(u0[i], stream)
        f.sync(stream)
    plain_s = time.perf_counter() - t2
    plain_rate = sum(evals) / plain_s
    log(f"plain per-request: {plain_s * 1e3:.1f} ms, {plain_rate:.3e} evals/s; served/plain {served_rate / plain_rate:.2f}x")

    # 4. the drop-in per request on host arrays; each robot's copy of its
    # particles is evaluated (only its weights change) and resampled into a
    # new array that is dropped, so every request evaluates the particle set
    # the served form starts from
    host_parts = {r: w.robots[r].particles.copy() for r in set(robots_of.tolist())}
    t3 = time.perf_counter()
    for i in range(n_req):
        r = robots_of[i]
        ev.evaluate(host_parts[r], w.robots[r].points, w.tf)
        tsl.evaluator.resample_systematic(host_parts[r], u0[i])
    dropin_s = time.perf_counter() - t3
    dropin_rate = sum(evals) / dropin_s
    log(f"drop-in per request on host arrays: {dropin_s * 1e3:.1f} ms, {dropin_rate:.3e} evals/s; "
        f"served/drop-in {served_rate / dropin_rate:.2f}x")
    out = {
        "metric": "particle*point evals/s",
        "value": served_rate,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": n_req,
        "warmup": 16,
        "ms_per_step": served_s * 1e3 / n_req,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (building map, shared ray-cast scans, particles around each robot)",
        "config": {"workload": "serving", "robots": R, "requests": n_req, "arrival_rate_per_s": rate,
                   "max_batch": args.max_batch, "chunk_points": args.chunk or 2048,
                   "map": "sparse 1024x1024x128 @0.05 m int16"},
        "latency_ms": {"p50": p50, "p99": p99, "mean_batch": float(batch.mean()), "failed": int(bad)},
        "latency_ms_by_rate": by_rate,
        "plain_value": plain_rate,
        "speedup_vs_plain": served_rate / plain_rate,
        "plain_dropin_value": dropin_rate,
        "speedup_vs_plain_dropin": served_rate / dropin_rate,
    }
    return out


def bench_mixed(args):
    """Four campus maps (12 x 11 building tiles, ~7.3e9 int16 voxels, ~14.6 GB
    each, block offsets past 2^31) served side by side on one GPU (one
    server per map), requests from a Poisson process split 9:1 between
    tracking robots (5,000 particles x 10,000 points) and global-localisation
    robots (1,000,000 particles x 4,000 points): p50/p99 latency per class
    under that load, the saturation throughput of the served form (every
    request submitted at once) and plain per-request launches of the same
    requests.  Every robot keeps its particle set between its requests, so
    a global robot's particles concentrate as its requests resample them and
    its later requests gather from fewer distinct map lines."""
    import tsdf_localization_amd as tsl
    from tsdf_localization_amd import synth
    from tsdf_localization_amd.server import Server

    t0 = time.time()
    w = synth.mixed(campus_tiles=None if args.mixed_small_maps else (12, 11))
    M = len(w.maps)
    T = len(w.tracking[0])
    log(f"workload mixed: {M} maps of {w.meta['map_bytes'][0] / 1e6:.0f} MB, {T} tracking + 1 global robot per map "
        f"(generated in {time.time() - t0:.1f}s)")
    evs = [tsl.TSDFEvaluator(m) for m in w.maps]
    rng = np.random.default_rng(5005)
    n_req = args.mixed_requests
    rate = args.mixed_rate
    is_global = rng.random(n_req) < 0.1
    map_of = rng.integers(0, M, n_req)
    trk_of = rng.integers(0, T, n_req)
    robot = [w.global_robots[map_of[i]] if is_global[i] else w.tracking[map_of[i]][trk_of[i]] for i in range(n_req)]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    motions = [tsl.evaluator.motion_from_delta(0.05, 0.0, 0.0, 0.0, 0.0, 0.01, 0.01, 0.005, seed=i)
               for i in range(n_req)]
    u0 = [np.float32(rng.uniform(0, 0.9 / robot[i].particles.shape[0])) for i in range(n_req)]
    evals = [robot[i].particles.shape[0] * robot[i].points.shape[0] for i in range(n_req)]
    gp = w.global_robots[0].particles.shape[0]

    def fresh_servers():
        srvs, ids = [], []
        for k in range(M):
            srv = Server(evs[k], hbm_budget=args.hbm_budget_gb * (1 << 30), max_robots=T + 1, max_particles=gp,
                         max_batch=8, chunk_len=2048)
            srv.set_slice(args.mixed_slice, args.mixed_reserve)
            trk = [srv.add_robot(r.particles, r.points.shape[0], w.tf) for r in w.tracking[k]]
            glb = srv.add_robot(w.global_robots[k].particles, w.global_robots[k].points.shape[0], w.tf)
            srvs.append(srv)
            ids.append((trk, glb))
        return srvs, ids

    def submit(srvs, ids, i):
        k = map_of[i]
        rid = ids[k][1] if is_global[i] else ids[k][0][trk_of[i]]
        srvs[k].submit(rid, motions[i], robot[i].points, u0[i], request_id=i)

    def collect(srvs):
        out = []
        for k in range(M):
            want = int(np.sum(map_of == k))
            out.extend(srvs[k].wait(want, timeout_ms=600_000))
        return out

    # 1. latency under Poisson arrivals
    srvs, ids = fresh_servers()
    used = sum(s.stats()["used_bytes"] for s in srvs)
    log(f"servers: {used / 2**30:.2f} GiB resident for {M * (T + 1)} robots")
    for k in range(M):  # warm-up: one request of each class per map
        srvs[k].submit(ids[k][0][0], motions[0], w.tracking[k][0].points, np.float32(0.1 / 5000),
                       request_id=n_req + 2 * k)
        srvs[k].submit(ids[k][1], motions[0], w.global_robots[k].points, np.float32(0.1 / gp),
                       request_id=n_req + 2 * k + 1)
        srvs[k].wait(2, timeout_ms=600_000)
    start = time.perf_counter()
    for i in range(n_req):
        delay = arrivals[i] - (time.perf_counter() - start)
        if delay > 0:
            time.sleep(delay)
        submit(srvs, ids, i)
    res = collect(srvs)
    wall = time.perf_counter() - start
    bad = sum(r["status"] != 0 for r in res)
    classes = {}
    for name, flag in (("tracking", False), ("global", True)):
        lat = np.array([r["done_ms"] - r["submit_ms"] for r in res if bool(is_global[r["id"]]) == flag])
        classes[name] = {"requests": int(lat.size), "p50": float(np.percentile(lat, 50)),
                         "p99": float(np.percentile(lat, 99))}
        log(f"poisson {rate}/s, {name}: {lat.size} requests, latency p50 {classes[name]['p50']:.2f} ms "
            f"p99 {classes[name]['p99']:.2f} ms")
    log(f"{n_req} requests in {wall:.2f} s, failed {bad}")
    for s in srvs:
        s.close()

    # 2. saturation: every request at once
    srvs, ids = fresh_servers()
    t1 = time.perf_counter()
    for i in range(n_req):
        submit(srvs, ids, i)
    collect(srvs)
    served_s = time.perf_counter() - t1
    served_rate = sum(evals) / served_s
    log(f"served saturation: {served_s * 1e3:.1f} ms for {n_req} requests, {served_rate:.3e} evals/s")
    for s in srvs:
        s.close()

    # 3. plain per-request launches, back to back
    stream = tsl.Stream()
    filters = {}
    for i in range(n_req):
        key = (int(map_of[i]), -1 if is_global[i] else int(trk_of[i]))
        if key not in filters:
            f = tsl.DeviceFilter(evs[key[0]], robot[i].particles.shape[0])
            f.set_particles(robot[i].particles)
            filters[key] = f
    t2 = time.perf_counter()
    for i in range(n_req):
        f = filters[(int(map_of[i]), -1 if is_global[i] else int(trk_of[i]))]
        f.motion_async(motions[i], stream)
        f.set_scan(robot[i].points, w.tf)
        f.update_async(stream)
        f.resample_async(u0[i], stream)
        f.sync(stream)
    plain_s = time.perf_counter() - t2
    plain_rate = sum(evals) / plain_s
    log(f"plain per-request: {plain_s * 1e3:.1f} ms, {plain_rate:.3e} evals/s; served/plain {served_rate / plain_rate:.2f}x")
    out = {
        "metric": "particle*point evals/s",
        "value": served_rate,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": n_req,
        "warmup": 2 * M,
        "ms_per_step": served_s * 1e3 / n_req,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (four seeded campus maps of building tiles, ray-cast scans, particles around each "
                "tracking robot and over the free space for each global robot)",
        "config": {"workload": "mixed", "maps": M, "map_bytes_each": w.meta["map_bytes"][0],
                   "map": ("sparse 1024x1024x128 @0.05 m int16 each" if args.mixed_small_maps else
                           "campus of 12x11 building tiles, 624x572x6.4 m @0.05 m int16 each (64-bit block offsets)"),
                   "classes": {"tracking": "5000 particles x 10000 points", "global": "1000000 x 4000"},
                   "class_ratio": "9:1", "requests": n_req, "arrival_rate_per_s": rate,
                   "global_slice_pairs": args.mixed_slice, "global_reserved_slots_per_cu": args.mixed_reserve},
        "latency_ms_by_class": classes,
        "failed": int(bad),
        "plain_value": plain_rate,
        "speedup_vs_plain": served_rate / plain_rate,
    }
    print(json.dumps(out), flush=True)


def bench_campus(args):
    """The global-localisation request of the mixed workload on its own: one
    campus map (12 x 11 building tiles, 7.3e9 int16 voxels, 14.6 GB, 64-bit
    block offsets: the map streams from HBM), 1,000,000 particles over its
    free space, a 4,000-point scan.  One step = restore + pose + evaluation
    + normalisation + resampling of the plain filter (k_eval); then the same
    robot served (k_eval_batch in slices), one request after the other."""
    import tsdf_localization_amd as tsl
    from tsdf_localization_amd import synth
    from tsdf_localization_amd.server import Server

    t0 = time.time()
    cp = synth.campus(0x5EED0005, tiles=(12, 11))
    rng = np.random.default_rng(0x5EED0005)
    tf = synth.tf_from_xyzrpy(**synth.SCANNER_TF)
    _, pts = synth.campus_scan(cp, rng, tf, 4000)
    parts = synth.campus_particles(cp, 0x5EED0005 + 5, args.particles or 1_000_000, np.radians(3.0))
    P, N = parts.shape[0], pts.shape[0]
    c = cp.map.coef()
    log(f"workload campus: {c.data_size} voxels ({2 * c.data_size / 1e9:.2f} GB), P={P} N={N} "
        f"(generated in {time.time() - t0:.1f}s)")
    ev = tsl.TSDFEvaluator(cp.map)
    f = tsl.DeviceFilter(ev, P)
    f.set_particles(parts)
    f.set_scan(pts, tf)
    f.snapshot()
    f.set_tile(args.tile_k, args.chunk, args.tile_p)
    f.set_order(args.order)
    stream = tsl.Stream()
    u0s = [np.float32(rng.uniform(0.0, 1.0 / P)) for _ in range(args.warmup + args.steps)]

    def step(i):
        f.restore_async(stream)
        f.update_async(stream)
        f.resample_async(u0s[i], stream)

    for i in range(args.warmup):
        step(i)
    f.sync(stream)
    f.set_timing(True)
    e0, e1 = tsl.Event(), tsl.Event()
    e0.record(stream)
    for i in range(args.steps):
        step(args.warmup + i)
    e1.record(stream)
    stream.synchronize()
    total_ms = e0.elapsed_ms(e1)
    kern_ms, launches, pairs = f.eval_timing()
    f.sync(stream)
    f.set_timing(False)
    ms_per_step = total_ms / args.steps
    kern_avg = kern_ms / max(launches, 1)
    evals = P * N
    achieved = BYTES_PER_EVAL * (pairs / max(launches, 1)) / (kern_avg / 1000.0) / 1e9
    log(f"plain: {ms_per_step:.2f} ms/step, {evals / ms_per_step * 1e3:.3e} evals/s; k_eval {kern_avg:.2f} ms "
        f"({achieved:.0f} GB/s algorithmic)")
    # served: the same request, one after the other; the robot is added
    # afresh before each one so that every request sees the uniform particle
    # set the plain steps restore (a resampled set collapses onto the best
    # poses, whose lookups share cache lines)
    # (a robot of 65,536+ particle slots is in the server's bulk class, whose
    # items are whole scans whatever the server's chunk length)
    srv = Server(ev, max_robots=2, max_particles=P, max_batch=4)
    srv.set_slice(args.mixed_slice, args.mixed_reserve)
    if args.served_tile:
        srv.set_tile(1, args.served_tile[0], args.served_tile[1])
    motion = tsl.evaluator.motion_from_delta(0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, seed=1)
    lat = []
    for i in range(args.warmup + args.steps):
        rid = srv.add_robot(parts, N, tf)
        srv.submit(rid, motion, pts, u0s[i], request_id=i)
        r = srv.wait(1)[0]
        if r["status"] != 0:
            raise RuntimeError(f"served campus request {i} failed: status {r['status']}")
        srv.remove_robot(rid)
        if i >= args.warmup:
            lat.append(r["done_ms"] - r["submit_ms"])
    srv.close()
    served_ms = float(np.median(lat))
    log(f"served: {served_ms:.2f} ms per request (median of {len(lat)}), {evals / served_ms * 1e3:.3e} evals/s")
    traffic = _profile_value(CAMPUS_TRAFFIC_FILE, "traffic_bytes_per_launch",
                             {"order": args.order, "k": args.tile_k, "p": args.tile_p, "chunk": args.chunk,
                              "lookup": -1})
    out = {
        "metric": "particle*point evals/s", "value": evals / ms_per_step * 1e3, "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (campus of seeded building tiles, ray-cast scan, particles over the free space)",
        "config": {"workload": "global localisation on a 14.6 GB campus map (the mixed workload's global class)",
                   "particles": P, "points": N, "map_voxels": int(c.data_size),
                   "map": "campus of 12x11 building tiles, 624x572x6.4 m @0.05 m int16 (64-bit block offsets)",
                   "tile": {"particle_order": {0: "caller", 1: "spatial", -1: "auto"}[args.order],
                            "particles_per_wavefront": args.tile_k, "points_per_lane": args.tile_p,
                            "chunk_points": args.chunk or "auto"}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_eval (64-bit map form)",
                     "kernel_avg_ms": kern_avg, "bytes_per_eval": BYTES_PER_EVAL,
                     "evals_per_launch": pairs // max(launches, 1)},
        "served": {"ms_per_request": served_ms, "evals_per_s": evals / served_ms * 1e3,
                   "slice_pairs": args.mixed_slice, "reserved_slots_per_cu": args.mixed_reserve,
                   "chunk_points": "whole scan (bulk class)",
                   "tile": (None if not args.served_tile else
                            {"particles_per_wavefront": args.served_tile[0],
                             "points_per_lane": args.served_tile[1]})},
    }
    print(json.dumps(out), flush=True)


CAMPUS_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_eval_campus_traffic.md")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="building",
                    choices=["building", "room", "batched", "serving", "mixed", "campus"])
    ap.add_argument("--mixed-requests", type=int, default=2000,
                    help="mixed: number of requests (2000: p99 of each class over more than 100 samples)")
    ap.add_argument("--mixed-rate", type=float, default=50.0, help="mixed: Poisson arrival rate (requests/s)")
    ap.add_argument("--mixed-slice", type=int, default=1 << 27,
                    help="mixed: particle-point pairs per launch of a global robot's evaluation (0 = one launch)")
    ap.add_argument("--mixed-reserve", type=int, default=2,
                    help="mixed: workgroup slots per CU a sliced global evaluation leaves to tracking requests")
    ap.add_argument("--served-tile", type=int, nargs=2, default=None, metavar=("K", "P"),
                    help="campus, serving: tile of the served evaluation (campus: the bulk class; serving: the "
                         "tracking class), particles per wavefront and points per lane (default: the server's)")
    ap.add_argument("--mixed-small-maps", action="store_true",
                    help="mixed: one 51.2 m building per map instead of a 12 x 11-tile campus")
    ap.add_argument("--requests", type=int, default=1000, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=200.0, help="serving: Poisson arrival rate (requests/s)")
    ap.add_argument("--extra-rates", type=float, nargs="*", default=[1000.0, 2000.0],
                    help="serving: further Poisson rates to report latency at")
    ap.add_argument("--max-batch", type=int, default=64, help="serving: requests per tick")
    ap.add_argument("--hbm-budget-gb", type=int, default=64, help="serving: device bytes the server may hold")
    ap.add_argument("--particles", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-dropin", action="store_true", help="skip the host-array drop-in timing")
    ap.add_argument("--graph-step", action="store_true", help="also time the step replayed from a HIP graph")
    ap.add_argument("--cpu-sample", type=int, default=10000)
    ap.add_argument("--cpu-repeats", type=int, default=10)
    ap.add_argument("--autotune", action="store_true",
                    help="pick the tile with tsl_filter_autotune (the room's default when no tile flag is given)")
    ap.add_argument("--no-autotune", action="store_true", help="room: the pinned tile instead of the autotuner's")
    ap.add_argument("--tile-k", type=int, default=None, help="particles per wavefront (1, 2)")
    ap.add_argument("--tile-p", type=int, default=None, help="points per lane per iteration (1, 2, 4)")
    ap.add_argument("--order", type=int, default=None,
                    help="particle order: 0 caller, 1 spatial, -1 auto (default: the pinned tile's on the "
                         "building and campus workloads, auto on the room)")
    ap.add_argument("--lookup", type=int, default=None, help="block lookup: 0 table, 1 LDS rank, -1 auto")
    ap.add_argument("--chunk", type=int, default=None, help="scan points per wavefront (0 = automatic)")
    ap.add_argument("--no-headline", action="store_true", help="skip the batched and serving forms in the default run")
    ap.add_argument("--headline-steps", type=int, default=10, help="batched steps timed in the default run")
    ap.add_argument("--no-probe", action="store_true", help="skip the gather probe")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run plain copies of this many MiB (counter calibration) and exit")
    args = ap.parse_args()
    # the room's update is a chain of small latency-bound kernels whose best
    # tile differs from the global-localisation workload's pinned one: the
    # library's autotuner picks it unless a tile flag is given
    tile_flags = (args.tile_k, args.tile_p, args.lookup, args.chunk, args.order)
    if args.workload == "room" and not args.no_autotune and all(v is None for v in tile_flags):
        args.autotune = True
    default_tile = CAMPUS_TILE if args.workload == "campus" else PINNED_TILE
    for name, key in (("tile_k", "k"), ("tile_p", "p"), ("lookup", "lookup"), ("chunk", "chunk")):
        if getattr(args, name) is None:
            setattr(args, name, default_tile[key])
    if args.order is None:
        # the room's 1,000 particles gain nothing from the spatial order's
        # sort kernels; the automatic choice takes the caller's order there
        args.order = -1 if args.workload == "room" else default_tile["order"]

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "filters are independent, so N GPUs would each run their own share.", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib, max(args.steps, 1))
        return

    if args.workload == "batched":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.workload == "serving":
        print(json.dumps(bench_serving(args)), flush=True)
        return
    if args.workload == "mixed":
        bench_mixed(args)
        return
    if args.workload == "campus":
        bench_campus(args)
        return

    import tsdf_localization_amd as tsl
    from tsdf_localization_amd import synth

    t0 = time.time()
    if args.workload == "building":
        w = synth.building(num_particles=args.particles or 100_000)
    else:
        w = synth.room(num_particles=args.particles or 1000)
    P, N = w.particles.shape[0], w.points.shape[0]
    log(f"workload {w.name}: P={P} N={N} map voxels={w.map.coef().data_size} meta={w.meta} "
        f"(generated in {time.time() - t0:.1f}s)")

    ev = tsl.TSDFEvaluator(w.map)
    f = tsl.DeviceFilter(ev, P)
    f.set_particles(w.particles)
    f.set_scan(w.points, w.tf)
    f.snapshot()
    stream = tsl.Stream()
    lookup = args.lookup
    if args.autotune:
        order, k, pp, chunk, tms = f.autotune(stream, reps=3)
        lookup = None  # the autotuner sets the block lookup too
        picked_by = "tsl_filter_autotune"
        log(f"autotuned: order={order} K={k} points/lane={pp} chunk={chunk or 'auto'} "
            f"({tms:.3f} ms per update for evaluation and particle order)")
    else:
        order, k, pp, chunk = args.order, args.tile_k, args.tile_p, args.chunk
        picked_by = "pinned" if (order, k, pp, chunk, lookup) == tuple(PINNED_TILE.values()) else "command line"
        f.set_tile(k, chunk, pp)
        f.set_order(order)
        f.set_block_lookup(lookup)
    tile = {"order": order, "k": k, "p": pp, "chunk": chunk or 0, "lookup": lookup}
    rng = np.random.default_rng(12345)
    u0s = [np.float32(rng.uniform(0.0, 1.0 / P)) for _ in range(args.warmup + args.steps)]

    def step(i):
        f.restore_async(stream)
        f.update_async(stream)
        f.resample_async(u0s[i], stream)

    for i in range(args.warmup):
        step(i)
    pose, n_after = f.sync(stream)
    log(f"warm-up done: pose {pose['position']}, resampled count {n_after}")

    f.set_timing(True)
    e0, e1 = tsl.Event(), tsl.Event()
    stream.synchronize()
    e0.record(stream)
    for i in range(args.steps):
        step(args.warmup + i)
    e1.record(stream)
    stream.synchronize()
    total_ms = e0.elapsed_ms(e1)
    kern_ms, launches, pairs = f.eval_timing()
    pose, n_after = f.sync(stream)  # raises on a failed update
    f.set_timing(False)

    ms_per_step = total_ms / args.steps
    evals = P * N
    value = evals * args.steps / (total_ms / 1000.0)
    kern_avg_ms = kern_ms / max(launches, 1)
    achieved = BYTES_PER_EVAL * (pairs / max(launches, 1)) / (kern_avg_ms / 1000.0) / 1e9
    log(f"{args.steps} steps in {total_ms:.2f} ms: {ms_per_step:.3f} ms/step, {value:.3e} evals/s; "
        f"eval kernel {kern_avg_ms:.3f} ms avg over {launches} launches ({achieved:.0f} GB/s algorithmic)")

    # the same step replayed from a HIP graph (a batch of this one filter):
    # reported beside the plain launches, which stay the measured value
    graph_ms = None
    if args.graph_step:
        b = tsl.Batch(ev, [f])
        b.set_tile(k if k in (1, 2) else 1, chunk, pp if pp in (1, 2, 4) else 1)
        one = [np.array([u], np.float32) for u in u0s]
        b.capture(stream, restore=True)
        for i in range(args.warmup):
            b.launch(one[i], stream)
        stream.synchronize()
        g0, g1 = tsl.Event(), tsl.Event()
        g0.record(stream)
        for i in range(args.steps):
            b.launch(one[args.warmup + i], stream)
        g1.record(stream)
        stream.synchronize()
        b.sync(stream)
        graph_ms = g0.elapsed_ms(g1) / args.steps
        log(f"graph-replayed step: {graph_ms:.3f} ms/step, {evals / graph_ms * 1e3:.3e} evals/s")

    # the drop-in form: host particle and point arrays through tsl_evaluate,
    # which copies them in and the weights out on every call like the
    # reference's CudaEvaluator (PCIe-inclusive; reported beside the value)
    dropin = None
    if not args.no_dropin:
        # median of 10 calls after a warm-up: pageable host copies vary
        # from call to call with the host's memory traffic
        reps = 10
        copies = [w.particles.copy() for _ in range(reps + 1)]
        ev.evaluate(copies[0], w.points, w.tf)
        times = []
        for i in range(reps):
            t0 = time.perf_counter()
            ev.evaluate(copies[i + 1], w.points, w.tf)
            times.append((time.perf_counter() - t0) * 1000.0)
        dropin_ms = float(np.median(times))
        dropin = {"ms_per_update": dropin_ms, "evals_per_s": evals / dropin_ms * 1e3,
                  "ms_min_max": [float(min(times)), float(max(times))], "calls": reps,
                  "form": "tsl_evaluate on host arrays: copy in, pose + evaluate + normalise, copy weights out"}
        log(f"drop-in evaluate on host arrays: {dropin_ms:.3f} ms/update (median of {reps}; "
            f"{min(times):.3f}-{max(times):.3f}), {evals / dropin_ms * 1e3:.3e} evals/s")

    cpu = None
    if not args.no_cpu:
        threads = host_cores()["used"]
        cpu = cpu_baseline(w, min(args.cpu_sample, P), args.cpu_repeats, threads)
        log(f"cpu baseline: {cpu['value']:.3e} evals/s on {threads} threads; faithful 8-thread "
            f"{cpu['faithful_8_threads_value']:.3e}")

    # the random-line gather probe sets the L2 view of the global-localisation
    # kernel (the room's update is latency bound; no line rate applies)
    probe = gather_probe() if not args.no_probe and w.name == "building" else None

    # the headline metric's other forms beside the plain update: batched
    # filters and the serving loop, each against plain per-filter /
    # per-request launches
    headline = {}
    if not args.no_headline and w.name == "building":
        hb = argparse.Namespace(**vars(args))
        hb.steps, hb.warmup = args.headline_steps, 2
        b = bench_batched(hb)
        headline["batched"] = {k2: b[k2] for k2 in ("value", "ms_per_step", "plain_value", "plain_ms_per_step",
                                                    "speedup_vs_plain", "batched_eager_ms_per_step",
                                                    "plain_dropin_value", "plain_dropin_ms_per_step",
                                                    "speedup_vs_plain_dropin")}
        headline["batched"]["config"] = b["config"]
        s = bench_serving(argparse.Namespace(**dict(vars(args), extra_rates=[1000.0])))
        headline["serving"] = {k2: s[k2] for k2 in ("value", "plain_value", "speedup_vs_plain", "latency_ms",
                                                    "latency_ms_by_rate", "plain_dropin_value",
                                                    "speedup_vs_plain_dropin")}
        headline["serving"]["config"] = s["config"]

    out = {
        "metric": "particle*point evals/s",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene, ray-cast scan, sampled particles)",
        "config": {"workload": "global localisation at scale" if w.name == "building" else "plain MCL update",
                   "form": "plain (one filter, per-update launches)",
                   "particles": P, "points": N, "map": "sparse 1024x1024x128 @0.05 m int16" if w.name == "building"
                   else "dense 256x256x64 @0.05 m int16",
                   "map_voxels": int(w.map.coef().data_size), "block_fraction": w.meta.get("block_fraction"),
                   "step": "restore + pose + evaluate + normalise + resample",
                   "tile": {"particle_order": {0: "caller", 1: "spatial", -1: "auto"}[order],
                            "particles_per_wavefront": k, "points_per_lane": pp, "chunk_points": chunk or "auto",
                            "block_lookup": {0: "table", 1: "LDS rank", -1: "auto", None: "autotuned"}[lookup],
                            "picked_by": picked_by, "profile_key": tile_key(tile)}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(tile) if w.name == "building" else None,
                     "kernel": "k_eval", "kernel_avg_ms": kern_avg_ms,
                     "bytes_per_eval": BYTES_PER_EVAL, "evals_per_launch": pairs // max(launches, 1)},
        "valu": valu_issue(kern_avg_ms, tile) if w.name == "building" else None,
        "l2": l2_view(kern_avg_ms, tile, probe) if w.name == "building" else None,
        "gather_probe_lines_per_s": probe,
        "cpu_baseline": cpu,
        "graph_step_ms": graph_ms,
        "dropin_host_buffers": dropin,
        "resampled_count": n_after,
    }
    # the limiter the counters show: the larger of the HBM, VALU-issue and
    # L2-request fractions (the roofline object keeps the HBM form)
    fracs = {"hbm": out["roofline"]["frac"]}
    if out["valu"]:
        fracs["valu_issue"] = out["valu"]["frac"]
    if out["l2"] and out["l2"]["frac_of_probe"]:
        fracs["l2_gather"] = out["l2"]["frac_of_probe"]
    out["roofline"]["measured_limiter"] = max(fracs, key=fracs.get)
    out["roofline"]["fractions"] = fracs
    out.update(headline)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
