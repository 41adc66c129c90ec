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
    t0 = time.perf_counter()
    om.evaluate(parts, w.points, w.tf, threads=8, faithful_copy=1)
    faithful = evals / (time.perf_counter() - t0)
    return {"value": tuned, "unit": "particle*point evals/s", "cores": threads, "kind": "port",
            "sample": f"{parts.shape[0]} of {w.particles.shape[0]} particles x {w.points.shape[0]} points, "
                      f"median of {repeats} after warm-up; oracle (C/OpenMP restatement of "
                      f"TSDFEvaluator::evaluate), {threads} threads",
            "faithful_8_threads_value": faithful,
            "host_cpu": _cpu_model()}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_eval_global_localisation_traffic.md")


def measured_traffic():
    """HBM bytes per evaluation-kernel launch from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (scaled by the plain-copy calibration),
    or None if that summary is absent."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def copy_calibration(mib, reps):
    """Plain device-to-device copies of known size, for scaling the
    FETCH_SIZE / WRITE_SIZE counters: each copy reads and writes mib MiB."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="building", choices=["building", "room"])
    ap.add_argument("--particles", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-sample", type=int, default=20000)
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--tile-k", type=int, default=None, help="particles per wavefront (1, 2, 4); default autotune")
    ap.add_argument("--chunk", type=int, default=0, help="scan points per wavefront (0 = automatic)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run plain copies of this many MiB (counter calibration) and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "filters are independent, so N GPUs would each run their own share.", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib, max(args.steps, 1))
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
    if args.tile_k is None:
        k, chunk, tms = f.autotune(stream, reps=3)
        log(f"autotuned tile: K={k} chunk={chunk or 'auto'} ({tms:.3f} ms per evaluation kernel)")
    else:
        k, chunk = args.tile_k, args.chunk
        f.set_tile(k, chunk)
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

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(w, min(args.cpu_sample, P), args.cpu_repeats, threads)
        log(f"cpu baseline: {cpu['value']:.3e} evals/s on {threads} threads; faithful 8-thread "
            f"{cpu['faithful_8_threads_value']:.3e}")

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
                   "tile": {"particles_per_wavefront": k, "chunk_points": chunk or "auto"}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "k_eval", "kernel_avg_ms": kern_avg_ms,
                     "bytes_per_eval": BYTES_PER_EVAL, "evals_per_launch": pairs // max(launches, 1)},
        "cpu_baseline": cpu,
        "resampled_count": n_after,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
