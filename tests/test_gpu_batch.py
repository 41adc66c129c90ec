"""Batched update: many device-resident filters on one map in one launch
sequence.  Each member must give exactly the bits it gives alone (same
chunk length), eager and replayed from a captured HIP graph, and match the
oracle within the weight tolerance."""
import numpy as np
import pytest

import oracle_binding as ob

pytestmark = pytest.mark.gpu

WEIGHT_RTOL = 1e-5
CHUNK = 320


@pytest.fixture(scope="module")
def tsl():
    import tsdf_localization_amd as t
    assert t.device_count() >= 1, "no HIP device"
    return t


@pytest.fixture(scope="module")
def room():
    from tsdf_localization_amd import synth
    return synth.room()


def _members(room, sizes, seed=5):
    """(particles, points) per member: uneven particle counts, different
    scan subsets."""
    rng = np.random.default_rng(seed)
    out = []
    for i, (P, N) in enumerate(sizes):
        pidx = rng.choice(room.particles.shape[0], P, replace=P > room.particles.shape[0])
        parts = room.particles[pidx].copy()
        parts[:, 6] = 1.0 / P
        pts = room.points[rng.permutation(room.points.shape[0])[:N]]
        out.append((parts, pts))
    return out


def _alone(tsl, ev, parts, pts, tf, u0, k, pp):
    f = tsl.DeviceFilter(ev, parts.shape[0])
    f.set_particles(parts)
    f.set_scan(pts, tf)
    f.set_tile(k, CHUNK, pp)
    f.set_order(0)
    s = tsl.Stream()
    f.update_async(s)
    pose, _ = f.sync(s)
    weighted = f.get_particles()
    f.resample_async(u0, s)
    f.sync(s)
    return weighted, f.get_particles(), pose


# includes the edges of the lone filter's one-workgroup finishing kernels
# (2048 particles: reduction in one workgroup, resampling not; 2049:
# neither; 2040: both), which must match the member kernels
SIZES = [(37, 300), (1000, 1500), (257, 64), (2500, 900), (64, 1), (2048, 700), (2049, 500), (2040, 200),
         (8000, 200)]


@pytest.mark.parametrize("k,pp", [(1, 1), (2, 1), (1, 4)])
def test_batch_member_equals_alone(tsl, room, k, pp):
    ev = tsl.TSDFEvaluator(room.map)
    mem = _members(room, SIZES)
    u0 = np.array([np.float32(0.37 / p.shape[0]) for p, _ in mem], np.float32)
    filters = []
    for parts, pts in mem:
        f = tsl.DeviceFilter(ev, parts.shape[0])
        f.set_particles(parts)
        f.set_scan(pts, room.tf)
        filters.append(f)
    b = tsl.Batch(ev, filters)
    b.set_tile(k, CHUNK, pp)
    s = tsl.Stream()
    b.update_async(s)
    poses_u, counts_u, status_u = b.sync(s)
    weighted = [f.get_particles() for f in filters]
    b.resample_async(u0, s)
    poses, counts, status = b.sync(s)
    assert np.all(status == 0)
    om = ob.OracleMap.from_host_map(room.map)
    for i, (parts, pts) in enumerate(mem):
        w_alone, r_alone, pose_alone = _alone(tsl, ev, parts, pts, room.tf, u0[i], k, pp)
        assert np.array_equal(weighted[i].view(np.uint32), w_alone.view(np.uint32)), f"member {i} weights"
        got = filters[i].get_particles()
        assert np.array_equal(got.view(np.uint32), r_alone.view(np.uint32)), f"member {i} resampled"
        assert counts[i] == got.shape[0]
        assert poses_u[i]["position"] == pose_alone["position"]
        _, w64 = om.evaluate(parts, pts, room.tf)
        ref = w64 / w64.sum()
        assert np.max(np.abs(weighted[i][:, 6].astype(np.float64) - ref) / ref) <= WEIGHT_RTOL
        idx = ob.resample(weighted[i][:, 6], u0[i])
        assert np.array_equal(got, weighted[i][idx])


def test_batch_autotune_keeps_results(tsl, room):
    """The batch tile autotuner returns one of its candidates, leaves the
    particles untouched (weights as one update gives them), and the update
    with the tile it picked equals each member updated alone with that tile,
    bit for bit."""
    ev = tsl.TSDFEvaluator(room.map)
    mem = _members(room, SIZES, seed=13)
    filters = []
    for parts, pts in mem:
        f = tsl.DeviceFilter(ev, parts.shape[0])
        f.set_particles(parts)
        f.set_scan(pts, room.tf)
        filters.append(f)
    b = tsl.Batch(ev, filters)
    s = tsl.Stream()
    k, pp, chunk, ms = b.autotune(s, reps=1)
    assert k in (1, 2) and pp in (1, 2, 4) and chunk in (0, 1024, 4096, 8192) and ms > 0.0
    for f, (parts, _) in zip(filters, mem):
        assert np.array_equal(f.get_particles()[:, :6], parts[:, :6])
    b.update_async(s)
    _, _, status = b.sync(s)
    assert np.all(status == 0)
    for i, (parts, pts) in enumerate(mem):
        alone = tsl.DeviceFilter(ev, parts.shape[0])
        alone.set_particles(parts)
        alone.set_scan(pts, room.tf)
        # the batch's automatic chunk: 2048 points, or the whole scan if shorter
        alone.set_tile(k, chunk or min(2048, (pts.shape[0] + 63) // 64 * 64), pp)
        alone.set_order(0)
        alone.update_async(s)
        alone.sync(s)
        assert np.array_equal(filters[i].get_particles().view(np.uint32), alone.get_particles().view(np.uint32)), i


def test_batch_graph_replays_match_eager(tsl, room):
    ev = tsl.TSDFEvaluator(room.map)
    mem = _members(room, SIZES, seed=9)
    filters = []
    for parts, pts in mem:
        f = tsl.DeviceFilter(ev, parts.shape[0])
        f.set_particles(parts)
        f.set_scan(pts, room.tf)
        f.snapshot()
        filters.append(f)
    b = tsl.Batch(ev, filters)
    b.set_tile(1, CHUNK, 1)
    s = tsl.Stream()
    rng = np.random.default_rng(3)
    draws = [np.array([np.float32(rng.uniform(0, 1.0 / p.shape[0])) for p, _ in mem], np.float32) for _ in range(3)]
    eager = []
    for u0 in draws:
        b.step_async(u0, restore=True, stream=s)
        b.sync(s)
        eager.append([f.get_particles() for f in filters])
    b.capture(s, restore=True)
    for rep in range(2):
        for u0, ref in zip(draws, eager):
            b.launch(u0, s)
            _, counts, status = b.sync(s)
            assert np.all(status == 0)
            for i, f in enumerate(filters):
                got = f.get_particles()
                assert np.array_equal(got.view(np.uint32), ref[i].view(np.uint32)), (rep, i)
            # eager work between replays
            b.update_async(s)
            b.sync(s)


@pytest.mark.parametrize("workload,trig", [("room", 0), ("room", 2), ("many", 0)])
def test_filter_motion_matches_oracle(tsl, room, workload, trig):
    """Device apply_model vs the oracle with the same counter-based draws.
    Under the glibc binding (the default) the float functions (sinf/cosf of
    the particle angles; asinf, atan2f and cosf in getAngleFromMat) are
    glibc's in the oracle and their bit-exact restatements on the device; the
    double sin/cos of the draws and the log/sqrt/cos that make them come from
    each side's math library but are rounded to float before they reach the
    pose, so a last-bit difference there almost never survives: every
    particle is bit for bit on these inputs (the inputs are fixed, so the
    test is deterministic), and the 4e-6 m / 2e-6 rad bound is the fallback
    statement for inputs where one would.  The double binding uses each
    side's double asin/atan2/cos in getAngleFromMat, rounded to float: also
    bit for bit here.  Workloads: the room's 1,000 particles, and 100,000
    seeded particles over a 50 m square with any heading."""
    ev = tsl.TSDFEvaluator(room.map, trig=trig)
    if workload == "room":
        parts = room.particles.copy()
    else:
        rng = np.random.default_rng(77)
        P = 100_000
        parts = np.zeros((P, 7), np.float32)
        parts[:, :2] = rng.uniform(0.0, 50.0, (P, 2))
        parts[:, 2] = rng.uniform(0.0, 2.0, P)
        parts[:, 3:5] = rng.normal(0.0, 0.05, (P, 2))
        parts[:, 5] = rng.uniform(-np.pi, np.pi, P)
        parts[:, 6] = 1.0 / P
    f = tsl.DeviceFilter(ev, parts.shape[0])
    f.set_particles(parts)
    m = tsl.evaluator.motion_from_delta(0.4, 0.05, 0.01, 0.01, -0.02, 0.15, 0.02, 0.01, seed=1234)
    s = tsl.Stream()
    f.motion_async(m, s)
    s.synchronize()
    got = f.get_particles()
    ref = ob.motion(parts, list(m.mean), list(m.stddev), m.seed, trig_mode=ob.trig_mode_for(trig))
    assert np.max(np.abs(got[:, :3] - ref[:, :3])) <= 4e-6
    ang = np.angle(np.exp(1j * (got[:, 3:6].astype(np.float64) - ref[:, 3:6])))
    assert np.max(np.abs(ang)) <= 2e-6
    assert np.array_equal(got[:, 6], ref[:, 6])
    frac_exact = np.mean(np.all(got.view(np.uint32) == ref.view(np.uint32), axis=1))
    print(f"{workload}, trig {trig}: bit-identical particles {frac_exact:.5f}")
    assert frac_exact == 1.0


def test_batch_motion_member_equals_alone(tsl, room):
    ev = tsl.TSDFEvaluator(room.map)
    mem = _members(room, SIZES, seed=21)
    motions = [tsl.evaluator.motion_from_delta(0.1 * i, 0.02, 0.0, 0.0, 0.0, 0.05 * i, 0.01, 0.005, seed=100 + i)
               for i in range(len(mem))]
    filters = []
    for parts, pts in mem:
        f = tsl.DeviceFilter(ev, parts.shape[0])
        f.set_particles(parts)
        f.set_scan(pts, room.tf)
        filters.append(f)
    b = tsl.Batch(ev, filters)
    s = tsl.Stream()
    b.motion_async(motions, s)
    b.sync(s)
    for i, (parts, pts) in enumerate(mem):
        alone = tsl.DeviceFilter(ev, parts.shape[0])
        alone.set_particles(parts)
        alone.motion_async(motions[i], s)
        s.synchronize()
        assert np.array_equal(filters[i].get_particles().view(np.uint32), alone.get_particles().view(np.uint32)), i
