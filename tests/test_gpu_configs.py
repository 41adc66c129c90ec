"""The benchmark workloads as parity cases, built by the same seeded
generators bench.py uses: global localisation (100,000 particles x 20,000
points on the sparse building map) checked against the oracle at full size,
and the batched-filters, serving and mixed (four maps, tracking and
global-localisation robots) workloads checked on members and robots against
the oracle, plus the bit-equality of each member or robot with the same
filter updated alone.

Bar: voxel indices and resampled indices bit-exact; weights within 1e-5
relative of the oracle's float64 sum."""
import numpy as np
import pytest

import oracle_binding as ob

pytestmark = pytest.mark.gpu

WEIGHT_RTOL = 1e-5


@pytest.fixture(scope="module")
def tsl():
    import tsdf_localization_amd as t
    assert t.device_count() >= 1, "no HIP device"
    return t


def rel_err(gpu, ref64):
    ref64 = np.asarray(ref64, np.float64)
    return np.abs(np.asarray(gpu, np.float64) - ref64) / np.maximum(np.abs(ref64), 1e-300)


@pytest.fixture(scope="module")
def building():
    from tsdf_localization_amd import synth
    return synth.building(num_particles=100_000)


def test_global_localisation_indices_subsample(tsl, building):
    """Every (particle, point) voxel index of 64 particles x the full scan."""
    w = building
    ev = tsl.TSDFEvaluator(w.map)
    om = ob.OracleMap.from_host_map(w.map)
    rng = np.random.default_rng(21)
    sub = w.particles[rng.choice(w.particles.shape[0], 64, replace=False)]
    print(f"global localisation: trace 64 x {w.points.shape[0]}", flush=True)
    assert np.array_equal(ev.trace_indices(sub, w.points, w.tf), om.trace(sub, w.points, w.tf))


def test_global_localisation_full_update(tsl, building):
    """The full update (2e9 pairs) on the device filter with the autotuned
    tile: every normalised weight against the oracle's float64 weights of
    all 100,000 particles, then the resampled set bit-exact."""
    w = building
    ev = tsl.TSDFEvaluator(w.map)
    om = ob.OracleMap.from_host_map(w.map)
    f = tsl.DeviceFilter(ev, w.particles.shape[0])
    f.set_particles(w.particles)
    f.set_scan(w.points, w.tf)
    s = tsl.Stream()
    f.autotune(s, reps=1)
    f.update_async(s)
    _, n = f.sync(s)
    got = f.get_particles()
    print("global localisation: oracle over 2e9 pairs", flush=True)
    _, w64 = om.evaluate(w.particles, w.points, w.tf)
    assert np.max(rel_err(got[:, 6], w64 / w64.sum())) <= WEIGHT_RTOL
    u0 = np.float32(0.37 / n)
    f.resample_async(u0, s)
    _, n2 = f.sync(s)
    ref_idx = ob.resample(got[:, 6], u0)
    assert n2 == ref_idx.size
    assert np.array_equal(f.get_particles(), got[ref_idx])


def test_batched_filters_members(tsl):
    """The 64-filter batch (one persistent launch per step): four members
    against the oracle, and each of them equal to the same filter alone with
    the batch's tile, bit for bit."""
    from tsdf_localization_amd import synth
    w = synth.batched()
    ev = tsl.TSDFEvaluator(w.map)
    om = ob.OracleMap.from_host_map(w.map)
    filters = []
    for r in w.robots:
        f = tsl.DeviceFilter(ev, r.particles.shape[0])
        f.set_particles(r.particles)
        f.set_scan(r.points, w.tf)
        filters.append(f)
    b = tsl.Batch(ev, filters)
    b.set_tile(2, 2048, 1)
    s = tsl.Stream()
    b.update_async(s)
    _, _, status = b.sync(s)
    assert set(status.tolist()) == {0}
    sizes = [r.particles.shape[0] * r.points.shape[0] for r in w.robots]
    picks = [int(np.argmin(sizes)), int(np.argsort(sizes)[len(sizes) // 2]), 5, 17]
    for i in picks:
        r = w.robots[i]
        got = filters[i].get_particles()
        print(f"batched member {i}: {r.particles.shape[0]} x {r.points.shape[0]}", flush=True)
        _, w64 = om.evaluate(r.particles, r.points, w.tf)
        assert np.max(rel_err(got[:, 6], w64 / w64.sum())) <= WEIGHT_RTOL, i
        alone = tsl.DeviceFilter(ev, r.particles.shape[0])
        alone.set_particles(r.particles)
        alone.set_scan(r.points, w.tf)
        alone.set_tile(2, 2048, 1)
        alone.set_order(0)
        alone.update_async(s)
        alone.sync(s)
        assert np.array_equal(alone.get_particles().view(np.uint32), got.view(np.uint32)), i


def test_serving_robots(tsl):
    """Requests of the serving workload through the server: each result
    equals the same robot's plain filter update bit for bit (motion, scan,
    update, resampling), and one robot's sensor update matches the oracle."""
    from tsdf_localization_amd import synth
    from tsdf_localization_amd.server import Server
    w = synth.serving()
    ev = tsl.TSDFEvaluator(w.map)
    picks = [0, 7, 64, 131, 255]
    srv = Server(ev, max_robots=len(picks), max_particles=20_000, max_batch=8, chunk_len=2048)
    try:
        ids = [srv.add_robot(w.robots[i].particles, 20_000, w.tf) for i in picks]
        motions = [tsl.evaluator.motion_from_delta(0.05, 0.0, 0.0, 0.0, 0.0, 0.01, 0.01, 0.005, seed=40 + k)
                   for k in range(len(picks))]
        u0 = [np.float32(0.61 / w.robots[i].particles.shape[0]) for i in picks]
        for k, i in enumerate(picks):
            srv.submit(ids[k], motions[k], w.robots[i].points, u0[k], request_id=k)
        res = {r["id"]: r for r in srv.wait(len(picks))}
        s = tsl.Stream()
        om = ob.OracleMap.from_host_map(w.map)
        for k, i in enumerate(picks):
            r = w.robots[i]
            assert res[k]["status"] == 0
            f = tsl.DeviceFilter(ev, r.particles.shape[0])
            f.set_particles(r.particles)
            f.set_tile(2, 2048, 1)
            f.set_order(0)
            f.motion_async(motions[k], s)
            s.synchronize()
            moved = f.get_particles()
            f.set_scan(r.points, w.tf)
            f.update_async(s)
            pose, _ = f.sync(s)
            weighted = f.get_particles()
            f.resample_async(u0[k], s)
            f.sync(s)
            assert res[k]["pose"] == pose, i
            assert np.array_equal(srv.get_particles(ids[k]).view(np.uint32), f.get_particles().view(np.uint32)), i
            if k == 1:
                print(f"serving robot {i}: oracle {r.particles.shape[0]} x {r.points.shape[0]}", flush=True)
                _, w64 = om.evaluate(moved, r.points, w.tf)
                assert np.max(rel_err(weighted[:, 6], w64 / w64.sum())) <= WEIGHT_RTOL
    finally:
        srv.close()


def test_mixed_classes(tsl):
    """The mixed workload at its stated size (four campus maps of ~7.3e9
    int16 voxels, ~14.6 GB each, block offsets past 2^31, one server each,
    tracking and global-localisation robots): one request of each class on
    two maps through the servers, each result equal to the same robot's
    plain filter update bit for bit; the global robot's weights (1,000,000 x
    4,000 pairs) and a tracking robot's against the oracle, which reads the
    campus through its int64 block table and int16 millimetres."""
    from tsdf_localization_amd import synth
    from tsdf_localization_amd.server import Server
    w = synth.mixed()
    s = tsl.Stream()
    for k in (0, 3):
        ev = tsl.TSDFEvaluator(w.maps[k])
        trk, glb = w.tracking[k][1], w.global_robots[k]
        gp = glb.particles.shape[0]
        srv = Server(ev, max_robots=len(w.tracking[k]) + 1, max_particles=gp, max_batch=8, chunk_len=2048)
        try:
            ids = [srv.add_robot(r.particles, r.points.shape[0], w.tf) for r in (trk, glb)]
            motions = [tsl.evaluator.motion_from_delta(0.05, 0.0, 0.0, 0.0, 0.0, 0.01, 0.01, 0.005, seed=90 + j)
                       for j in range(2)]
            u0 = [np.float32(0.43 / r.particles.shape[0]) for r in (trk, glb)]
            for j, r in enumerate((trk, glb)):
                srv.submit(ids[j], motions[j], r.points, u0[j], request_id=j)
            res = {r["id"]: r for r in srv.wait(2)}
            for j, r in enumerate((trk, glb)):
                assert res[j]["status"] == 0
                f = tsl.DeviceFilter(ev, r.particles.shape[0])
                f.set_particles(r.particles)
                # the server's chunking: 2048 points for the tracking robot,
                # the whole scan for the bulk (global) one
                f.set_tile(2, 2048 if j == 0 else (r.points.shape[0] + 63) // 64 * 64, 1)
                f.set_order(0)
                f.motion_async(motions[j], s)
                s.synchronize()
                moved = f.get_particles()
                f.set_scan(r.points, w.tf)
                f.update_async(s)
                pose, _ = f.sync(s)
                weighted = f.get_particles()
                f.resample_async(u0[j], s)
                f.sync(s)
                assert res[j]["pose"] == pose, (k, j)
                assert np.array_equal(srv.get_particles(ids[j]).view(np.uint32), f.get_particles().view(np.uint32))
                if k == 0:
                    print(f"mixed map {k} robot class {j}: oracle {r.particles.shape[0]} x {r.points.shape[0]}",
                          flush=True)
                    om = (ob.OracleMap64 if w.maps[k].wide else ob.OracleMap).from_host_map(w.maps[k])
                    _, w64 = om.evaluate(moved, r.points, w.tf)
                    assert np.max(rel_err(weighted[:, 6], w64 / w64.sum())) <= WEIGHT_RTOL, (k, j)
        finally:
            srv.close()


def test_batched_filters_graph_autotuned(tsl):
    """The batched workload exactly as run: 64 filters, the tile the batch
    autotuner picks, one step (restore + update + resample) captured in a
    HIP graph and replayed three times with eager work between replays.
    Every replay gives the same bits; four members equal the same filter
    alone with that tile, update and resampling bit for bit; their weights
    are within 1e-5 of the oracle."""
    from tsdf_localization_amd import synth
    w = synth.batched()
    ev = tsl.TSDFEvaluator(w.map)
    om = ob.OracleMap.from_host_map(w.map)
    s = tsl.Stream()
    filters = []
    for r in w.robots:
        f = tsl.DeviceFilter(ev, r.particles.shape[0])
        f.set_particles(r.particles)
        f.set_scan(r.points, w.tf)
        f.snapshot()
        filters.append(f)
    b = tsl.Batch(ev, filters)
    k, p, chunk, _ = b.autotune(s)
    print(f"batched autotune: K={k} points/lane={p} chunk={chunk or 'auto'}", flush=True)
    for f in filters:  # the autotuner left the members updated: back to the snapshots
        f.restore_async(s)
    s.synchronize()
    b.prepare()
    rng = np.random.default_rng(9)
    u0 = np.array([np.float32(rng.uniform(0, 0.9 / r.particles.shape[0])) for r in w.robots], np.float32)
    b.capture(s, restore=True)
    replays = []
    for rep in range(3):
        b.launch(u0, s)
        _, counts, status = b.sync(s)
        assert set(status.tolist()) == {0}, rep
        replays.append([f.get_particles() for f in filters])
        f0 = tsl.DeviceFilter(ev, 64)  # eager work between replays
        f0.set_particles(w.robots[0].particles[:64])
    for rep in (1, 2):
        for a, c in zip(replays[0], replays[rep]):
            assert np.array_equal(a.view(np.uint32), c.view(np.uint32)), rep
    sizes = [r.particles.shape[0] * r.points.shape[0] for r in w.robots]
    picks = [int(np.argmin(sizes)), int(np.argmax(sizes)), 5, 40]
    for i in picks:
        r = w.robots[i]
        n = r.points.shape[0]
        eff_chunk = chunk or min(2048, (n + 63) // 64 * 64)
        alone = tsl.DeviceFilter(ev, r.particles.shape[0])
        alone.set_particles(r.particles)
        alone.set_scan(r.points, w.tf)
        alone.set_tile(k, eff_chunk, p)
        alone.set_order(0)
        alone.update_async(s)
        alone.sync(s)
        weighted = alone.get_particles()
        _, w64 = om.evaluate(r.particles, r.points, w.tf)
        print(f"batched member {i}: {r.particles.shape[0]} x {n}", flush=True)
        assert np.max(rel_err(weighted[:, 6], w64 / w64.sum())) <= WEIGHT_RTOL, i
        alone.resample_async(u0[i], s)
        alone.sync(s)
        assert np.array_equal(alone.get_particles().view(np.uint32), replays[0][i].view(np.uint32)), i
