"""Serving loop: resident robots, continuous batching, HBM admission.
A robot's result must be the same bits alone and in a mix, equal to the
single-filter path (motion, scan, update, resampling with the same chunk),
and the evaluation must match the oracle."""
import numpy as np
import pytest

import oracle_binding as ob

pytestmark = pytest.mark.gpu

CHUNK = 512


@pytest.fixture(scope="module")
def tsl():
    import tsdf_localization_amd as t
    assert t.device_count() >= 1, "no HIP device"
    return t


@pytest.fixture(scope="module")
def room():
    from tsdf_localization_amd import synth
    return synth.room()


def _robots(room, k, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        P = int(rng.integers(100, 1000))
        parts = room.particles[rng.choice(room.particles.shape[0], P, replace=False)].copy()
        parts[:, 6] = 1.0 / P
        N = int(rng.integers(200, room.points.shape[0]))
        pts = room.points[np.sort(rng.choice(room.points.shape[0], N, replace=False))]
        out.append((parts, pts))
    return out


def _motion(tsl, i):
    return tsl.evaluator.motion_from_delta(0.05 * (i % 3), 0.01, 0.0, 0.0, 0.0, 0.02 * (i % 5), 0.01, 0.005,
                                           seed=1000 + i)


@pytest.mark.parametrize("chunk", [CHUNK, 0], ids=["chunk512", "whole-scan"])
def test_server_alone_equals_mix_equals_filter(tsl, room, chunk):
    from tsdf_localization_amd.server import Server
    ev = tsl.TSDFEvaluator(room.map)
    robots = _robots(room, 8)
    srv = Server(ev, max_robots=32, max_particles=2000, max_batch=16, chunk_len=chunk)
    try:
        ids_a = [srv.add_robot(p, 3000, room.tf) for p, _ in robots]
        ids_b = [srv.add_robot(p, 3000, room.tf) for p, _ in robots]  # twins
        u0 = [np.float32(0.31 / p.shape[0]) for p, _ in robots]
        # robot 0 alone
        srv.submit(ids_a[0], _motion(tsl, 0), robots[0][1], u0[0], request_id=1)
        alone = srv.wait(1)[0]
        assert alone["batch"] == 1 and alone["status"] == 0
        # all twins together
        for i, (p, pts) in enumerate(robots):
            srv.submit(ids_b[i], _motion(tsl, i), pts, u0[i], request_id=100 + i)
        mix = {r["id"]: r for r in srv.wait(len(robots))}
        assert max(r["batch"] for r in mix.values()) > 1
        assert mix[100]["pose"] == alone["pose"] and mix[100]["count"] == alone["count"]
        pa = srv.get_particles(ids_a[0])
        pb = srv.get_particles(ids_b[0])
        assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
        for r in mix.values():
            assert r["status"] == 0
            assert r["submit_ms"] <= r["start_ms"] <= r["done_ms"]
        # single-filter path with the same steps
        for i in (0, 3, 7):
            parts, pts = robots[i]
            f = tsl.DeviceFilter(ev, parts.shape[0])
            f.set_particles(parts)
            f.set_tile(1, chunk or (pts.shape[0] + 63) // 64 * 64, 1)
            f.set_order(0)
            s = tsl.Stream()
            f.motion_async(_motion(tsl, i), s)
            s.synchronize()
            moved = f.get_particles()
            f.set_scan(pts, room.tf)
            f.update_async(s)
            pose, _ = f.sync(s)
            weighted = f.get_particles()
            f.resample_async(u0[i], s)
            f.sync(s)
            assert np.array_equal(srv.get_particles(ids_b[i]).view(np.uint32), f.get_particles().view(np.uint32)), i
            assert mix[100 + i]["pose"] == pose
            # the evaluation against the oracle
            om = ob.OracleMap.from_host_map(room.map)
            _, w64 = om.evaluate(moved, pts, room.tf)
            ref = w64 / w64.sum()
            assert np.max(np.abs(weighted[:, 6].astype(np.float64) - ref) / ref) <= 1e-5
    finally:
        srv.close()


def test_server_admission_by_budget(tsl, room):
    from tsdf_localization_amd.server import Server
    ev = tsl.TSDFEvaluator(room.map)
    parts = room.particles[:500]
    probe = Server(ev, max_robots=4, max_particles=1000, max_batch=4)
    try:
        base = probe.stats()["used_bytes"]
        r = probe.add_robot(parts, 3000, room.tf)
        per_robot = probe.stats()["used_bytes"] - base
    finally:
        probe.close()
    srv = Server(ev, hbm_budget=base + 2 * per_robot + per_robot // 2, max_robots=8, max_particles=1000, max_batch=4)
    try:
        a = srv.add_robot(parts, 3000, room.tf)
        b = srv.add_robot(parts, 3000, room.tf)
        with pytest.raises(tsl.TslError, match="bytes"):
            srv.add_robot(parts, 3000, room.tf)
        srv.remove_robot(a)
        c = srv.add_robot(parts, 3000, room.tf)
        assert c == a
        srv.submit(b, _motion(tsl, 1), room.points, np.float32(0.1 / 500), request_id=7)
        srv.submit(c, _motion(tsl, 2), room.points, np.float32(0.2 / 500), request_id=8)
        res = srv.wait(2)
        assert sorted(r["id"] for r in res) == [7, 8] and all(r["status"] == 0 for r in res)
    finally:
        srv.close()


def test_server_many_requests_per_robot_in_order(tsl, room):
    """Requests of one robot run one after the other, in submission order."""
    from tsdf_localization_amd.server import Server
    ev = tsl.TSDFEvaluator(room.map)
    robots = _robots(room, 3, seed=11)
    srv = Server(ev, max_robots=4, max_particles=2000, max_batch=8, chunk_len=CHUNK)
    try:
        ids = [srv.add_robot(p, 3000, room.tf) for p, _ in robots]
        for step in range(4):
            for i, (p, pts) in enumerate(robots):
                srv.submit(ids[i], _motion(tsl, step), pts, np.float32(0.5 / p.shape[0]), request_id=step * 10 + i)
        res = srv.wait(12)
        for i in range(3):
            mine = [r for r in res if r["robot"] == ids[i]]
            assert [r["id"] for r in sorted(mine, key=lambda r: r["done_ms"])] == [s * 10 + i for s in range(4)]
    finally:
        srv.close()


def test_server_tick_scan_preparation_paths_agree(tsl, room):
    """A tick of one scan sorts it on its own; larger ticks sort all their
    points together (by key, then by scan).  Both give the same points in the
    same order: every robot run alone equals its twin run in a large tick,
    bit for bit."""
    from tsdf_localization_amd.server import Server
    ev = tsl.TSDFEvaluator(room.map)
    robots = _robots(room, 12, seed=11)
    srv = Server(ev, max_robots=32, max_particles=2000, max_batch=16)
    try:
        ids_a = [srv.add_robot(p, 3000, room.tf) for p, _ in robots]
        ids_b = [srv.add_robot(p, 3000, room.tf) for p, _ in robots]
        u0 = [np.float32(0.27 / p.shape[0]) for p, _ in robots]
        alone = {}
        for i, (p, pts) in enumerate(robots):
            srv.submit(ids_a[i], _motion(tsl, i), pts, u0[i], request_id=i)
            r = srv.wait(1)[0]
            assert r["batch"] == 1 and r["status"] == 0
            alone[i] = r
        for i, (p, pts) in enumerate(robots):
            srv.submit(ids_b[i], _motion(tsl, i), pts, u0[i], request_id=100 + i)
        mix = {r["id"] - 100: r for r in srv.wait(len(robots))}
        assert max(r["batch"] for r in mix.values()) >= 4, "no tick with several scans"
        for i in range(len(robots)):
            assert mix[i]["status"] == 0
            assert mix[i]["pose"] == alone[i]["pose"] and mix[i]["count"] == alone[i]["count"], i
            pa = srv.get_particles(ids_a[i])
            pb = srv.get_particles(ids_b[i])
            assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32)), i
    finally:
        srv.close()


def test_bulk_robot_sliced_equals_unsliced_and_tracking_interleaves(tsl):
    """A robot of 70,000 particles (the bulk class: its own stream, its
    evaluation cut into launches) gives the same bits with slices of 2^22
    pairs as in one launch, and as the plain filter; a tracking robot's
    request submitted right after it completes on the other stream."""
    from tsdf_localization_amd import synth
    from tsdf_localization_amd.server import Server
    w = synth.building(num_particles=70_000, target_points=6000)
    ev = tsl.TSDFEvaluator(w.map)
    trk = w.particles[:800].copy()
    trk[:, 6] = 1.0 / trk.shape[0]
    motion = _motion(tsl, 7)
    u0 = np.float32(0.29 / w.particles.shape[0])
    outs = []
    for slice_pairs in (1 << 22, 0):
        srv = Server(ev, max_robots=4, max_particles=70_000, max_batch=4, chunk_len=2048)
        try:
            srv.set_slice(slice_pairs)
            big = srv.add_robot(w.particles, w.points.shape[0], w.tf)
            small = srv.add_robot(trk, w.points.shape[0], w.tf)
            srv.submit(big, motion, w.points, u0, request_id=1)
            srv.submit(small, _motion(tsl, 8), w.points[:2000], np.float32(0.5 / 800), request_id=2)
            res = {r["id"]: r for r in srv.wait(2)}
            assert res[1]["status"] == 0 and res[2]["status"] == 0
            if slice_pairs:
                # the tracking request ran on the other stream, not behind the bulk one
                assert res[2]["done_ms"] <= res[1]["done_ms"]
            outs.append((res[1]["pose"], srv.get_particles(big)))
        finally:
            srv.close()
    assert outs[0][0] == outs[1][0]
    assert np.array_equal(outs[0][1].view(np.uint32), outs[1][1].view(np.uint32))
    # the plain filter with the same chunk (a bulk robot's: the whole scan)
    s = tsl.Stream()
    f = tsl.DeviceFilter(ev, w.particles.shape[0])
    f.set_particles(w.particles)
    f.set_tile(2, (w.points.shape[0] + 63) // 64 * 64, 1)
    f.motion_async(motion, s)
    f.set_scan(w.points, w.tf)
    f.update_async(s)
    pose, _ = f.sync(s)
    f.resample_async(u0, s)
    f.sync(s)
    assert outs[0][0] == pose
    assert np.array_equal(outs[0][1].view(np.uint32), f.get_particles().view(np.uint32))


def test_remove_while_submitting_and_reading(tsl, room):
    """Requests submitted from one thread and particles read from another
    while the robot is removed: submits after the removal fail with "no
    robot", reads never see a freed set, and the server then closes."""
    import threading
    import time
    from tsdf_localization_amd.server import Server
    ev = tsl.TSDFEvaluator(room.map)
    srv = Server(ev, max_robots=4, max_particles=2000, max_batch=4)
    rid = srv.add_robot(room.particles, room.points.shape[0], room.tf)
    other = srv.add_robot(room.particles[:500], room.points.shape[0], room.tf)
    stop = threading.Event()
    errors = {"submit": 0, "read": 0, "read_ok": 0}

    def submitter():
        i = 0
        while not stop.is_set():
            try:
                srv.submit(rid, _motion(tsl, i), room.points, np.float32(0.3 / 1000), request_id=i)
            except tsl.TslError:
                errors["submit"] += 1
            i += 1
            time.sleep(0.0005)

    def reader():
        while not stop.is_set():
            try:
                p = srv.get_particles(rid)
                assert p.shape[1] == 7 and p.shape[0] > 0
                errors["read_ok"] += 1
            except tsl.TslError:
                errors["read"] += 1
            time.sleep(0.001)

    ts = [threading.Thread(target=submitter), threading.Thread(target=reader)]
    for t in ts:
        t.start()
    time.sleep(0.2)
    removed = False
    t_end = time.time() + 20.0
    while time.time() < t_end:
        srv.poll(1024, wait_ms=1)
        try:
            srv.remove_robot(rid)
            removed = True
            break
        except tsl.TslError as e:
            assert "pending" in str(e)
    time.sleep(0.1)
    stop.set()
    for t in ts:
        t.join(timeout=30)
        assert not t.is_alive()
    assert removed, "the robot was never without pending requests"
    assert errors["submit"] > 0 and errors["read_ok"] > 0, errors
    srv.submit(other, _motion(tsl, 1), room.points, np.float32(0.3 / 500), request_id=10**6)
    res = []
    for _ in range(400):
        res += [r for r in srv.poll(1024, wait_ms=20) if r["id"] == 10**6]
        if res:
            break
    assert res and res[0]["status"] == 0
    closer = threading.Thread(target=srv.close)
    closer.start()
    closer.join(timeout=60)
    assert not closer.is_alive(), "server did not close"
