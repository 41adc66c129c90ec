"""The float form of a map: a CudaSubVoxelMap whose float data holds more
distinct values than the 4096-entry palette (cuda_sub_voxel_map.h:141,166
takes any float data) is kept on the device as the floats themselves.
Indices and values bit-exact against the oracle, weights within 1e-5 of its
float64 sum, every tile shape giving the same bits, bricked and linear
layouts, the exact tiers (negative band, far particles) included.
"""
import math

import numpy as np
import pytest

import oracle_binding as ob
from test_gpu_parity import oracle_values, rel_err, WEIGHT_RTOL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tsl():
    import tsdf_localization_amd as t
    assert t.device_count() >= 1, "no HIP device"
    return t


def _spread(data, seed):
    """The map's values with a distinct perturbation per voxel: far more
    than 4096 distinct values (init_value voxels included)."""
    rng = np.random.default_rng(seed)
    return (data * (np.float32(1.0) + rng.uniform(-1e-3, 1e-3, data.shape).astype(np.float32))
            + rng.uniform(0.0, 1e-4, data.shape).astype(np.float32)).astype(np.float32)


@pytest.fixture(scope="module")
def room_f32():
    from tsdf_localization_amd import synth
    room = synth.room()
    data = _spread(ob.float_data(room.map), 31)
    assert np.unique(data.view(np.uint32)).size > 100_000
    return room, data


def _weights(tsl, ev, parts, pts, tf, k, pp, chunk, order):
    f = tsl.DeviceFilter(ev, parts.shape[0])
    f.set_particles(parts)
    f.set_scan(pts, tf)
    f.set_tile(k, chunk, pp)
    f.set_order(order)
    s = tsl.Stream()
    f.update_async(s)
    f.sync(s)
    return f.get_particles()[:, 6]


def test_room_float_form_indices_values_weights(tsl, room_f32):
    room, data = room_f32
    m = tsl.CudaSubVoxelMap(room.map.coef(), room.map.grid_occ, data=data)
    ev = tsl.TSDFEvaluator(m)
    om = ob.OracleMap(room.map.coef(), room.map.grid_occ, data)
    parts = room.particles[:300]
    idx = ev.trace_indices(parts, room.points, room.tf)
    ref = om.trace(parts, room.points, room.tf)
    assert np.array_equal(idx, ref)
    vals = ev.trace_values(parts, room.points, room.tf)
    assert np.array_equal(vals.view(np.uint32), oracle_values(om, ref).view(np.uint32))
    raw = ev.evaluate_raw(parts, room.points, room.tf)
    _, w64 = om.evaluate(parts, room.points, room.tf)
    assert np.max(rel_err(raw, w64)) <= WEIGHT_RTOL


def test_room_float_form_tiles_bit_identical(tsl, room_f32):
    """For a given chunk length every tile shape and particle order gives the
    same weights bit for bit, each within the bar of the oracle."""
    room, data = room_f32
    m = tsl.CudaSubVoxelMap(room.map.coef(), room.map.grid_occ, data=data)
    ev = tsl.TSDFEvaluator(m)
    om = ob.OracleMap(room.map.coef(), room.map.grid_occ, data)
    parts, pts, tf = room.particles, room.points, room.tf
    _, w64 = om.evaluate(parts, pts, tf)
    ref_norm = w64 / w64.sum()
    for chunk in (320, 0):
        first = None
        for k, pp, order in [(1, 1, 0), (1, 2, 1), (2, 4, 0), (2, 1, 1), (1, 4, 1)]:
            w = _weights(tsl, ev, parts, pts, tf, k, pp, chunk, order)
            assert np.max(rel_err(w, ref_norm)) <= WEIGHT_RTOL, (k, pp, chunk, order)
            if first is None:
                first = w
            else:
                assert np.array_equal(w.view(np.uint32), first.view(np.uint32)), (k, pp, chunk, order)


def test_linear_layout_and_exact_tiers(tsl):
    """sub_dim 13 (no bricks), setData-built block table, particles around
    the low faces (negative band) and far away (general tier)."""
    rng = np.random.default_rng(41)
    cells = rng.uniform((0.0, 0.0, 0.0), (6.4, 6.4, 3.2), (40_000, 3)).astype(np.float32)
    # positive values, as likelihood maps hold (a sum that cancels to near
    # zero has no relative bound)
    values = rng.uniform(0.0, 1.0, 40_000).astype(np.float32)
    m = tsl.CudaSubVoxelMap.from_cells((0, 0, 0), (6.4, 6.4, 3.2), 0.08, 0.003, cells, values)
    c = m.coef()
    assert c.sub_dim % 4 != 0
    assert np.unique(m.data.view(np.uint32)).size > 4096
    ev = tsl.TSDFEvaluator(m)
    om = ob.OracleMap(c, m.grid_occ, m.data)
    P = 240
    parts = np.zeros((P, 7), np.float32)
    parts[:, 0] = rng.uniform(-3, 7, P)
    parts[:, 1] = rng.uniform(-3, 7, P)
    parts[:, 2] = rng.uniform(-2, 3, P)
    parts[:, 3:5] = rng.uniform(-0.3, 0.3, (P, 2))
    parts[:, 5] = rng.uniform(-math.pi, math.pi, P)
    parts[:, 6] = 1.0 / P
    parts[:6, 0] += 2e7
    pts = rng.uniform(-6, 6, (1500, 3)).astype(np.float32)
    from tsdf_localization_amd import synth
    tf = synth.tf_from_xyzrpy(0.1, 0.0, 0.4, 0.0, 0.02, 0.03)
    idx = ev.trace_indices(parts, pts, tf)
    ref = om.trace(parts, pts, tf)
    assert np.count_nonzero(ref != c.data_size) > 10_000
    assert np.array_equal(idx, ref)
    vals = ev.trace_values(parts, pts, tf)
    assert np.array_equal(vals.view(np.uint32), oracle_values(om, ref).view(np.uint32))
    raw = ev.evaluate_raw(parts, pts, tf)
    _, w64 = om.evaluate(parts, pts, tf)
    assert np.max(rel_err(raw, w64)) <= WEIGHT_RTOL


def test_palette_and_float_forms_agree(tsl, room_f32):
    """A map at the palette's limit (4096 distinct values) and the same map
    with one more distinct value (float form) give the same bits wherever
    the extra value is not looked up."""
    room, _ = room_f32
    data = ob.float_data(room.map).copy()
    grid = room.map.grid_occ
    ds = data.size
    # 4096 distinct values in all: 4095 spread values and init_value
    init = np.float32(room.map.coef().init_value)
    rng = np.random.default_rng(5)
    pal = np.unique(rng.uniform(0.0, 2.0, 5000).astype(np.float32))[:4095]
    assert pal.size == 4095 and not np.any(pal == init)
    data = np.where(data == init, init, pal[np.arange(ds) % 4095]).astype(np.float32)
    assert np.unique(data.view(np.uint32)).size == 4096
    ev_p = tsl.TSDFEvaluator(tsl.CudaSubVoxelMap(room.map.coef(), grid, data=data))
    parts = room.particles[:200]
    idx = ev_p.trace_indices(parts, room.points, room.tf)
    # one voxel no lookup of this test reaches takes a 4097th value
    unused = np.setdiff1d(np.arange(ds, dtype=np.uint64), idx.ravel())
    assert unused.size > 0
    extra = data.copy()
    extra[int(unused[-1])] = np.float32(7.25)
    ev_f = tsl.TSDFEvaluator(tsl.CudaSubVoxelMap(room.map.coef(), grid, data=extra))
    assert np.array_equal(ev_f.trace_indices(parts, room.points, room.tf), idx)
    wp = ev_p.evaluate_raw(parts, room.points, room.tf)
    wf = ev_f.evaluate_raw(parts, room.points, room.tf)
    assert np.array_equal(wp.view(np.uint32), wf.view(np.uint32))


def test_float_form_batch_and_server(tsl, room_f32):
    """The batched and served forms on a float-form map (their own kernel
    instances): every batch member and every served robot gives the bits of
    the same update run alone with the same chunk length."""
    from tsdf_localization_amd.server import Server
    room, data = room_f32
    m = tsl.CudaSubVoxelMap(room.map.coef(), room.map.grid_occ, data=data)
    ev = tsl.TSDFEvaluator(m)
    rng = np.random.default_rng(9)
    chunk = 320
    mem = []
    for P, N in [(300, 900), (1000, 2000), (37, 64)]:
        parts = room.particles[rng.choice(room.particles.shape[0], P, replace=False)].copy()
        parts[:, 6] = 1.0 / P
        mem.append((parts, room.points[rng.permutation(room.points.shape[0])[:N]]))
    u0 = [np.float32(0.41 / p.shape[0]) for p, _ in mem]
    s = tsl.Stream()

    def alone(parts, pts, u, motion=None, ch=chunk):
        f = tsl.DeviceFilter(ev, parts.shape[0])
        f.set_particles(parts)
        f.set_tile(2, ch, 1)
        f.set_order(0)
        if motion is not None:
            f.motion_async(motion, s)
        f.set_scan(pts, room.tf)
        f.update_async(s)
        pose, _ = f.sync(s)
        f.resample_async(u, s)
        f.sync(s)
        return pose, f.get_particles()

    # batch
    for k, pp in [(1, 2), (2, 1)]:
        filters = []
        for parts, pts in mem:
            f = tsl.DeviceFilter(ev, parts.shape[0])
            f.set_particles(parts)
            f.set_scan(pts, room.tf)
            filters.append(f)
        b = tsl.Batch(ev, filters)
        b.set_tile(k, chunk, pp)
        b.update_async(s)
        b.sync(s)
        b.resample_async(np.array(u0, np.float32), s)
        b.sync(s)
        for i, (parts, pts) in enumerate(mem):
            _, ref = alone(parts, pts, u0[i])
            assert np.array_equal(filters[i].get_particles().view(np.uint32), ref.view(np.uint32)), (k, pp, i)
    # server
    srv = Server(ev, max_robots=4, max_particles=2000, max_batch=4, chunk_len=chunk)
    try:
        ids = [srv.add_robot(p, 3000, room.tf) for p, _ in mem]
        motions = [tsl.evaluator.motion_from_delta(0.05, 0.01, 0.0, 0.0, 0.0, 0.02, 0.01, 0.005, seed=300 + i)
                   for i in range(len(mem))]
        for i, (p, pts) in enumerate(mem):
            srv.submit(ids[i], motions[i], pts, u0[i], request_id=i)
        res = {r["id"]: r for r in srv.wait(len(mem))}
        for i, (parts, pts) in enumerate(mem):
            assert res[i]["status"] == 0
            pose, ref = alone(parts, pts, u0[i], motions[i])
            assert res[i]["pose"] == pose, i
            assert np.array_equal(srv.get_particles(ids[i]).view(np.uint32), ref.view(np.uint32)), i
    finally:
        srv.close()
