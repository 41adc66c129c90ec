"""Parity of the HIP sensor update with the CPU oracle, through the C ABI.

Bar (per the path's contract): voxel indices and resampled indices
bit-exact; per-particle weights within 1e-5 relative of the oracle's float64
sum of the same float terms.
"""
import math

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


@pytest.fixture(scope="module")
def room():
    from tsdf_localization_amd import synth
    return synth.room()


@pytest.fixture(scope="module")
def room_oracle(room):
    return ob.OracleMap.from_host_map(room.map)


def rel_err(gpu, ref64):
    ref64 = np.asarray(ref64, np.float64)
    return np.abs(np.asarray(gpu, np.float64) - ref64) / np.maximum(np.abs(ref64), 1e-300)


def test_pose_matrices_bit_exact(tsl, room):
    L = tsl.pose_matrices(room.particles, room.tf)
    ref = np.stack([ob.pose_matrix(p[:6], room.tf) for p in room.particles])
    assert np.array_equal(L.view(np.uint32), ref.view(np.uint32))


def test_room_voxel_indices_bit_exact(tsl, room, room_oracle):
    ev = tsl.TSDFEvaluator(room.map)
    gpu = ev.trace_indices(room.particles, room.points, room.tf)
    ref = room_oracle.trace(room.particles, room.points, room.tf)
    assert gpu.shape == (1000, room.points.shape[0])
    mism = np.count_nonzero(gpu != ref)
    assert mism == 0, f"{mism} of {gpu.size} voxel indices differ"


def test_room_raw_weights(tsl, room, room_oracle):
    ev = tsl.TSDFEvaluator(room.map)
    raw = ev.evaluate_raw(room.particles, room.points, room.tf)
    w32, w64 = room_oracle.evaluate(room.particles, room.points, room.tf)
    assert np.max(rel_err(raw, w64)) <= WEIGHT_RTOL
    # The reference's own float sequential sum over ~2k terms is itself only
    # within ~2e-5 of the float64 sum; the device sum is far tighter.
    assert np.max(rel_err(raw, w64)) < np.max(rel_err(w32, w64))


def test_room_float_map_path(tsl, room, room_oracle):
    """Map uploaded from the reference's float voxel array (palette path)."""
    m = tsl.CudaSubVoxelMap(room.map.coef(), room.map.grid_occ, data=ob.float_data(room.map))
    ev = tsl.TSDFEvaluator(m)
    raw = ev.evaluate_raw(room.particles[:200], room.points, room.tf)
    _, w64 = room_oracle.evaluate(room.particles[:200], room.points, room.tf)
    assert np.max(rel_err(raw, w64)) <= WEIGHT_RTOL


@pytest.mark.parametrize("params", [(0.9, 0.1, 0.0, 100.0), (1.0, 0.0, 0.0, 100.0), (0.5, 0.3, 0.2, 6.0)])
def test_room_evaluate_normalised(tsl, room, room_oracle, params):
    a_hit, a_range, a_max, max_range = params
    ev = tsl.TSDFEvaluator(room.map, a_hit=a_hit, a_range=a_range, a_max=a_max, max_range=max_range)
    parts = room.particles.copy()
    pose = ev.evaluate(parts, room.points, room.tf)
    _, w64 = room_oracle.evaluate(room.particles, room.points, room.tf, a_hit, a_range, a_max, max_range)
    ref_norm = w64 / w64.sum()
    assert np.max(rel_err(parts[:, 6], ref_norm)) <= WEIGHT_RTOL
    assert np.array_equal(parts[:, :6], room.particles[:, :6])
    wn, pref = ob.normalize(room.particles, w64.astype(np.float32))
    assert np.allclose(pose["position"], pref[:3], rtol=0, atol=1e-4)
    q = np.array(pose["orientation"])
    assert abs(np.linalg.norm(q) - 1.0) < 1e-6


def test_negative_band_and_outside(tsl):
    """Particles around and outside the map's low faces: offsets below zero
    take the x86-semantics path; indices and values must match bit for
    bit."""
    from tsdf_localization_amd import synth
    sc = synth.Scene()
    sc.add_box((-5, -5, -5), (20, 20, 0.3))
    sc.add_box((-5, -5, -5), (0.25, 20, 20))
    sc.add_box((-5, -5, -5), (20, 0.25, 20))
    sc.add_cylinder(2.0, 2.0, 0.0, 2.5, 0.4)
    m = sc.tsdf_map((0, 0, 0), (6.4, 6.4, 3.2), 0.05, 300, dense=True)
    rng = np.random.default_rng(7)
    P = 300
    parts = np.zeros((P, 7), np.float32)
    parts[:, 0] = rng.uniform(-3, 4, P)
    parts[:, 1] = rng.uniform(-3, 4, P)
    parts[:, 2] = rng.uniform(-2, 2, P)
    parts[:, 3:5] = rng.uniform(-0.3, 0.3, (P, 2))
    parts[:, 5] = rng.uniform(-math.pi, math.pi, P)
    parts[:, 6] = 1.0 / P
    pts = rng.uniform(-8, 8, (1500, 3)).astype(np.float32)
    tf = synth.tf_from_xyzrpy(0.1, 0.0, 0.4, 0.0, 0.02, 0.03)
    ev = tsl.TSDFEvaluator(m)
    om = ob.OracleMap.from_host_map(m)
    gpu = ev.trace_indices(parts, pts, tf)
    ref = om.trace(parts, pts, tf)
    ds = m.coef().data_size
    assert np.count_nonzero(ref != ds) > 1000  # the case exercises real lookups
    mism = np.count_nonzero(gpu != ref)
    assert mism == 0, f"{mism} mismatches"
    got = ev.trace_values(parts, pts, tf)
    assert np.array_equal(got.view(np.uint32), oracle_values(om, ref).view(np.uint32))
    raw = ev.evaluate_raw(parts, pts, tf)
    _, w64 = om.evaluate(parts, pts, tf)
    assert np.max(rel_err(raw, w64)) <= WEIGHT_RTOL


@pytest.mark.parametrize("P,N", [(1, 1), (3, 65), (67, 129), (257, 1000), (1000, 1)])
def test_ragged_sizes(tsl, room, room_oracle, P, N):
    ev = tsl.TSDFEvaluator(room.map)
    parts = room.particles[:P]
    pts = room.points[:N]
    raw = ev.evaluate_raw(parts, pts, room.tf)
    _, w64 = room_oracle.evaluate(parts, pts, room.tf)
    assert np.max(rel_err(raw, w64)) <= WEIGHT_RTOL


def _filter_weights(tsl, ev, parts, pts, tf, k, pp, chunk, order=0, lookup=0):
    f = tsl.DeviceFilter(ev, parts.shape[0])
    f.set_particles(parts)
    f.set_scan(pts, tf)
    f.set_tile(k, chunk, pp)
    f.set_order(order)
    f.set_block_lookup(lookup)
    s = tsl.Stream()
    f.update_async(s)
    f.sync(s)
    return f.get_particles()[:, 6]


def _edge_scene(tsl):
    from tsdf_localization_amd import synth
    sc = synth.Scene()
    sc.add_box((-5, -5, -5), (20, 20, 0.3))
    sc.add_box((-5, -5, -5), (0.25, 20, 20))
    sc.add_cylinder(2.0, 2.0, 0.0, 2.5, 0.4)
    m = sc.tsdf_map((0, 0, 0), (6.4, 6.4, 3.2), 0.05, 300, dense=False)
    rng = np.random.default_rng(17)
    P = 200
    parts = np.zeros((P, 7), np.float32)
    parts[:, 0] = rng.uniform(-3, 4, P)
    parts[:, 1] = rng.uniform(-3, 4, P)
    parts[:, 2] = rng.uniform(-2, 2, P)
    parts[:, 5] = rng.uniform(-math.pi, math.pi, P)
    parts[:, 6] = 1.0 / P
    parts[:8, 0] += 2e7  # a few particles in the general tier
    pts = rng.uniform(-8, 8, (1300, 3)).astype(np.float32)
    tf = synth.tf_from_xyzrpy(0.1, 0.0, 0.4, 0.0, 0.02, 0.03)
    return m, parts, pts, tf


@pytest.mark.parametrize("scene", ["room", "edges"])
def test_tile_shapes_bit_identical(tsl, room, room_oracle, scene):
    """Every (particle order, particles per wavefront, points per lane)
    shape gives the same weights bit for bit for a given chunk length, on the
    fast tier (room) and with the wrap/general tiers mixed in (edges), within
    1e-5 of the oracle."""
    if scene == "room":
        m, parts, pts, tf, om = room.map, room.particles, room.points, room.tf, room_oracle
    else:
        m, parts, pts, tf = _edge_scene(tsl)
        om = ob.OracleMap.from_host_map(m)
    ev = tsl.TSDFEvaluator(m)
    _, w64 = om.evaluate(parts, pts, tf)
    ref_norm = w64 / w64.sum()
    for chunk in (320, 0):
        first = None
        for order, lookup in ((0, 0), (1, 0), (0, 1), (1, 1)):
            for k in (1, 2):
                for pp in (1, 2, 4):
                    w = _filter_weights(tsl, ev, parts, pts, tf, k, pp, chunk, order, lookup)
                    key = (order, lookup, k, pp, chunk)
                    assert np.max(rel_err(w, ref_norm)) <= WEIGHT_RTOL, key
                    if first is None:
                        first = w
                    else:
                        assert np.array_equal(w.view(np.uint32), first.view(np.uint32)), key


def test_per_point_tile_same_result(tsl, room, room_oracle):
    ev = tsl.TSDFEvaluator(room.map, per_point=True)
    raw = ev.evaluate_raw(room.particles, room.points, room.tf)
    _, w64 = room_oracle.evaluate(room.particles, room.points, room.tf)
    assert np.max(rel_err(raw, w64)) <= WEIGHT_RTOL


def test_empty_scan_returns_default_pose(tsl, room):
    ev = tsl.TSDFEvaluator(room.map)
    parts = room.particles.copy()
    pose = ev.evaluate(parts, np.zeros((0, 3), np.float32), room.tf)
    assert pose["position"] == (0.0, 0.0, 0.0)
    assert np.array_equal(parts, room.particles)


def test_no_particles_raises(tsl, room):
    # a non-empty scan over an empty particle set leaves the weight sum at
    # zero, which the reference rejects (src/cuda/cuda_evaluator.cu:366-368,
    # src/evaluation/tsdf_evaluator.cpp:159-161)
    ev = tsl.TSDFEvaluator(room.map)
    with pytest.raises(RuntimeError, match="No particle is valid!"):
        ev.evaluate(np.zeros((0, room.particles.shape[1]), np.float32), room.points, room.tf)


def test_zero_weight_sum_raises(tsl, room):
    ev = tsl.TSDFEvaluator(room.map, a_hit=0.0, a_range=0.0, a_max=0.0)
    parts = room.particles.copy()
    with pytest.raises(RuntimeError, match="No particle is valid!"):
        ev.evaluate(parts, room.points, room.tf)


# ---------------------------------------------------------------- resampling

def _weights(kind, M, rng):
    if kind == "random":
        w = rng.gamma(0.5, 1.0, M)
    elif kind == "uniform":
        w = np.ones(M)
    elif kind == "onehot":
        w = np.zeros(M)
        w[M // 3] = 1.0
    elif kind == "tiny":  # forces rounding in the double running sum
        w = rng.uniform(0, 1, M) * np.where(rng.uniform(0, 1, M) < 0.5, 1e-12, 1.0)
    elif kind == "sparse":
        w = np.where(rng.uniform(0, 1, M) < 0.01, rng.uniform(0, 1, M), 0.0)
        w[0] = 1.0
    w = (w / w.sum()).astype(np.float32)
    return w


@pytest.mark.parametrize("kind", ["random", "uniform", "onehot", "tiny", "sparse"])
@pytest.mark.parametrize("M", [1, 7, 1000, 4096, 100_000])
def test_resample_bit_exact(tsl, kind, M):
    rng = np.random.default_rng(M + len(kind))
    w = _weights(kind, M, rng)
    parts = np.zeros((M, 7), np.float32)
    parts[:, 0] = np.arange(M)
    parts[:, 6] = w
    for u_frac in (0.0, 0.37, 0.999):
        u0 = np.float32(u_frac / M)
        if u0 >= np.float32(1.0 / M):
            u0 = np.nextafter(np.float32(1.0 / M), np.float32(0))
        new, idx = tsl.resample_systematic(parts, u0)
        ref = ob.resample(w, u0)
        assert np.array_equal(idx, ref), f"{kind} M={M} u0={u0}: {idx.size} vs {ref.size}"
        assert np.array_equal(new[:, 0], ref.astype(np.float32))


def test_resample_large_m_count_differs_from_m(tsl):
    """At M = 1e6 the float U recurrence drifts: the reference emits fewer
    than M particles; the device must emit exactly the same set."""
    M = 1_000_000
    rng = np.random.default_rng(3)
    w = _weights("random", M, rng)
    parts = np.zeros((M, 7), np.float32)
    parts[:, 6] = w
    u0 = np.float32(0.5 / M)
    _, idx = tsl.resample_systematic(parts, u0)
    ref = ob.resample(w, u0)
    assert ref.size != M
    assert np.array_equal(idx, ref)


def test_device_filter_update_and_resample(tsl, room, room_oracle):
    ev = tsl.TSDFEvaluator(room.map)
    f = tsl.DeviceFilter(ev, room.particles.shape[0])
    f.set_particles(room.particles)
    f.set_scan(room.points, room.tf)
    s = tsl.Stream()
    f.update_async(s)
    pose, n = f.sync(s)
    got = f.get_particles()
    _, w64 = room_oracle.evaluate(room.particles, room.points, room.tf)
    assert np.max(rel_err(got[:, 6], w64 / w64.sum())) <= WEIGHT_RTOL
    u0 = np.float32(0.25 / n)
    f.resample_async(u0, s)
    _, n2 = f.sync(s)
    after = f.get_particles()
    ref_idx = ob.resample(got[:, 6], u0)
    assert n2 == ref_idx.size
    assert np.array_equal(after, got[ref_idx])


def test_general_tier_huge_and_nan(tsl, room, room_oracle):
    """Offsets beyond 2^23 m, infinities and NaN take the general tier, which
    spells out every x86-64 conversion; indices must still match."""
    rng = np.random.default_rng(11)
    parts = room.particles[:64].copy()
    parts[0, 0] = 1e8
    parts[1, 1] = -1e8
    parts[2, 2] = np.inf
    parts[3, 0] = np.nan
    parts[4, 5] = np.nan
    parts[5, 0] = 3e38
    parts[6, 0] = -9.3e18
    parts[7, 0] = 1.9e19
    parts[8:40, :3] += rng.uniform(-4e6, 4e6, (32, 3)).astype(np.float32)
    pts = np.concatenate([room.points[:300], rng.uniform(-1e7, 1e7, (40, 3)).astype(np.float32)])
    ev = tsl.TSDFEvaluator(room.map)
    gpu = ev.trace_indices(parts, pts, room.tf)
    ref = room_oracle.trace(parts, pts, room.tf)
    assert np.array_equal(gpu, ref), f"{np.count_nonzero(gpu != ref)} mismatches"


def oracle_values(om, idx):
    """getEntry (cuda_sub_voxel_map.tcc:140-158) from the oracle's indices."""
    ds = om.coef.data_size
    vals = np.full(idx.shape, np.float32(om.coef.init_value), np.float32)
    ok = idx < ds
    vals[ok] = om.data[idx[ok].astype(np.int64)]
    return vals


def test_room_values_bit_exact(tsl, room, room_oracle):
    """Bricked device layout (room map: 20^3 blocks, aligned offsets)."""
    ev = tsl.TSDFEvaluator(room.map)
    got = ev.trace_values(room.particles[:300], room.points, room.tf)
    ref = oracle_values(room_oracle, room_oracle.trace(room.particles[:300], room.points, room.tf))
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_values_linear_layout_odd_sub_dim(tsl):
    """resolution 0.03 -> 34^3 blocks: codes stay in the reference's linear layout."""
    from tsdf_localization_amd import synth
    sc = synth.Scene()
    sc.add_box((-5, -5, -5), (20, 20, 0.3))
    sc.add_box((-5, -5, -5), (0.35, 20, 20))
    sc.add_cylinder(1.7, 1.5, 0.0, 2.0, 0.3)
    m = sc.tsdf_map((0, 0, 0), (3.2, 3.2, 2.4), 0.03, 150, dense=False)
    assert m.coef().sub_dim == 34
    rng = np.random.default_rng(5)
    P = 64
    parts = np.zeros((P, 7), np.float32)
    parts[:, :3] = rng.uniform([0.5, 0.5, 0.5], [2.5, 2.5, 1.5], (P, 3))
    parts[:, 5] = rng.uniform(-math.pi, math.pi, P)
    pts = rng.uniform(-2.5, 2.5, (700, 3)).astype(np.float32)
    tf = synth.tf_from_xyzrpy(0, 0, 0, 0, 0, 0)
    ev = tsl.TSDFEvaluator(m)
    om = ob.OracleMap.from_host_map(m)
    idx_ref = om.trace(parts, pts, tf)
    assert np.array_equal(ev.trace_indices(parts, pts, tf), idx_ref)
    got = ev.trace_values(parts, pts, tf)
    assert np.array_equal(got.view(np.uint32), oracle_values(om, idx_ref).view(np.uint32))


def test_values_unaligned_block_offsets(tsl, room):
    """A block table whose offsets are not multiples of sub_dim^3 (legal in the
    reference's representation) keeps the linear layout and exact values."""
    c = room.map.coef()
    grid = room.map.grid_occ.copy()
    data = ob.float_data(room.map).copy()
    sd3 = c.sub_dim ** 3
    # shift every block by 7 voxels into a longer array
    grid = np.where(grid >= 0, grid + 7, grid).astype(np.int32)
    data = np.concatenate([np.arange(7, dtype=np.float32) + 100.0, data]).astype(np.float32)
    c2 = type(c)()
    for name, _ in type(c)._fields_:
        setattr(c2, name, getattr(c, name))
    c2.data_size = data.size
    m = tsl.CudaSubVoxelMap(c2, grid, data=data)
    assert data.size % sd3 != 0
    om = ob.OracleMap(c2, grid, data)
    ev = tsl.TSDFEvaluator(m)
    parts = room.particles[:100]
    idx_ref = om.trace(parts, room.points, room.tf)
    assert np.array_equal(ev.trace_indices(parts, room.points, room.tf), idx_ref)
    got = ev.trace_values(parts, room.points, room.tf)
    assert np.array_equal(got.view(np.uint32), oracle_values(om, idx_ref).view(np.uint32))


def test_values_wrap_and_general_tiers(tsl, room, room_oracle):
    rng = np.random.default_rng(21)
    parts = room.particles[:96].copy()
    parts[:32, :3] -= rng.uniform(1.0, 14.0, (32, 3)).astype(np.float32)   # wrap band
    parts[32:40, 0] = rng.uniform(-3e7, 3e7, 8).astype(np.float32)          # general tier
    pts = room.points
    ev = tsl.TSDFEvaluator(room.map)
    idx_ref = room_oracle.trace(parts, pts, room.tf)
    assert np.array_equal(ev.trace_indices(parts, pts, room.tf), idx_ref)
    got = ev.trace_values(parts, pts, room.tf)
    assert np.array_equal(got.view(np.uint32), oracle_values(room_oracle, idx_ref).view(np.uint32))
