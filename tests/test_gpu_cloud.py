"""The PointCloud2 form of the evaluator (CudaEvaluator::evaluate(particles,
cloud, tf), src/cuda/cuda_evaluator.cu:78-116): the device reduction of a raw
cloud to distinct 0.064 m voxel centres against the oracle, bit for bit and
in the documented order, and the evaluation of a cloud against the oracle's
evaluation of those centres."""
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


def _raw_cloud(seed, room):
    from tsdf_localization_amd import synth
    sc = synth.Scene()
    sc.add_box((-5, -5, -1), (5, 5, 0.0))
    sc.add_box((-5, 4.0, -1), (5, 5, 3))
    sc.add_box((3.5, -5, -1), (5, 5, 3))
    sc.add_cylinder(1.0, 1.5, -1.0, 2.0, 0.3)
    raw, rings = sc.raycast(np.array([0.0, 0.0, 0.8, 0.0, 0.0, 0.2]), 16, 1024, -15.0, 15.0, 30.0, 0.01, seed)
    return raw.astype(np.float32), rings.astype(np.int32)


def test_raycast_cloud_reduction_bit_exact(tsl, room):
    raw, rings = _raw_cloud(3, room)
    ev = tsl.TSDFEvaluator(room.map)
    f = tsl.DeviceFilter(ev, 16)
    f.set_scan_cloud(raw, rings, room.tf)
    got = f.get_scan()
    ref = ob.reduce_cloud(raw, rings)
    assert 100 < ref.shape[0] < raw.shape[0]
    assert got.shape == ref.shape
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_adversarial_clouds(tsl, room):
    """Duplicated centres across rings, negative and large rings, NaN and
    huge coordinates, ragged sizes."""
    ev = tsl.TSDFEvaluator(room.map)
    f = tsl.DeviceFilter(ev, 16)
    rng = np.random.default_rng(21)
    for trial in range(6):
        n = int(rng.integers(1, 5000))
        pts = rng.normal(0, 4, (n, 3)).astype(np.float32)
        pts[: n // 2] = np.round(pts[: n // 2] * 9) / 9
        if trial == 2:
            pts[3] = np.nan
            pts[7, 1] = np.nan
            pts[11] = np.float32(3e38)
        rings = rng.integers(-3 if trial % 2 else 0, 40000 if trial == 3 else 16, n).astype(np.int32)
        if trial == 4:
            rings[:] = -2**31 + rng.integers(0, 3, n)
        f.set_scan_cloud(pts, rings, room.tf)
        got = f.get_scan()
        ref = ob.reduce_cloud(pts, rings)
        assert got.shape == ref.shape, trial
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), trial


def test_evaluate_cloud_matches_oracle(tsl, room):
    raw, rings = _raw_cloud(4, room)
    # the cloud in the room's scanner frame: the room scene ray-cast scan
    # shifted so it overlaps walls of the room map
    raw = raw + np.float32(0.05)
    ev = tsl.TSDFEvaluator(room.map)
    parts = room.particles.copy()
    pose = ev.evaluate_cloud(parts, raw, rings, room.tf)
    centres = ob.reduce_cloud(raw, rings)
    om = ob.OracleMap.from_host_map(room.map)
    _, w64 = om.evaluate(room.particles, centres, room.tf)
    ref = w64 / w64.sum()
    assert np.max(np.abs(parts[:, 6].astype(np.float64) - ref) / ref) <= WEIGHT_RTOL
    _, pref = ob.normalize(room.particles, w64.astype(np.float32))
    assert np.allclose(pose["position"], pref[:3], rtol=0, atol=1e-4)
    # the same as evaluating the centres directly
    parts2 = room.particles.copy()
    ev.evaluate(parts2, centres, room.tf)
    assert np.array_equal(parts[:, 6], parts2[:, 6])


def test_empty_cloud_returns_default_pose(tsl, room):
    ev = tsl.TSDFEvaluator(room.map)
    parts = room.particles.copy()
    pose = ev.evaluate_cloud(parts, np.zeros((0, 3), np.float32), np.zeros(0, np.int32), room.tf)
    assert pose["position"] == (0.0, 0.0, 0.0)
    assert np.array_equal(parts, room.particles)
