"""Device scan reduction (tsl_filter_set_scan_raw) against the oracle, bit
for bit, on ray-cast scans and on adversarial small scans (duplicates,
skipped points that shift the ring iterator, NaN)."""
import numpy as np
import pytest

import oracle_binding as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tsl():
    import tsdf_localization_amd as t
    assert t.device_count() >= 1, "no HIP device"
    return t


@pytest.fixture(scope="module")
def room():
    from tsdf_localization_amd import synth
    return synth.room()


def _raw_scan(seed, beams=16, azimuths=1024):
    from tsdf_localization_amd import synth
    sc = synth.Scene()
    sc.add_box((-5, -5, -1), (5, 5, 0.0))
    sc.add_box((-5, 4.0, -1), (5, 5, 3))
    sc.add_box((3.5, -5, -1), (5, 5, 3))
    sc.add_cylinder(1.0, 1.5, -1.0, 2.0, 0.3)
    pose = np.array([0.0, 0.0, 0.8, 0.0, 0.0, 0.2])
    raw, rings = sc.raycast(pose, beams, azimuths, -15.0, 15.0, 30.0, 0.01, seed)
    return raw.astype(np.float32), rings.astype(np.int16)


@pytest.mark.parametrize("cell", [0.064, 0.05, 0.15])
def test_raycast_scan_reduction_bit_exact(tsl, room, cell):
    raw, rings = _raw_scan(3)
    # some points inside 1 m so that the ring iterator lags
    raw[::37] *= np.float32(0.2)
    ev = tsl.TSDFEvaluator(room.map)
    f = tsl.DeviceFilter(ev, 16)
    f.set_scan_raw(raw, rings, room.tf, cell)
    got = f.get_scan()
    ref = ob.reduce_scan(raw, rings, cell)
    assert ref.shape[0] > 100
    assert got.shape == ref.shape
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_small_adversarial_scans(tsl, room):
    ev = tsl.TSDFEvaluator(room.map)
    f = tsl.DeviceFilter(ev, 16)
    rng = np.random.default_rng(8)
    for trial in range(6):
        n = int(rng.integers(1, 3000))
        pts = rng.normal(0, 3, (n, 3)).astype(np.float32)
        pts[: n // 2] = np.round(pts[: n // 2] * 3) / 3
        if trial == 2:
            pts[5] = np.nan
            pts[9] = np.nan
        rings = rng.integers(0, 64 if trial % 2 else 3, n).astype(np.int16)
        f.set_scan_raw(pts, rings, room.tf, 0.064)
        got = f.get_scan()
        ref = ob.reduce_scan(pts, rings, 0.064)
        assert got.shape == ref.shape, trial
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), trial


def test_ring_out_of_range_raises(tsl, room):
    ev = tsl.TSDFEvaluator(room.map)
    f = tsl.DeviceFilter(ev, 16)
    pts = np.array([[2.0, 0.0, 0.0], [3.0, 0.0, 0.0]], np.float32)
    with pytest.raises(tsl.TslError, match="ring"):
        f.set_scan_raw(pts, np.array([1, 70], np.int16), room.tf, 0.064)


def test_raw_scan_update_matches_reduced_scan(tsl, room):
    """Evaluating after the device reduction equals evaluating the oracle's
    reduced scan through the plain entry point."""
    raw, rings = _raw_scan(5)
    ev = tsl.TSDFEvaluator(room.map)
    a = tsl.DeviceFilter(ev, room.particles.shape[0])
    a.set_particles(room.particles)
    a.set_scan_raw(raw, rings, room.tf, 0.064)
    b = tsl.DeviceFilter(ev, room.particles.shape[0])
    b.set_particles(room.particles)
    b.set_scan(ob.reduce_scan(raw, rings, 0.064), room.tf)
    s = tsl.Stream()
    for f in (a, b):
        f.update_async(s)
        f.sync(s)
    assert np.array_equal(a.get_particles().view(np.uint32), b.get_particles().view(np.uint32))
