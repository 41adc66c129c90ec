"""The 64-bit map form on the device (tsl_map_create_tsdf_wide): a map in
that form gives the same bits as in the reference's int32 form, and a
full-size campus (~7.3e9 voxels, ~14.6 GB, block offsets past 2^31) gives
the oracle's voxel indices bit for bit and weights within 1e-5."""
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


def _wide(tsl, m):
    return tsl.CudaSubVoxelMap(m.coef(), m.grid_occ.astype(np.int64), tsdf_mm=m.tsdf_mm,
                               truncation_mm=m.truncation_mm, sigma=m.sigma)


def test_wide_form_equals_reference_form(tsl):
    from tsdf_localization_amd import synth
    room = synth.room()
    w = synth.building(num_particles=4000, target_points=4000)
    for work in (room, w):
        wide = _wide(tsl, work.map)
        ev32 = tsl.TSDFEvaluator(work.map)
        ev64 = tsl.TSDFEvaluator(wide)
        om64 = ob.OracleMap64.from_host_map(wide)
        parts, pts = work.particles, work.points
        print(f"wide form of {work.name}: P={parts.shape[0]} N={pts.shape[0]}", flush=True)
        idx = ev64.trace_indices(parts[:64], pts, work.tf)
        assert np.array_equal(idx, ev32.trace_indices(parts[:64], pts, work.tf))
        assert np.array_equal(idx, om64.trace(parts[:64], pts, work.tf))
        assert np.array_equal(ev64.trace_values(parts[:64], pts, work.tf).view(np.uint32),
                              ev32.trace_values(parts[:64], pts, work.tf).view(np.uint32))
        raw = ev64.evaluate_raw(parts, pts, work.tf)
        assert np.array_equal(raw.view(np.uint32), ev32.evaluate_raw(parts, pts, work.tf).view(np.uint32))
        _, w64 = om64.evaluate(parts, pts, work.tf)
        assert np.max(np.abs(raw.astype(np.float64) - w64) / np.abs(w64)) <= WEIGHT_RTOL
        # a resident filter through update and resampling, and a batch member
        s = tsl.Stream()
        outs = []
        for ev in (ev32, ev64):
            f = tsl.DeviceFilter(ev, parts.shape[0])
            f.set_particles(parts)
            f.set_scan(pts, work.tf)
            f.update_async(s)
            f.resample_async(np.float32(0.37 / parts.shape[0]), s)
            f.sync(s)
            outs.append(f.get_particles())
        assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


def test_full_size_campus(tsl):
    """A 12 x 11 campus (the mixed workload's map size): 64 particles
    spread over the campus (most blocks past offset 2^31) traced against the
    oracle's 64-bit lookup, and 20,000 particles' weights (x 4,000 points)
    against the oracle, read through int16 millimetres."""
    from tsdf_localization_amd import synth
    cp = synth.campus(0x5EED0005, tiles=(12, 11))
    c = cp.map.coef()
    assert c.data_size > 2 ** 31 and cp.map.grid_occ.max() > 2 ** 31
    rng = np.random.default_rng(3)
    tf = synth.tf_from_xyzrpy(**synth.SCANNER_TF)
    _, pts = synth.campus_scan(cp, rng, tf, 4000)
    parts = synth.campus_particles(cp, 17, 20000, np.radians(3.0))
    ev = tsl.TSDFEvaluator(cp.map)
    print(f"campus: {c.data_size} voxels, {cp.map.device_map().device_bytes() / 1e9:.2f} GB on the device", flush=True)
    om64 = ob.OracleMap64.from_host_map(cp.map)
    idx = ev.trace_indices(parts[:64], pts, tf)
    ref = om64.trace(parts[:64], pts, tf)
    assert np.array_equal(idx, ref)
    inside = ref[ref < c.data_size]
    assert inside.size > 1000 and np.count_nonzero(inside > 2 ** 31) > inside.size // 2
    raw = ev.evaluate_raw(parts, pts, tf)
    _, w64 = om64.evaluate(parts, pts, tf)
    assert np.max(np.abs(raw.astype(np.float64) - w64) / np.abs(w64)) <= WEIGHT_RTOL
