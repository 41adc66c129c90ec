"""The allocation-free entry point (tsl_evaluate_raw_noalloc) against the
entry point it varies, tsl_evaluate_raw: device-resident inputs and output,
every intermediate in the library's fixed per-GPU scratch array.  Calls in
a row with shrinking inputs (the array keeps the previous call's contents),
bit-identical raw weights each time, within 1e-5 of the oracle's float64
sums, the capacity error, and an empty scan."""
import ctypes

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


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libtsl.so uses: the
    test's own device buffers (the entry point under test allocates none)."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so")
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(nbytes, 4))) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        if a.nbytes:
            assert self.rt.hipMemcpy(p, ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(a.nbytes), 1) == 0
        return p

    def download(self, p, n, dtype):
        out = np.empty(n, dtype)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(ctypes.c_void_p(out.ctypes.data), p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        assert self.rt.hipDeviceSynchronize() == 0
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _noalloc(ev, parts, pts, tf, P=None):
    from tsdf_localization_amd._lib import lib
    hip = _Hip()
    try:
        d_p = hip.upload(np.ascontiguousarray(parts, np.float32))
        d_pts = hip.upload(np.ascontiguousarray(pts, np.float32).reshape(-1, 3))
        n = parts.shape[0] if P is None else P
        d_w = hip.upload(np.full(parts.shape[0], -1.0, np.float32))
        tf = np.ascontiguousarray(tf, np.float32)
        rc = lib.tsl_evaluate_raw_noalloc(ev.handle, d_p, n, d_pts if pts.shape[0] else None, pts.shape[0],
                                          ctypes.c_void_p(tf.ctypes.data), d_w, None)
        if rc != 0:
            return rc, None
        return 0, hip.download(d_w, parts.shape[0], np.float32)
    finally:
        hip.free()


def test_noalloc_matches_evaluate_raw(tsl, room):
    ev = tsl.TSDFEvaluator(room.map)
    om = ob.OracleMap.from_host_map(room.map)
    rng = np.random.default_rng(3)
    big = np.tile(room.particles, (20, 1))
    big[:, :3] += rng.normal(0, 0.05, (big.shape[0], 3)).astype(np.float32)
    cases = [(big, room.points), (big[:3000], room.points), (room.particles[:700], room.points[:900])]
    for parts, pts in cases:
        print(f"noalloc case P={parts.shape[0]} N={pts.shape[0]}", flush=True)
        rc, got = _noalloc(ev, parts, pts, room.tf)
        assert rc == 0, tsl._lib.lib.tsl_last_error()
        want = ev.evaluate_raw(parts, pts, room.tf)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        _, w64 = om.evaluate(parts, pts, room.tf)
        assert np.max(np.abs(got.astype(np.float64) - w64) / np.abs(w64)) <= WEIGHT_RTOL


def test_noalloc_empty_scan_and_capacity(tsl, room):
    from tsdf_localization_amd._lib import lib
    ev = tsl.TSDFEvaluator(room.map)
    rc, got = _noalloc(ev, room.particles[:100], np.zeros((0, 3), np.float32), room.tf)
    assert rc == 0 and np.all(got == 0.0)
    # 2,000,000 particles need 96 MB of pose matrices: refused on the host
    # before any launch (the buffers hold all 2,000,000 anyway)
    big = np.zeros((2_000_000, 7), np.float32)
    rc, _ = _noalloc(ev, big, room.points, room.tf)
    assert rc == 5  # TSL_ERR_CAPACITY
    assert b"scratch" in lib.tsl_last_error()
    # the scratch still serves a call after the refusal
    rc, got = _noalloc(ev, room.particles, room.points, room.tf)
    assert rc == 0
    want = ev.evaluate_raw(room.particles, room.points, room.tf)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_noalloc_other_map_forms(tsl, room):
    """The float form (a map with more distinct values than the palette) and
    the 64-bit form (int64 block table): the allocation-free entry equals
    tsl_evaluate_raw bit for bit on each, two calls in a row, the second
    smaller."""
    rng = np.random.default_rng(17)
    data = ob.float_data(room.map)
    data = (data * (np.float32(1.0) + rng.uniform(-1e-3, 1e-3, data.shape).astype(np.float32))).astype(np.float32)
    maps = {
        "float": tsl.CudaSubVoxelMap(room.map.coef(), room.map.grid_occ, data=data),
        "wide": tsl.CudaSubVoxelMap(room.map.coef(), room.map.grid_occ.astype(np.int64), tsdf_mm=room.map.tsdf_mm,
                                    truncation_mm=room.map.truncation_mm, sigma=room.map.sigma),
    }
    for name, m in maps.items():
        ev = tsl.TSDFEvaluator(m)
        for parts, pts in ((room.particles, room.points), (room.particles[:300], room.points[:700])):
            print(f"noalloc {name} map P={parts.shape[0]} N={pts.shape[0]}", flush=True)
            rc, got = _noalloc(ev, parts, pts, room.tf)
            assert rc == 0, tsl._lib.lib.tsl_last_error()
            want = ev.evaluate_raw(parts, pts, room.tf)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), name
