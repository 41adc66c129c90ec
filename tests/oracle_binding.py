"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP path against it, and pin it
against closed forms, brute force and the native behaviour it restates.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

if not os.path.exists(ORACLE_PATH):
    raise ImportError(f"{ORACLE_PATH} is not built (make -C oracle)")

olib = ctypes.CDLL(ORACLE_PATH)

# sin/cos of the particle angles: the C library's sinf/cosf (the reference's
# unqualified sin(float)), or the double-evaluated routine
TRIG_LIBM = 0
TRIG_DOUBLE = 2


def trig_mode_for(library_mode):
    """Oracle trig mode that restates a library TSL_TRIG_* mode: the glibc
    modes are the C library itself (the variant this process's glibc chose)."""
    return TRIG_DOUBLE if int(library_mode) == 2 else TRIG_LIBM


class OrcCoef(ctypes.Structure):
    _fields_ = [
        ("dim_x", ctypes.c_uint64), ("dim_y", ctypes.c_uint64), ("dim_z", ctypes.c_uint64),
        ("min_x", ctypes.c_float), ("min_y", ctypes.c_float), ("min_z", ctypes.c_float),
        ("max_x", ctypes.c_float), ("max_y", ctypes.c_float), ("max_z", ctypes.c_float),
        ("resolution", ctypes.c_float), ("init_value", ctypes.c_float),
        ("up_dim_x", ctypes.c_uint64), ("up_dim_y", ctypes.c_uint64), ("up_dim_z", ctypes.c_uint64),
        ("up_dim_2", ctypes.c_uint64), ("sub_dim", ctypes.c_uint64), ("sub_dim_2", ctypes.c_uint64),
        ("grid_occ_size", ctypes.c_uint64), ("data_size", ctypes.c_uint64),
    ]


class OrcParams(ctypes.Structure):
    _fields_ = [("a_hit", ctypes.c_float), ("a_range", ctypes.c_float), ("a_max", ctypes.c_float),
                ("max_range", ctypes.c_float)]


_P = ctypes.c_void_p
_F = ctypes.c_float
_S = ctypes.c_size_t
_sigs = {
    "orc_cvt_f32_u64": (ctypes.c_uint64, [_F]),
    "orc_native_cvt_f32_u64": (ctypes.c_uint64, [_F]),
    "orc_cvt_u64_f32": (_F, [ctypes.c_uint64]),
    "orc_sincosf": (None, [_F, ctypes.POINTER(_F), ctypes.POINTER(_F)]),
    "orc_coef_size": (_S, []),
    "orc_coef_init": (None, [_F, _F, _F, _F, _F, _F, _F, _F, ctypes.POINTER(OrcCoef)]),
    "orc_get_index": (ctypes.c_uint64, [ctypes.POINTER(OrcCoef), _P, _F, _F, _F]),
    "orc_get_entry": (_F, [ctypes.POINTER(OrcCoef), _P, _P, _F, _F, _F]),
    "orc_set_data": (ctypes.c_int, [ctypes.POINTER(OrcCoef), _P, _P, _P, _S, ctypes.POINTER(_P)]),
    "orc_free": (None, [_P]),
    "orc_likelihood_init": (_F, [_F]),
    "orc_likelihood_value": (_F, [ctypes.c_int16, _F]),
    "orc_chunks_to_cells": (None, [_P, _P, _S, _F, ctypes.POINTER(OrcCoef), _P, _P, ctypes.POINTER(_S), _P,
                                   ctypes.POINTER(_S)]),
    "orc_pose_matrix": (None, [_P, _P, ctypes.c_int, _P]),
    "orc_evaluate": (None, [ctypes.POINTER(OrcCoef), _P, _P, _P, _S, _P, _S, _P, ctypes.POINTER(OrcParams),
                            ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P]),
    "orc_trace_indices": (None, [ctypes.POINTER(OrcCoef), _P, _P, _S, _P, _S, _P, ctypes.c_int, _P]),
    "orc_normalize": (ctypes.c_int, [_P, _S, _P, ctypes.c_int, ctypes.c_int, _P]),
    "orc_resample_systematic": (ctypes.c_long, [_P, _S, _F, _P, _S]),
    "orc_systematic_u": (None, [_S, _F, _P, _S]),
    "orc_cumsum": (None, [_P, _S, _P]),
    "orc_motion_normal": (ctypes.c_double, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32]),
    "orc_motion": (None, [_P, _S, _P, _P, ctypes.c_uint64, ctypes.c_int, _P]),
    "orc_reduce_cloud": (ctypes.c_long, [_P, _P, _S, _P]),
    "orc_get_index64": (ctypes.c_uint64, [ctypes.POINTER(OrcCoef), _P, _F, _F, _F]),
    "orc_evaluate64": (None, [ctypes.POINTER(OrcCoef), _P, _P, _P, ctypes.c_int, _P, _S, _P, _S, _P,
                              ctypes.POINTER(OrcParams), ctypes.c_int, ctypes.c_int, _P, _P]),
    "orc_trace_indices64": (None, [ctypes.POINTER(OrcCoef), _P, _P, _S, _P, _S, _P, ctypes.c_int, _P]),
    "orc_trig_binding_diff": (None, [ctypes.POINTER(OrcCoef), _P, _P, _P, _S, _P, _S, _P, ctypes.POINTER(OrcParams),
                                     ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double)]),
    "orc_reduce_scan": (ctypes.c_long, [_P, _P, _S, _F, _P]),
}
for _n, (_r, _a) in _sigs.items():
    _f = getattr(olib, _n)
    _f.restype = _r
    _f.argtypes = _a

assert olib.orc_coef_size() == 120


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def coef_from(c):
    """OrcCoef from any object with MapCoef's fields."""
    o = OrcCoef()
    for name, _ in OrcCoef._fields_:
        setattr(o, name, getattr(c, name))
    return o


def coef_init(min_xyz, max_xyz, res, init=0.0):
    c = OrcCoef()
    olib.orc_coef_init(*map(float, min_xyz), *map(float, max_xyz), float(res), float(init), ctypes.byref(c))
    return c


def sincosf(x):
    s, c = ctypes.c_float(), ctypes.c_float()
    olib.orc_sincosf(float(x), ctypes.byref(s), ctypes.byref(c))
    return np.float32(s.value), np.float32(c.value)


def likelihood_lut(truncation_mm, sigma=0.1):
    """float value for every mm in (-trunc, trunc), index mm + trunc - 1."""
    t = truncation_mm
    return np.array([olib.orc_likelihood_value(mm, sigma) for mm in range(-(t - 1), t)], np.float32)


def float_data(host_map):
    """The reference's float voxel array for a map given as int16 mm
    (map_util.h:116-129 values, init elsewhere), computed by the oracle."""
    if host_map.data is not None:
        return host_map.data
    t = host_map.truncation_mm
    lut = likelihood_lut(t, host_map.sigma)
    mm = host_map.tsdf_mm.astype(np.int32)
    valid = (mm > -t) & (mm < t) & (mm != -32768)
    data = np.full(mm.shape, np.float32(host_map.coef().init_value), np.float32)
    data[valid] = lut[mm[valid] + t - 1]
    return data


def pose_matrix(pose6, tf16, trig_mode=0):
    p = np.ascontiguousarray(pose6, np.float32)
    tf = np.ascontiguousarray(tf16, np.float32)
    out = np.empty(12, np.float32)
    olib.orc_pose_matrix(_ptr(p), _ptr(tf), trig_mode, _ptr(out))
    return out


class OracleMap:
    def __init__(self, coef, grid_occ, data):
        self.coef = coef_from(coef)
        self.grid = np.ascontiguousarray(grid_occ, np.int32)
        self.data = np.ascontiguousarray(data, np.float32)

    @classmethod
    def from_host_map(cls, m):
        return cls(m.coef(), m.grid_occ, float_data(m))

    def get_index(self, x, y, z):
        return olib.orc_get_index(ctypes.byref(self.coef), _ptr(self.grid), float(x), float(y), float(z))

    def get_entry(self, x, y, z):
        return np.float32(olib.orc_get_entry(ctypes.byref(self.coef), _ptr(self.grid), _ptr(self.data),
                                             float(x), float(y), float(z)))

    def evaluate(self, particles, points, tf, a_hit=0.9, a_range=0.1, a_max=0.0, max_range=100.0, trig_mode=0,
                 threads=0, faithful_copy=0):
        p = np.ascontiguousarray(particles, np.float32)
        pts = np.ascontiguousarray(points, np.float32)
        tf = np.ascontiguousarray(tf, np.float32)
        prm = OrcParams(a_hit, a_range, a_max, max_range)
        w32 = np.empty(p.shape[0], np.float32)
        w64 = np.empty(p.shape[0], np.float64)
        olib.orc_evaluate(ctypes.byref(self.coef), _ptr(self.grid), _ptr(self.data), _ptr(p), p.shape[0],
                          _ptr(pts), pts.shape[0], _ptr(tf), ctypes.byref(prm), trig_mode, threads, faithful_copy,
                          _ptr(w32), _ptr(w64))
        return w32, w64

    def trig_binding_diff(self, particles, points, tf, mode_a, mode_b, a_hit=0.9, a_range=0.1, a_max=0.0,
                          max_range=100.0):
        """(index mismatches, pose-matrix mismatches, max relative weight
        difference) between two trig modes on the same inputs."""
        p = np.ascontiguousarray(particles, np.float32)
        pts = np.ascontiguousarray(points, np.float32)
        tf = np.ascontiguousarray(tf, np.float32)
        prm = OrcParams(a_hit, a_range, a_max, max_range)
        bad = ctypes.c_uint64()
        badm = ctypes.c_uint64()
        worst = ctypes.c_double()
        olib.orc_trig_binding_diff(ctypes.byref(self.coef), _ptr(self.grid), _ptr(self.data), _ptr(p), p.shape[0],
                                   _ptr(pts), pts.shape[0], _ptr(tf), ctypes.byref(prm), int(mode_a), int(mode_b),
                                   ctypes.byref(bad), ctypes.byref(badm), ctypes.byref(worst))
        return bad.value, badm.value, worst.value

    def trace(self, particles, points, tf, trig_mode=0):
        p = np.ascontiguousarray(particles, np.float32)
        pts = np.ascontiguousarray(points, np.float32)
        tf = np.ascontiguousarray(tf, np.float32)
        out = np.empty((p.shape[0], pts.shape[0]), np.uint64)
        olib.orc_trace_indices(ctypes.byref(self.coef), _ptr(self.grid), _ptr(p), p.shape[0], _ptr(pts),
                               pts.shape[0], _ptr(tf), trig_mode, _ptr(out))
        return out


class OracleMap64:
    """A map in the 64-bit form (int64 block table, int16 TSDF millimetres
    read through the likelihood values), without a float copy of the voxels."""

    def __init__(self, coef, grid_occ64, tsdf_mm, truncation_mm, sigma=0.1):
        self.coef = coef_from(coef)
        self.grid = np.ascontiguousarray(grid_occ64, np.int64)
        self.mm = np.ascontiguousarray(tsdf_mm, np.int16)
        self.t = int(truncation_mm)
        self.lut = likelihood_lut(self.t, sigma)

    @classmethod
    def from_host_map(cls, m):
        return cls(m.coef(), m.grid_occ, m.tsdf_mm, m.truncation_mm, m.sigma)

    def get_index(self, x, y, z):
        return olib.orc_get_index64(ctypes.byref(self.coef), _ptr(self.grid), float(x), float(y), float(z))

    def evaluate(self, particles, points, tf, a_hit=0.9, a_range=0.1, a_max=0.0, max_range=100.0, trig_mode=0,
                 threads=0):
        p = np.ascontiguousarray(particles, np.float32)
        pts = np.ascontiguousarray(points, np.float32)
        tf = np.ascontiguousarray(tf, np.float32)
        prm = OrcParams(a_hit, a_range, a_max, max_range)
        w32 = np.empty(p.shape[0], np.float32)
        w64 = np.empty(p.shape[0], np.float64)
        olib.orc_evaluate64(ctypes.byref(self.coef), _ptr(self.grid), _ptr(self.mm), _ptr(self.lut), self.t,
                            _ptr(p), p.shape[0], _ptr(pts), pts.shape[0], _ptr(tf), ctypes.byref(prm), trig_mode,
                            threads, _ptr(w32), _ptr(w64))
        return w32, w64

    def trace(self, particles, points, tf, trig_mode=0):
        p = np.ascontiguousarray(particles, np.float32)
        pts = np.ascontiguousarray(points, np.float32)
        tf = np.ascontiguousarray(tf, np.float32)
        out = np.empty((p.shape[0], pts.shape[0]), np.uint64)
        olib.orc_trace_indices64(ctypes.byref(self.coef), _ptr(self.grid), _ptr(p), p.shape[0], _ptr(pts),
                                 pts.shape[0], _ptr(tf), trig_mode, _ptr(out))
        return out


def chunks_to_cells(positions, raw, sigma=0.1):
    """createTSDFMap's cells: (coef, cells (C,3), values (C,), free (F,3))."""
    pos = np.ascontiguousarray(positions, np.int32).reshape(-1, 3)
    words = np.ascontiguousarray(raw, np.uint32).reshape(pos.shape[0], 64 ** 3)
    cap = words.size
    cells = np.empty((cap, 3), np.float32)
    vals = np.empty(cap, np.float32)
    free = np.empty((cap, 3), np.float32)
    c = OrcCoef()
    nc, nf = ctypes.c_size_t(0), ctypes.c_size_t(0)
    olib.orc_chunks_to_cells(_ptr(pos), _ptr(words), pos.shape[0], float(sigma), ctypes.byref(c), _ptr(cells),
                             _ptr(vals), ctypes.byref(nc), _ptr(free), ctypes.byref(nf))
    return c, cells[: nc.value].copy(), vals[: nc.value].copy(), free[: nf.value].copy()


def set_data(min_xyz, max_xyz, res, init, cells_xyz, values):
    c = coef_init(min_xyz, max_xyz, res, init)
    grid = np.empty(c.grid_occ_size, np.int32)
    cells = np.ascontiguousarray(cells_xyz, np.float32)
    vals = np.ascontiguousarray(values, np.float32)
    out = ctypes.c_void_p()
    rc = olib.orc_set_data(ctypes.byref(c), _ptr(grid), _ptr(cells), _ptr(vals), cells.shape[0], ctypes.byref(out))
    if rc == -2:
        raise OverflowError("block offsets exceed int32")
    if rc != 0:
        raise RuntimeError("Upper voxel index overflow!")
    n = int(c.data_size)
    if out.value:
        data = np.ctypeslib.as_array((ctypes.c_float * max(n, 1)).from_address(out.value))[:n].copy()
        olib.orc_free(out)
    else:
        data = np.zeros(0, np.float32)
    return c, grid, data


def normalize(particles, weights, sum_mode=1, trig_mode=0):
    p = np.ascontiguousarray(particles, np.float32)
    w = np.ascontiguousarray(weights, np.float32).copy()
    pose = np.empty(6, np.float64)
    rc = olib.orc_normalize(_ptr(p), p.shape[0], _ptr(w), sum_mode, trig_mode, _ptr(pose))
    if rc != 0:
        raise RuntimeError("No particle is valid!")
    return w, pose


def resample(weights, u0, capacity=None):
    w = np.ascontiguousarray(weights, np.float32)
    M = w.shape[0]
    cap = capacity or (2 * M + 1024)
    out = np.empty(cap, np.int32)
    n = olib.orc_resample_systematic(_ptr(w), M, float(u0), _ptr(out), cap)
    if n < 0:
        raise RuntimeError("capacity exceeded")
    return out[:n].copy()


def systematic_u(M, u0, n):
    out = np.empty(n, np.float32)
    olib.orc_systematic_u(M, float(u0), _ptr(out), n)
    return out


def cumsum(weights):
    w = np.ascontiguousarray(weights, np.float32)
    out = np.empty(w.shape[0], np.float64)
    olib.orc_cumsum(_ptr(w), w.shape[0], _ptr(out))
    return out


def motion(particles, mean, stddev, seed, trig_mode=0):
    """ParticleCloud::apply_model with the counter-based draws."""
    p = np.ascontiguousarray(particles, np.float32)
    out = np.empty_like(p)
    m = np.ascontiguousarray(mean, np.float64)
    sd = np.ascontiguousarray(stddev, np.float64)
    olib.orc_motion(_ptr(p), p.shape[0], _ptr(m), _ptr(sd), int(seed), int(trig_mode), _ptr(out))
    return out


def reduce_cloud(points, rings):
    """Point reduction of CudaEvaluator::evaluate(PointCloud2): distinct
    0.064 m voxel centres, ordered by (ring, index) of their first point."""
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    r = np.ascontiguousarray(rings, np.int32).reshape(-1)
    out = np.empty((max(p.shape[0], 1), 3), np.float32)
    n = olib.orc_reduce_cloud(_ptr(p), _ptr(r), p.shape[0], _ptr(out))
    return out[:n].copy()


def reduce_scan(points, rings, cell=0.064):
    """Scan reduction of TSDFEvaluator::evaluateParticles; None if a ring is
    out of range."""
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    r = np.ascontiguousarray(rings, np.int16).reshape(-1)
    out = np.empty((max(p.shape[0], 1), 3), np.float32)
    n = olib.orc_reduce_scan(_ptr(p), _ptr(r), p.shape[0], float(cell), _ptr(out))
    if n < 0:
        return None
    return out[:n].copy()

