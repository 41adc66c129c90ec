"""Python mirror of the reference's evaluator / resampler interface.

Names, argument meaning and error behaviour follow uos/tsdf_localization:
  CudaSubVoxelMap   include/tsdf_localization/cuda/cuda_sub_voxel_map.h:17-342
  TSDFEvaluator     include/tsdf_localization/evaluation/tsdf_evaluator.h:38-115
  ParticleCloud     include/tsdf_localization/particle_cloud.h:46-259 (container part)
  SystematicResampler include/tsdf_localization/resampling/novel_resampling.h:38-73
Every computation runs in libtsl.so on the GPU; there is no CPU path.

Particles are float32 arrays of shape (P, 7): x, y, z, roll, pitch, yaw,
weight (the byte layout of the reference's Particle).  Points are float32
arrays of shape (N, 3) (CudaPoint).  tf matrices are 16 floats, row-major.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import TRIG_DOUBLE, TRIG_GLIBC_FMA, TRIG_GLIBC_NOFMA, MapCoef, Motion, Pose, check, lib, ptr  # noqa: F401


def _particles(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != 7:
        raise ValueError("particles must have shape (P, 7)")
    return a


def _points(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError("points must have shape (N, 3)")
    return a


def _tf(tf):
    tf = np.ascontiguousarray(tf, dtype=np.float32).reshape(-1)
    if tf.size == 12:
        tf = np.concatenate([tf, np.array([0, 0, 0, 1], np.float32)])
    if tf.size != 16:
        raise ValueError("tf_matrix must hold 16 (or 12) floats")
    return tf


def coef_init(min_xyz, max_xyz, resolution, init_value=0.0):
    """CudaSubVoxelMap constructor arithmetic (cuda_sub_voxel_map.tcc:5-45)."""
    c = MapCoef()
    check(lib.tsl_map_coef_init(*[float(v) for v in min_xyz], *[float(v) for v in max_xyz],
                                float(resolution), float(init_value), ctypes.byref(c)))
    return c


def likelihood_value(tsdf_mm, sigma=0.1):
    return lib.tsl_likelihood_value(int(tsdf_mm), float(sigma))


def likelihood_init(sigma=0.1):
    return lib.tsl_likelihood_init(float(sigma))


class CudaSubVoxelMap:
    """Host-side map in the reference's representation.

    Holds the MapCoef, the int32 block table and either one float per voxel
    (the reference's data_) or int16 TSDF millimetres plus the likelihood
    parameters, from which the device map is built.  A block table given as
    int64 is the 64-bit form (maps past 2^31 voxels, int16 millimetres only,
    tsl_map_create_tsdf_wide).
    """

    def __init__(self, coef, grid_occ, data=None, tsdf_mm=None, truncation_mm=None, sigma=0.1):
        self._coef = coef
        self.wide = np.asarray(grid_occ).dtype == np.int64
        self.grid_occ = np.ascontiguousarray(grid_occ, dtype=np.int64 if self.wide else np.int32)
        if self.wide and data is not None:
            raise ValueError("the 64-bit map form takes int16 TSDF millimetres")
        self.data = None if data is None else np.ascontiguousarray(data, dtype=np.float32)
        self.tsdf_mm = None if tsdf_mm is None else np.ascontiguousarray(tsdf_mm, dtype=np.int16)
        self.truncation_mm = truncation_mm
        self.sigma = sigma
        if self.grid_occ.size != coef.grid_occ_size:
            raise ValueError("grid_occ size does not match coef.grid_occ_size")
        n = self.data.size if self.data is not None else self.tsdf_mm.size
        if n != coef.data_size:
            raise ValueError("voxel array size does not match coef.data_size")
        self._device = None

    @classmethod
    def from_cells(cls, min_xyz, max_xyz, resolution, init_value, cells_xyz, values):
        """Constructor + setData (cuda_sub_voxel_map.tcc:5-45,171-231)."""
        c = coef_init(min_xyz, max_xyz, resolution, init_value)
        grid = np.empty(c.grid_occ_size, np.int32)
        cells_xyz = np.ascontiguousarray(cells_xyz, np.float32).reshape(-1, 3)
        values = np.ascontiguousarray(values, np.float32).reshape(-1)
        rc = lib.tsl_map_set_data(ctypes.byref(c), ptr(cells_xyz), ptr(values), cells_xyz.shape[0],
                                  ptr(grid), None, 0)
        if rc not in (_lib.TSL_OK, _lib.TSL_ERR_CAPACITY):
            check(rc)
        data = np.empty(max(int(c.data_size), 1), np.float32)
        check(lib.tsl_map_set_data(ctypes.byref(c), ptr(cells_xyz), ptr(values), cells_xyz.shape[0],
                                   ptr(grid), ptr(data), data.size))
        return cls(c, grid, data=data[: c.data_size])

    def coef(self):
        return self._coef

    def rawGridOcc(self):
        return self.grid_occ

    def rawData(self):
        return self.data

    def gridOccSize(self):
        return int(self._coef.grid_occ_size)

    def dataSize(self):
        return int(self._coef.data_size)

    def device_map(self):
        """Upload once (the reference uploads in CudaEvaluator's constructor)."""
        if self._device is None:
            self._device = DeviceMap(self)
        return self._device


class DeviceMap:
    def __init__(self, host_map):
        self.handle = ctypes.c_void_p()
        c = host_map.coef()
        if host_map.data is not None:
            check(lib.tsl_map_create(ctypes.byref(c), ptr(host_map.grid_occ), ptr(host_map.data),
                                     ctypes.byref(self.handle)))
        elif host_map.wide:
            check(lib.tsl_map_create_tsdf_wide(ctypes.byref(c), ptr(host_map.grid_occ), ptr(host_map.tsdf_mm),
                                               int(host_map.truncation_mm), float(host_map.sigma),
                                               ctypes.byref(self.handle)))
        else:
            check(lib.tsl_map_create_tsdf(ctypes.byref(c), ptr(host_map.grid_occ), ptr(host_map.tsdf_mm),
                                          int(host_map.truncation_mm), float(host_map.sigma),
                                          ctypes.byref(self.handle)))

    def device_bytes(self):
        b = ctypes.c_uint64()
        check(lib.tsl_map_device_bytes(self.handle, ctypes.byref(b)))
        return b.value

    def __del__(self):
        if getattr(self, "handle", None):
            lib.tsl_map_destroy(self.handle)
            self.handle = None


def _pose_dict(p):
    return {"position": tuple(p.position), "orientation": tuple(p.orientation),
            "covariance": tuple(p.covariance)}


class TSDFEvaluator:
    """Sensor update with a TSDF map (tsdf_evaluator.h:72-115)."""

    def __init__(self, map_ptr, per_point=False, a_hit=0.9, a_range=0.1, a_max=0.0, max_range=100.0,
                 reduction_cell_size=0.064, trig=TRIG_GLIBC_FMA):
        """trig: sin/cos of the particle angles, TRIG_GLIBC_FMA (default: the
        C library's sinf/cosf the reference CPU evaluator calls, as glibc
        computes them on x86-64 with FMA3), TRIG_GLIBC_NOFMA or TRIG_DOUBLE."""
        self.map = map_ptr
        self._dev_map = map_ptr.device_map()
        self.reduction_cell_size = reduction_cell_size
        self.handle = ctypes.c_void_p()
        check(lib.tsl_evaluator_create(self._dev_map.handle, int(bool(per_point)), float(a_hit), float(a_range),
                                       float(a_max), float(max_range), ctypes.byref(self.handle)))
        check(lib.tsl_evaluator_set_trig(self.handle, int(trig)))
        self.trig = int(trig)

    def evaluate(self, particles, points, tf_matrix, use_cuda=True):
        """TSDFEvaluator::evaluate: writes normalised weights into particles[:, 6]
        in place and returns the pose estimate.  Raises RuntimeError
        ("No particle is valid!") on a zero weight sum.  use_cuda is accepted for
        signature compatibility; this build has one (GPU) code path."""
        if not isinstance(particles, np.ndarray) or particles.dtype != np.float32 or \
                not particles.flags.c_contiguous:
            raise ValueError("particles must be a C-contiguous float32 (P, 7) array (updated in place)")
        _particles(particles)
        pts = _points(points)
        tf = _tf(tf_matrix)
        pose = Pose()
        check(lib.tsl_evaluate(self.handle, ptr(particles), particles.shape[0], ptr(pts), pts.shape[0],
                               ptr(tf), ctypes.byref(pose)))
        return _pose_dict(pose)

    def evaluate_cloud(self, particles, points, rings, tf_matrix):
        """Evaluator::evaluate(particles, PointCloud2, tf) of the reference's
        CUDA evaluator (cuda_evaluator.cu:78-116): the cloud's points reduced
        to distinct 0.064 m voxel centres on the device, then evaluated as
        evaluate(); weights normalised in place, pose returned."""
        if not isinstance(particles, np.ndarray) or particles.dtype != np.float32 or \
                not particles.flags.c_contiguous:
            raise ValueError("particles must be a C-contiguous float32 (P, 7) array (updated in place)")
        _particles(particles)
        pts = _points(points)
        r = np.ascontiguousarray(rings, np.int32).reshape(-1)
        if r.size != pts.shape[0]:
            raise ValueError("one ring number per point")
        tf = _tf(tf_matrix)
        pose = Pose()
        check(lib.tsl_evaluate_cloud(self.handle, ptr(particles), particles.shape[0], ptr(pts), ptr(r),
                                     pts.shape[0], ptr(tf), ctypes.byref(pose)))
        return _pose_dict(pose)

    def evaluate_raw(self, particles, points, tf_matrix):
        """Per-particle raw weights (before normalisation)."""
        p = _particles(particles)
        pts = _points(points)
        tf = _tf(tf_matrix)
        out = np.empty(p.shape[0], np.float32)
        check(lib.tsl_evaluate_raw(self.handle, ptr(p), p.shape[0], ptr(pts), pts.shape[0], ptr(tf), ptr(out)))
        return out

    def trace_indices(self, particles, points, tf_matrix):
        """getIndex result of every (particle, point) pair, shape (P, N), uint64."""
        p = _particles(particles)
        pts = _points(points)
        tf = _tf(tf_matrix)
        out = np.empty((p.shape[0], pts.shape[0]), np.uint64)
        check(lib.tsl_trace_indices(self.handle, ptr(p), p.shape[0], ptr(pts), pts.shape[0], ptr(tf), ptr(out)))
        return out

    def trace_values(self, particles, points, tf_matrix):
        """getEntry result of every (particle, point) pair, shape (P, N), float32."""
        p = _particles(particles)
        pts = _points(points)
        tf = _tf(tf_matrix)
        out = np.empty((p.shape[0], pts.shape[0]), np.float32)
        check(lib.tsl_trace_values(self.handle, ptr(p), p.shape[0], ptr(pts), pts.shape[0], ptr(tf), ptr(out)))
        return out

    def __del__(self):
        if getattr(self, "handle", None):
            lib.tsl_evaluator_destroy(self.handle)
            self.handle = None


def pose_matrices(particles, tf_matrix, trig=TRIG_GLIBC_FMA):
    """Pose matrices L (P, 12) computed on the device (tsdf_evaluator.cpp:102-145)."""
    p = _particles(particles)
    tf = _tf(tf_matrix)
    out = np.empty((p.shape[0], 12), np.float32)
    check(lib.tsl_pose_matrices(ptr(p), p.shape[0], ptr(tf), int(trig), ptr(out)))
    return out


def trig(angles, mode=TRIG_GLIBC_FMA, on_device=True):
    """(sin, cos) of float32 angles in a trig mode, on the device or by the
    host build of the same code."""
    x = np.ascontiguousarray(angles, np.float32).reshape(-1)
    s = np.empty_like(x)
    c = np.empty_like(x)
    check(lib.tsl_trig(int(mode), ptr(x), x.size, ptr(s), ptr(c), int(bool(on_device))))
    return s, c


def check_trig_libm(mode, stride):
    """Host build of a trig mode against this process's sinf/cosf over every
    non-NaN float whose bit pattern is a multiple of stride: (mismatches, checked)."""
    bad = ctypes.c_uint64()
    n = ctypes.c_uint64()
    check(lib.tsl_check_trig_libm(int(mode), int(stride), ctypes.byref(bad), ctypes.byref(n)))
    return bad.value, n.value


def check_invtrig_libm(stride, atan2_samples):
    """Host build of the motion model's asinf and atanf against this
    process's C library over every non-NaN float whose bit pattern is a
    multiple of stride, and of its atan2f on atan2_samples seeded random
    pairs: (mismatches, checked)."""
    bad = ctypes.c_uint64()
    n = ctypes.c_uint64()
    check(lib.tsl_check_invtrig_libm(int(stride), int(atan2_samples), ctypes.byref(bad), ctypes.byref(n)))
    return bad.value, n.value


def resample_systematic(particles, u0):
    """Device systematic resampling; returns (new_particles, source_indices)."""
    p = _particles(particles)
    M = p.shape[0]
    cap = ctypes.c_size_t()
    check(lib.tsl_resample_capacity(M, float(u0), ctypes.byref(cap)))
    out = np.empty((max(cap.value, 1), 7), np.float32)
    idx = np.empty(max(cap.value, 1), np.int32)
    n = ctypes.c_size_t()
    check(lib.tsl_resample_systematic(ptr(p), M, float(u0), ptr(out), out.shape[0], ptr(idx), ctypes.byref(n)))
    return out[: n.value].copy(), idx[: n.value].copy()


class ParticleCloud:
    """Container part of ParticleCloud (particle_cloud.h:46-259)."""

    def __init__(self, particles=None):
        self._p = np.zeros((0, 7), np.float32) if particles is None else _particles(particles).copy()

    def particles(self):
        return self._p

    def set_particles(self, p):
        self._p = _particles(p)

    def size(self):
        return self._p.shape[0]

    def __len__(self):
        return self.size()

    def __getitem__(self, i):
        if i >= self.size():
            raise IndexError("Index exceeds number of particles")
        return self._p[i]


class SystematicResampler:
    """SystematicResampler (novel_resampling.h:38-73) on the device.

    The draw U ~ U[0, 1/M) comes from a seeded numpy generator, or from the
    caller through resample(cloud, u0=...)."""

    def __init__(self, seed=None):
        self._rng = np.random.default_rng(seed)

    def resample(self, particle_cloud, u0=None):
        M = particle_cloud.size()
        if u0 is None:
            u0 = np.float32(self._rng.uniform(0.0, 1.0 / M))
        new, _ = resample_systematic(particle_cloud.particles(), u0)
        particle_cloud.set_particles(new)
        return particle_cloud


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(lib.tsl_stream_create(ctypes.byref(self.handle)))

    def synchronize(self):
        check(lib.tsl_stream_synchronize(self.handle))

    def __del__(self):
        if getattr(self, "handle", None):
            lib.tsl_stream_destroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        check(lib.tsl_event_create(ctypes.byref(self.handle)))

    def record(self, stream):
        check(lib.tsl_event_record(self.handle, stream.handle))

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        check(lib.tsl_event_elapsed_ms(self.handle, end.handle, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        if getattr(self, "handle", None):
            lib.tsl_event_destroy(self.handle)
            self.handle = None


class DeviceFilter:
    """Device-resident particle set + scan bound to an evaluator."""

    def __init__(self, evaluator, capacity):
        self.evaluator = evaluator
        self.handle = ctypes.c_void_p()
        check(lib.tsl_filter_create(evaluator.handle, int(capacity), ctypes.byref(self.handle)))

    def set_particles(self, particles):
        p = _particles(particles)
        check(lib.tsl_filter_set_particles(self.handle, ptr(p), p.shape[0]))

    def get_particles(self, capacity=None):
        n = ctypes.c_size_t()
        if capacity is None:
            lib.tsl_filter_get_particles(self.handle, None, 0, ctypes.byref(n))
            capacity = n.value
        out = np.empty((max(capacity, 1), 7), np.float32)
        check(lib.tsl_filter_get_particles(self.handle, ptr(out), capacity, ctypes.byref(n)))
        return out[: n.value].copy()

    def set_scan(self, points, tf_matrix):
        pts = _points(points)
        tf = _tf(tf_matrix)
        check(lib.tsl_filter_set_scan(self.handle, ptr(pts), pts.shape[0], ptr(tf)))

    def set_scan_raw(self, points, rings, tf_matrix, cell=0.064):
        """Raw scan + ring numbers through the reference's reduction
        (tsdf_evaluator.cpp:304-376; cell = reduction_cell_size, default
        0.064 as TSDFEvaluator's constructor, tsdf_evaluator.h:72)."""
        pts = _points(points)
        r = np.ascontiguousarray(rings, np.int16).reshape(-1)
        if r.size != pts.shape[0]:
            raise ValueError("one ring number per point")
        tf = _tf(tf_matrix)
        check(lib.tsl_filter_set_scan_raw(self.handle, ptr(pts), ptr(r), pts.shape[0], float(cell), ptr(tf)))

    def set_scan_cloud(self, points, rings, tf_matrix):
        """Raw cloud through CudaEvaluator's centre reduction
        (cuda_evaluator.cu:78-116): distinct 0.064 m voxel centres."""
        pts = _points(points)
        r = np.ascontiguousarray(rings, np.int32).reshape(-1)
        if r.size != pts.shape[0]:
            raise ValueError("one ring number per point")
        tf = _tf(tf_matrix)
        check(lib.tsl_filter_set_scan_cloud(self.handle, ptr(pts), ptr(r), pts.shape[0], ptr(tf)))

    def get_scan(self):
        n = ctypes.c_size_t()
        lib.tsl_filter_get_scan(self.handle, None, 0, ctypes.byref(n))
        out = np.empty((max(n.value, 1), 3), np.float32)
        check(lib.tsl_filter_get_scan(self.handle, ptr(out), out.shape[0], ctypes.byref(n)))
        return out[: n.value].copy()

    def update_async(self, stream=None):
        check(lib.tsl_filter_update_async(self.handle, None if stream is None else stream.handle))

    def resample_async(self, u0, stream=None):
        check(lib.tsl_filter_resample_async(self.handle, float(u0), None if stream is None else stream.handle))

    def motion_async(self, motion, stream=None):
        """ParticleCloud::apply_model on the device (motion: a tsl_motion)."""
        check(lib.tsl_filter_motion_async(self.handle, ctypes.byref(motion), _stream(stream)))

    def snapshot(self):
        check(lib.tsl_filter_snapshot(self.handle))

    def restore_async(self, stream=None):
        check(lib.tsl_filter_restore_async(self.handle, None if stream is None else stream.handle))

    def set_tile(self, k=0, chunk_len=0, points_per_lane=0):
        check(lib.tsl_filter_set_tile(self.handle, int(k), int(points_per_lane), int(chunk_len)))

    def set_block_lookup(self, mode=-1):
        """Block offsets: 0 global table, 1 LDS occupancy-rank table, -1 automatic."""
        check(lib.tsl_filter_set_block_lookup(self.handle, int(mode)))

    def set_order(self, mode=-1):
        """Particle evaluation order: 0 caller, 1 spatial, -1 automatic."""
        check(lib.tsl_filter_set_order(self.handle, int(mode)))

    def autotune(self, stream=None, reps=3):
        """Pick the fastest evaluation shape on the current particles and
        scan; returns (order, K, points per lane, chunk_len (0 = automatic),
        ms per update of the evaluation kernel plus the particle order)."""
        o = ctypes.c_int()
        k = ctypes.c_int()
        pp = ctypes.c_int()
        c = ctypes.c_uint32()
        ms = ctypes.c_float()
        check(lib.tsl_filter_autotune(self.handle, None if stream is None else stream.handle, int(reps),
                                      ctypes.byref(o), ctypes.byref(k), ctypes.byref(pp), ctypes.byref(c),
                                      ctypes.byref(ms)))
        return o.value, k.value, pp.value, c.value, ms.value

    def set_timing(self, enable=True):
        check(lib.tsl_filter_set_timing(self.handle, int(bool(enable))))

    def eval_timing(self):
        """(kernel ms, launches, particle*point evaluations) since set_timing."""
        ms = ctypes.c_double()
        n = ctypes.c_uint64()
        pairs = ctypes.c_uint64()
        check(lib.tsl_filter_eval_timing(self.handle, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(pairs)))
        return ms.value, n.value, pairs.value

    def sync(self, stream=None):
        pose = Pose()
        n = ctypes.c_size_t()
        check(lib.tsl_filter_sync(self.handle, None if stream is None else stream.handle, ctypes.byref(pose),
                                  ctypes.byref(n)))
        return _pose_dict(pose), n.value

    def __del__(self):
        if getattr(self, "handle", None):
            lib.tsl_filter_destroy(self.handle)
            self.handle = None


def _stream(stream):
    return None if stream is None else stream.handle


DEFAULT_MOTION_PARAMS = (0.1,) * 12  # a_1 .. a_12 (particle_cloud.h:60-71)


def motion_from_delta(dx, dy, dz, roll, pitch, yaw, lin_base_noise=0.0, ang_base_noise=0.0,
                      a=DEFAULT_MOTION_PARAMS, seed=0):
    """Distributions of ParticleCloud::motion_update (particle_cloud.cpp:456-497)."""
    aa = np.ascontiguousarray(a, np.float32)
    m = Motion()
    check(lib.tsl_motion_from_delta(float(dx), float(dy), float(dz), float(roll), float(pitch), float(yaw),
                                    float(lin_base_noise), float(ang_base_noise), ptr(aa), int(seed), ctypes.byref(m)))
    return m


def motion_from_odom(linear_velocity, angular_velocity, time_diff, a=DEFAULT_MOTION_PARAMS, seed=0):
    """Distributions of ParticleCloud::motionUpdate(odom) (particle_cloud.cpp:153-207)."""
    aa = np.ascontiguousarray(a, np.float32)
    m = Motion()
    check(lib.tsl_motion_from_odom(float(linear_velocity), float(angular_velocity), float(time_diff), ptr(aa),
                                   int(seed), ctypes.byref(m)))
    return m


def motion(mean, stddev, seed=0):
    """A tsl_motion from explicit means and standard deviations."""
    m = Motion()
    for k in range(6):
        m.mean[k] = float(mean[k])
        m.stddev[k] = float(stddev[k])
    m.seed = int(seed)
    return m


class Batch:
    """Many DeviceFilters sharing one evaluator, updated and resampled by one
    fixed sequence of launches (optionally a captured HIP graph)."""

    def __init__(self, evaluator, filters):
        self.evaluator = evaluator
        self.filters = list(filters)
        arr = (ctypes.c_void_p * len(self.filters))(*[f.handle.value for f in self.filters])
        self.handle = ctypes.c_void_p()
        check(lib.tsl_batch_create(evaluator.handle, arr, len(self.filters), ctypes.byref(self.handle)))

    def __len__(self):
        return len(self.filters)

    def _u0(self, u0):
        u = np.ascontiguousarray(u0, np.float32).reshape(-1)
        if u.size != len(self.filters):
            raise ValueError("one draw per filter")
        return u

    def set_tile(self, k=1, chunk_len=0, points_per_lane=1):
        check(lib.tsl_batch_set_tile(self.handle, int(k), int(points_per_lane), int(chunk_len)))

    def autotune(self, stream=None, reps=2):
        """Time the batched update for every tile shape and keep the fastest;
        returns (K, points per lane, chunk_len (0 = automatic), ms per update)."""
        k, p, c, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint32(), ctypes.c_float()
        check(lib.tsl_batch_autotune(self.handle, _stream(stream), int(reps), ctypes.byref(k), ctypes.byref(p),
                                     ctypes.byref(c), ctypes.byref(ms)))
        return k.value, p.value, c.value, ms.value

    def prepare(self):
        check(lib.tsl_batch_prepare(self.handle))

    def restore_async(self, stream=None):
        check(lib.tsl_batch_restore_async(self.handle, _stream(stream)))

    def update_async(self, stream=None):
        check(lib.tsl_batch_update_async(self.handle, _stream(stream)))

    def resample_async(self, u0, stream=None):
        u = self._u0(u0)
        check(lib.tsl_batch_resample_async(self.handle, ptr(u), _stream(stream)))

    def step_async(self, u0, restore=False, stream=None):
        u = self._u0(u0)
        check(lib.tsl_batch_step_async(self.handle, ptr(u), int(bool(restore)), _stream(stream)))

    def capture(self, stream, restore=False):
        check(lib.tsl_batch_capture(self.handle, int(bool(restore)), _stream(stream)))

    def launch(self, u0, stream):
        u = self._u0(u0)
        check(lib.tsl_batch_launch(self.handle, ptr(u), _stream(stream)))

    def motion_async(self, motions, stream=None):
        arr = (Motion * len(self.filters))(*motions)
        check(lib.tsl_batch_motion_async(self.handle, arr, _stream(stream)))

    def sync(self, stream=None, raise_on_error=True):
        """(poses, counts, status) of every member after `stream` drains."""
        n = len(self.filters)
        poses = (Pose * n)()
        counts = np.zeros(n, np.uint64)
        status = np.zeros(n, np.int32)
        rc = lib.tsl_batch_sync(self.handle, _stream(stream), poses, ptr(counts), ptr(status))
        if raise_on_error:
            check(rc)
        return [_pose_dict(p) for p in poses], counts, status

    def __del__(self):
        if getattr(self, "handle", None):
            lib.tsl_batch_destroy(self.handle)
            self.handle = None

