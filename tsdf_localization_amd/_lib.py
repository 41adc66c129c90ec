"""ctypes binding of libtsl.so (the C ABI declared in include/tsl.h).

The library is loaded from this package directory only.  There is no
fallback: if libtsl.so is missing or fails to load, importing this module
raises, so nothing can silently run a non-HIP path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtsl.so")

TSL_OK = 0
TSL_ERR_INVALID_ARGUMENT = 1
TSL_ERR_HIP = 2
TSL_ERR_NO_VALID_PARTICLE = 3
TSL_ERR_OUT_OF_MEMORY = 4
TSL_ERR_CAPACITY = 5
TSL_ERR_UNSUPPORTED = 6
TSL_ERR_ADMISSION = 7

# sin/cos of the particle angles (include/tsl.h TSL_TRIG_*)
TRIG_GLIBC_FMA = 0
TRIG_GLIBC_NOFMA = 1
TRIG_DOUBLE = 2


class MapCoef(ctypes.Structure):
    """tsl_map_coef == CudaSubVoxelMap<float,float>::MapCoef (cuda_sub_voxel_map.h:22-50)."""
    _fields_ = [
        ("dim_x", ctypes.c_uint64), ("dim_y", ctypes.c_uint64), ("dim_z", ctypes.c_uint64),
        ("min_x", ctypes.c_float), ("min_y", ctypes.c_float), ("min_z", ctypes.c_float),
        ("max_x", ctypes.c_float), ("max_y", ctypes.c_float), ("max_z", ctypes.c_float),
        ("resolution", ctypes.c_float), ("init_value", ctypes.c_float),
        ("up_dim_x", ctypes.c_uint64), ("up_dim_y", ctypes.c_uint64), ("up_dim_z", ctypes.c_uint64),
        ("up_dim_2", ctypes.c_uint64),
        ("sub_dim", ctypes.c_uint64), ("sub_dim_2", ctypes.c_uint64),
        ("grid_occ_size", ctypes.c_uint64), ("data_size", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Pose(ctypes.Structure):
    """tsl_pose == geometry_msgs::PoseWithCovariance (position, quaternion xyzw, 6x6 covariance)."""
    _fields_ = [("position", ctypes.c_double * 3), ("orientation", ctypes.c_double * 4),
                ("covariance", ctypes.c_double * 36)]


class Motion(ctypes.Structure):
    """tsl_motion: the six normal distributions of a motion step and the noise seed."""
    _fields_ = [("mean", ctypes.c_double * 6), ("stddev", ctypes.c_double * 6), ("seed", ctypes.c_uint64)]


class Request(ctypes.Structure):
    """tsl_request: one serving request."""
    _fields_ = [("id", ctypes.c_uint64), ("robot", ctypes.c_uint32), ("pad", ctypes.c_uint32), ("motion", Motion),
                ("points", ctypes.c_void_p), ("num_points", ctypes.c_size_t), ("u0", ctypes.c_float),
                ("pad2", ctypes.c_float)]


class Result(ctypes.Structure):
    """tsl_result: one finished serving request."""
    _fields_ = [("id", ctypes.c_uint64), ("robot", ctypes.c_uint32), ("status", ctypes.c_int32), ("pose", Pose),
                ("count", ctypes.c_uint64), ("submit_ms", ctypes.c_double), ("start_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double), ("batch", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


assert ctypes.sizeof(MapCoef) == 120

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                      "(or python -m tsdf_localization_amd.build)")

lib = ctypes.CDLL(LIB_PATH)

_P = ctypes.c_void_p
_S = ctypes.c_size_t
_F = ctypes.c_float
_I = ctypes.c_int

_SIGS = {
    "tsl_last_error": (ctypes.c_char_p, []),
    "tsl_version": (_I, []),
    "tsl_device_count": (_I, [ctypes.POINTER(_I)]),
    "tsl_set_device": (_I, [_I]),
    "tsl_map_coef_init": (_I, [_F, _F, _F, _F, _F, _F, _F, _F, ctypes.POINTER(MapCoef)]),
    "tsl_map_create": (_I, [ctypes.POINTER(MapCoef), _P, _P, ctypes.POINTER(_P)]),
    "tsl_map_create_tsdf": (_I, [ctypes.POINTER(MapCoef), _P, _P, _I, _F, ctypes.POINTER(_P)]),
    "tsl_map_create_tsdf_wide": (_I, [ctypes.POINTER(MapCoef), _P, _P, _I, _F, ctypes.POINTER(_P)]),
    "tsl_map_set_data": (_I, [ctypes.POINTER(MapCoef), _P, _P, _S, _P, _P, _S]),
    "tsl_map_chunks_to_cells": (_I, [_P, _P, _S, _F, ctypes.POINTER(MapCoef), _P, _P, _S, ctypes.POINTER(_S),
                                     _P, _S, ctypes.POINTER(_S)]),
    "tsl_likelihood_value": (_F, [ctypes.c_int16, _F]),
    "tsl_likelihood_init": (_F, [_F]),
    "tsl_map_device_bytes": (_I, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "tsl_map_destroy": (_I, [_P]),
    "tsl_evaluator_create": (_I, [_P, _I, _F, _F, _F, _F, ctypes.POINTER(_P)]),
    "tsl_evaluator_destroy": (_I, [_P]),
    "tsl_evaluate": (_I, [_P, _P, _S, _P, _S, _P, ctypes.POINTER(Pose)]),
    "tsl_evaluate_raw": (_I, [_P, _P, _S, _P, _S, _P, _P]),
    "tsl_evaluate_raw_noalloc": (_I, [_P, _P, _S, _P, _S, _P, _P, _P]),
    "tsl_trace_indices": (_I, [_P, _P, _S, _P, _S, _P, _P]),
    "tsl_trace_values": (_I, [_P, _P, _S, _P, _S, _P, _P]),
    "tsl_check_magic_divisor": (_I, [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]),
    "tsl_pose_matrices": (_I, [_P, _S, _P, _I, _P]),
    "tsl_evaluator_set_trig": (_I, [_P, _I]),
    "tsl_evaluator_get_trig": (_I, [_P, ctypes.POINTER(_I)]),
    "tsl_trig": (_I, [_I, _P, _S, _P, _P, _I]),
    "tsl_check_trig_libm": (_I, [_I, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64),
                                 ctypes.POINTER(ctypes.c_uint64)]),
    "tsl_check_invtrig_libm": (_I, [ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                    ctypes.POINTER(ctypes.c_uint64)]),
    "tsl_evaluate_cloud": (_I, [_P, _P, _S, _P, _P, _S, _P, ctypes.POINTER(Pose)]),
    "tsl_filter_set_scan_cloud": (_I, [_P, _P, _P, _S, _P]),
    "tsl_resample_systematic": (_I, [_P, _S, _F, _P, _S, _P, ctypes.POINTER(_S)]),
    "tsl_resample_capacity": (_I, [_S, _F, ctypes.POINTER(_S)]),
    "tsl_filter_create": (_I, [_P, _S, ctypes.POINTER(_P)]),
    "tsl_filter_destroy": (_I, [_P]),
    "tsl_filter_set_particles": (_I, [_P, _P, _S]),
    "tsl_filter_get_particles": (_I, [_P, _P, _S, ctypes.POINTER(_S)]),
    "tsl_filter_set_scan": (_I, [_P, _P, _S, _P]),
    "tsl_filter_update_async": (_I, [_P, _P]),
    "tsl_filter_resample_async": (_I, [_P, _F, _P]),
    "tsl_filter_sync": (_I, [_P, _P, ctypes.POINTER(Pose), ctypes.POINTER(_S)]),
    "tsl_filter_snapshot": (_I, [_P]),
    "tsl_filter_restore_async": (_I, [_P, _P]),
    "tsl_filter_set_tile": (_I, [_P, _I, _I, ctypes.c_uint32]),
    "tsl_filter_autotune": (_I, [_P, _P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I),
                                 ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(_F)]),
    "tsl_filter_set_order": (_I, [_P, _I]),
    "tsl_filter_set_block_lookup": (_I, [_P, _I]),
    "tsl_batch_autotune": (_I, [_P, _P, _I, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_uint32),
                                ctypes.POINTER(_F)]),
    "tsl_server_create": (_I, [_P, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                               ctypes.POINTER(_P)]),
    "tsl_server_destroy": (_I, [_P]),
    "tsl_server_add_robot": (_I, [_P, _P, _S, ctypes.c_uint32, _P, ctypes.POINTER(ctypes.c_uint32)]),
    "tsl_server_remove_robot": (_I, [_P, ctypes.c_uint32]),
    "tsl_server_submit": (_I, [_P, ctypes.POINTER(Request)]),
    "tsl_server_poll": (_I, [_P, _P, _S, ctypes.POINTER(_S), _I]),
    "tsl_server_get_particles": (_I, [_P, ctypes.c_uint32, _P, _S, ctypes.POINTER(_S)]),
    "tsl_server_stats": (_I, [_P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                              ctypes.POINTER(ctypes.c_uint64)]),
    "tsl_server_set_slice": (_I, [_P, ctypes.c_uint64, ctypes.c_uint32]),
    "tsl_server_set_tile": (_I, [_P, _I, _I, _I]),
    "tsl_mcl_write": (_I, [ctypes.c_char_p, _P, _P, _S, _P, _S, _P, _P]),
    "tsl_mcl_read": (_I, [ctypes.c_char_p, _P, _P, _S, _P, _S, _P, _P, ctypes.POINTER(_S), ctypes.POINTER(_S)]),
    "tsl_filter_set_scan_raw": (_I, [_P, _P, _P, _S, _F, _P]),
    "tsl_filter_get_scan": (_I, [_P, _P, _S, ctypes.POINTER(_S)]),
    "tsl_batch_create": (_I, [_P, _P, _S, ctypes.POINTER(_P)]),
    "tsl_batch_destroy": (_I, [_P]),
    "tsl_batch_set_tile": (_I, [_P, _I, _I, ctypes.c_uint32]),
    "tsl_batch_prepare": (_I, [_P]),
    "tsl_batch_restore_async": (_I, [_P, _P]),
    "tsl_batch_update_async": (_I, [_P, _P]),
    "tsl_batch_resample_async": (_I, [_P, _P, _P]),
    "tsl_batch_step_async": (_I, [_P, _P, _I, _P]),
    "tsl_batch_capture": (_I, [_P, _I, _P]),
    "tsl_batch_launch": (_I, [_P, _P, _P]),
    "tsl_batch_sync": (_I, [_P, _P, _P, _P, _P]),
    "tsl_batch_motion_async": (_I, [_P, _P, _P]),
    "tsl_motion_from_delta": (_I, [_F, _F, _F, _F, _F, _F, _F, _F, _P, ctypes.c_uint64, ctypes.POINTER(Motion)]),
    "tsl_motion_from_odom": (_I, [_F, _F, _F, _P, ctypes.c_uint64, ctypes.POINTER(Motion)]),
    "tsl_filter_motion_async": (_I, [_P, ctypes.POINTER(Motion), _P]),
    "tsl_filter_set_timing": (_I, [_P, _I]),
    "tsl_filter_eval_timing": (_I, [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64),
                                    ctypes.POINTER(ctypes.c_uint64)]),
    "tsl_check_sub_table": (_I, [_F, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64),
                                 ctypes.POINTER(ctypes.c_uint32)]),
    "tsl_probe_gather": (_I, [ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_double)]),
    "tsl_probe_copy": (_I, [ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    "tsl_stream_create": (_I, [ctypes.POINTER(_P)]),
    "tsl_stream_destroy": (_I, [_P]),
    "tsl_stream_synchronize": (_I, [_P]),
    "tsl_event_create": (_I, [ctypes.POINTER(_P)]),
    "tsl_event_destroy": (_I, [_P]),
    "tsl_event_record": (_I, [_P, _P]),
    "tsl_event_elapsed_ms": (_I, [_P, _P, ctypes.POINTER(_F)]),
}

for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


class TslError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


def check(rc):
    """Raise TslError for a non-zero status, with the library's message."""
    if rc != TSL_OK:
        msg = lib.tsl_last_error().decode(errors="replace")
        raise TslError(rc, msg or f"tsl error {rc}")
    return rc


def ptr(a):
    """Data pointer of a contiguous numpy array (or None)."""
    if a is None:
        return None
    return ctypes.c_void_p(a.ctypes.data)


def device_count():
    n = ctypes.c_int(0)
    check(lib.tsl_device_count(ctypes.byref(n)))
    return n.value
