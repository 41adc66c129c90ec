"""MI355X (gfx950) MCL sensor update for uos/tsdf_localization.

The hot path (pose transform of every scan point by every particle, TSDF
voxel lookup, score, per-particle sum, normalisation, systematic resampling)
runs in hand-written HIP kernels in libtsl.so behind the C ABI of
include/tsl.h.  This package is the Python mirror of the reference's
evaluator/resampler interface over that ABI.
"""
from .evaluator import (Batch, CudaSubVoxelMap, DeviceFilter, Event, ParticleCloud, Stream, SystematicResampler,
                        TSDFEvaluator, check_trig_libm, coef_init, likelihood_init, likelihood_value, pose_matrices,
                        resample_systematic, trig)
from ._lib import TRIG_DOUBLE, TRIG_GLIBC_FMA, TRIG_GLIBC_NOFMA, TslError, device_count
from . import evaluator  # noqa: F401  (motion helpers: evaluator.motion_from_delta, ...)

__all__ = ["Batch", "CudaSubVoxelMap", "DeviceFilter", "Event", "ParticleCloud", "Stream", "SystematicResampler",
           "TRIG_DOUBLE", "TRIG_GLIBC_FMA", "TRIG_GLIBC_NOFMA", "TSDFEvaluator", "TslError", "check_trig_libm",
           "coef_init", "device_count", "likelihood_init", "likelihood_value", "pose_matrices", "resample_systematic",
           "trig"]
