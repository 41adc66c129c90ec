"""Reader/writer of the reference's .mcl snapshot files (MCLFile,
src/util/mcl_file.cpp:14-113) over the C ABI (tsl_mcl_read/tsl_mcl_write)."""
import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import TSL_ERR_CAPACITY, TSL_OK, check, lib, ptr


@dataclass
class MCLSnapshot:
    points: np.ndarray     # (N, 3) float32
    rings: np.ndarray      # (N,) int32
    particles: np.ndarray  # (P, 7) float32
    tf: np.ndarray         # (16,) float32, row-major scanner -> base
    pose: np.ndarray       # (7,) float32: x, y, z, q1..q4


def write(path, points, rings, particles, tf, pose):
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    rg = np.ascontiguousarray(rings, np.int32).reshape(-1)
    ps = np.ascontiguousarray(particles, np.float32).reshape(-1, 7)
    t = np.ascontiguousarray(tf, np.float32).reshape(16)
    q = np.ascontiguousarray(pose, np.float32).reshape(7)
    if rg.size != pts.shape[0]:
        raise ValueError("one ring per point")
    check(lib.tsl_mcl_write(str(path).encode(), ptr(pts), ptr(rg), pts.shape[0], ptr(ps), ps.shape[0], ptr(t),
                            ptr(q)))


def read(path):
    n_pts, n_par = ctypes.c_size_t(), ctypes.c_size_t()
    rc = lib.tsl_mcl_read(str(path).encode(), None, None, 0, None, 0, None, None, ctypes.byref(n_pts),
                          ctypes.byref(n_par))
    if rc not in (TSL_OK, TSL_ERR_CAPACITY):
        check(rc)
    pts = np.empty((max(n_pts.value, 1), 3), np.float32)
    rg = np.empty(max(n_pts.value, 1), np.int32)
    ps = np.empty((max(n_par.value, 1), 7), np.float32)
    t = np.empty(16, np.float32)
    q = np.empty(7, np.float32)
    check(lib.tsl_mcl_read(str(path).encode(), ptr(pts), ptr(rg), pts.shape[0], ptr(ps), ps.shape[0], ptr(t), ptr(q),
                           ctypes.byref(n_pts), ctypes.byref(n_par)))
    return MCLSnapshot(pts[: n_pts.value].copy(), rg[: n_pts.value].copy(), ps[: n_par.value].copy(), t, q)
