"""Python handle of the serving loop (tsl_server_* in include/tsl.h): robots'
particle sets resident on the device, requests batched continuously by a
native worker thread."""
import ctypes

import numpy as np

from ._lib import Request, Result, check, lib, ptr
from .evaluator import _particles, _points, _pose_dict, _tf


class Server:
    def __init__(self, evaluator, hbm_budget=0, max_robots=256, max_particles=20_000, max_batch=64, chunk_len=0):
        self.evaluator = evaluator
        self.handle = ctypes.c_void_p()
        check(lib.tsl_server_create(evaluator.handle, int(hbm_budget), int(max_robots), int(max_particles),
                                    int(max_batch), int(chunk_len), ctypes.byref(self.handle)))

    def add_robot(self, particles, max_points, tf_matrix):
        p = _particles(particles)
        tf = _tf(tf_matrix)
        r = ctypes.c_uint32()
        check(lib.tsl_server_add_robot(self.handle, ptr(p), p.shape[0], int(max_points), ptr(tf), ctypes.byref(r)))
        return r.value

    def remove_robot(self, robot):
        check(lib.tsl_server_remove_robot(self.handle, int(robot)))

    def submit(self, robot, motion, points, u0, request_id=0):
        pts = _points(points)
        req = Request()
        req.id = int(request_id)
        req.robot = int(robot)
        req.motion = motion
        req.points = pts.ctypes.data
        req.num_points = pts.shape[0]
        req.u0 = float(u0)
        check(lib.tsl_server_submit(self.handle, ctypes.byref(req)))

    def poll(self, max_results=1024, wait_ms=0):
        buf = (Result * max_results)()
        n = ctypes.c_size_t()
        check(lib.tsl_server_poll(self.handle, buf, max_results, ctypes.byref(n), int(wait_ms)))
        out = []
        for r in buf[: n.value]:
            out.append({"id": r.id, "robot": r.robot, "status": r.status, "pose": _pose_dict(r.pose),
                        "count": r.count, "submit_ms": r.submit_ms, "start_ms": r.start_ms, "done_ms": r.done_ms,
                        "batch": r.batch})
        return out

    def wait(self, n, timeout_ms=60_000):
        """Collect n results (raises on timeout)."""
        import time
        out = []
        t_end = time.time() + timeout_ms / 1000.0
        while len(out) < n:
            out.extend(self.poll(max(n - len(out), 1), wait_ms=50))
            if time.time() > t_end:
                raise TimeoutError(f"{len(out)} of {n} results")
        return out

    def get_particles(self, robot):
        n = ctypes.c_size_t()
        lib.tsl_server_get_particles(self.handle, int(robot), None, 0, ctypes.byref(n))
        out = np.empty((max(n.value, 1), 7), np.float32)
        check(lib.tsl_server_get_particles(self.handle, int(robot), ptr(out), out.shape[0], ctypes.byref(n)))
        return out[: n.value].copy()

    def set_slice(self, pairs, reserved_per_cu=2):
        """Particle-point pairs per launch of a bulk robot's evaluation (robots
        of 65,536+ particle slots; 0 = one launch per tick) and the workgroup
        slots per compute unit those launches leave to the other robots."""
        check(lib.tsl_server_set_slice(self.handle, int(pairs), int(reserved_per_cu)))

    def set_tile(self, bulk, k, points_per_lane):
        """Evaluation tile (particles per wavefront, points per lane) of the
        bulk class (robots of 65,536+ particle slots) or of the others."""
        check(lib.tsl_server_set_tile(self.handle, int(bool(bulk)), int(k), int(points_per_lane)))

    def stats(self):
        used, ticks, reqs = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib.tsl_server_stats(self.handle, ctypes.byref(used), ctypes.byref(ticks), ctypes.byref(reqs)))
        return {"used_bytes": used.value, "ticks": ticks.value, "requests": reqs.value}

    def close(self):
        if getattr(self, "handle", None):
            lib.tsl_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()
