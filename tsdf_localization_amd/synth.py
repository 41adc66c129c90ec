"""Seeded synthetic workloads (scene -> TSDF map, ray-cast scan, particles).

Wraps libsynth.so.  The named workloads follow the benchmark configurations
of this project:
  room()      dense 256x256x64 @ 0.05 m int16 TSDF (trunc 0.3 m) of a room with
              walls, floor, ceiling and 20 boxes/cylinders; 16x1024 ray-cast
              scan voxel-filtered to ~2k points; 1k particles ~ N(true; 0.2 m, 5 deg)
  building()  sparse 1024x1024x128 @ 0.05 m int16 TSDF of a multi-room building;
              128x2048 scan reduced to ~20k points; 100k particles uniform over
              free space, yaw uniform, roll/pitch within +-3 deg
  batched()   64 independent filters on one 512x512x96 @ 0.05 m map of a
              multi-room floor: particle counts log-uniform in [500, 50k],
              scans of 2k-30k points, each filter tracking its own robot
"""
import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

from ._lib import MapCoef
from .evaluator import CudaSubVoxelMap, coef_init

_HERE = os.path.dirname(os.path.abspath(__file__))
SYNTH_PATH = os.path.join(_HERE, "libsynth.so")
if not os.path.exists(SYNTH_PATH):
    raise ImportError(f"{SYNTH_PATH} is not built: run __graft_entry__.build()")
_s = ctypes.CDLL(SYNTH_PATH)

_P = ctypes.c_void_p
_D = ctypes.c_double
_sigs = {
    "synth_scene_new": (_P, []),
    "synth_scene_free": (None, [_P]),
    "synth_free": (None, [_P]),
    "synth_scene_add_box": (ctypes.c_int, [_P, _D, _D, _D, _D, _D, _D]),
    "synth_scene_add_cylinder": (ctypes.c_int, [_P, _D, _D, _D, _D, _D]),
    "synth_scene_count": (ctypes.c_int, [_P]),
    "synth_scene_build_room": (None, [_P, ctypes.c_uint64, _D, _D, _D, ctypes.c_int, _D, _D, _D]),
    "synth_scene_build_building": (None, [_P, ctypes.c_uint64, _D, _D, _D, _D, ctypes.c_int, _D]),
    "synth_sdf": (_D, [_P, _D, _D, _D]),
    "synth_sdf_batch": (None, [_P, _P, ctypes.c_size_t, _P]),
    "synth_map_build": (ctypes.c_int, [_P, ctypes.POINTER(MapCoef), ctypes.c_int, ctypes.c_int, _P,
                                       ctypes.POINTER(_P)]),
    "synth_raycast": (ctypes.c_size_t, [_P, _P, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, ctypes.c_uint64,
                                        _P, _P]),
    "synth_particles_normal": (None, [ctypes.c_uint64, ctypes.c_size_t, _P, _P, _P]),
    "synth_particles_free": (ctypes.c_size_t, [_P, ctypes.c_uint64, ctypes.c_size_t, _P, _P, _D, _D, _P]),
    "synth_voxel_filter": (ctypes.c_size_t, [_P, _P, ctypes.c_size_t, _D, _D, _P, _P]),
}
for _n, (_r, _a) in _sigs.items():
    _f = getattr(_s, _n)
    _f.restype = _r
    _f.argtypes = _a


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class Scene:
    def __init__(self):
        self.h = _s.synth_scene_new()

    def __del__(self):
        if getattr(self, "h", None):
            _s.synth_scene_free(self.h)
            self.h = None

    def add_box(self, lo, hi):
        return _s.synth_scene_add_box(self.h, *map(float, lo), *map(float, hi))

    def add_cylinder(self, cx, cy, z0, z1, r):
        return _s.synth_scene_add_cylinder(self.h, float(cx), float(cy), float(z0), float(z1), float(r))

    def count(self):
        return _s.synth_scene_count(self.h)

    def sdf(self, xyz):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        out = np.empty(xyz.shape[0], np.float64)
        _s.synth_sdf_batch(self.h, _ptr(xyz), xyz.shape[0], _ptr(out))
        return out

    def tsdf_map(self, min_xyz, max_xyz, resolution, truncation_mm, sigma=0.1, dense=False, init_value=None):
        """CudaSubVoxelMap of this scene with int16 TSDF voxels."""
        if init_value is None:
            from .evaluator import likelihood_init
            init_value = likelihood_init(sigma)
        c = coef_init(min_xyz, max_xyz, resolution, init_value)
        grid = np.empty(c.grid_occ_size, np.int32)
        out = ctypes.c_void_p()
        rc = _s.synth_map_build(self.h, ctypes.byref(c), int(truncation_mm), int(bool(dense)), _ptr(grid),
                                ctypes.byref(out))
        if rc != 0:
            raise RuntimeError(f"synth_map_build failed ({rc})")
        n = int(c.data_size)
        buf = (ctypes.c_int16 * max(n, 1)).from_address(out.value)
        tsdf = np.frombuffer(buf, dtype=np.int16, count=n).copy()
        _s.synth_free(out)
        return CudaSubVoxelMap(c, grid, tsdf_mm=tsdf, truncation_mm=truncation_mm, sigma=sigma)

    def raycast(self, sensor_pose, beams, azimuths, el_min_deg, el_max_deg, max_range, noise_sigma=0.0, seed=0):
        total = beams * azimuths
        pts = np.empty((total, 3), np.float32)
        rings = np.empty(total, np.int32)
        pose = np.ascontiguousarray(sensor_pose, np.float64)
        n = _s.synth_raycast(self.h, _ptr(pose), beams, azimuths, el_min_deg, el_max_deg, max_range, noise_sigma,
                             seed, _ptr(pts), _ptr(rings))
        return pts[:n].copy(), rings[:n].copy()

    def particles_free(self, seed, n, lo, hi, clearance, rp):
        out = np.empty((n, 7), np.float32)
        lo = np.ascontiguousarray(lo, np.float64)
        hi = np.ascontiguousarray(hi, np.float64)
        _s.synth_particles_free(self.h, seed, n, _ptr(lo), _ptr(hi), clearance, rp, _ptr(out))
        return out


def particles_normal(seed, n, mean, sigma):
    out = np.empty((n, 7), np.float32)
    m = np.ascontiguousarray(mean, np.float64)
    s = np.ascontiguousarray(sigma, np.float64)
    _s.synth_particles_normal(seed, n, _ptr(m), _ptr(s), _ptr(out))
    return out


def voxel_filter(pts, rings, cell, min_range=1.0):
    pts = np.ascontiguousarray(pts, np.float32)
    rings = np.ascontiguousarray(rings, np.int32)
    out = np.empty_like(pts)
    out_r = np.empty_like(rings)
    n = _s.synth_voxel_filter(_ptr(pts), _ptr(rings), pts.shape[0], cell, min_range, _ptr(out), _ptr(out_r))
    return out[:n].copy(), out_r[:n].copy()


def rpy_matrix(roll, pitch, yaw):
    sa, ca = math.sin(roll), math.cos(roll)
    sb, cb = math.sin(pitch), math.cos(pitch)
    sg, cg = math.sin(yaw), math.cos(yaw)
    return np.array([[cb * cg, sa * sb * cg - ca * sg, ca * sb * cg + sa * sg],
                     [cb * sg, sa * sb * sg + ca * cg, ca * sb * sg - sa * cg],
                     [-sb, sa * cb, ca * cb]])


def tf_from_xyzrpy(x, y, z, roll, pitch, yaw):
    """Row-major 4x4 homogeneous transform (float32, 16 entries)."""
    T = np.eye(4)
    T[:3, :3] = rpy_matrix(roll, pitch, yaw)
    T[:3, 3] = (x, y, z)
    return T.astype(np.float32).reshape(-1)


def compose_pose(base, tf16):
    """Scanner pose in the map frame = base pose (x,y,z,r,p,y) composed with tf."""
    B = np.eye(4)
    B[:3, :3] = rpy_matrix(*base[3:6])
    B[:3, 3] = base[:3]
    W = B @ np.asarray(tf16, np.float64).reshape(4, 4)
    R = W[:3, :3]
    pitch = -math.asin(max(-1.0, min(1.0, R[2, 0])))
    roll = math.atan2(R[2, 1], R[2, 2])
    yaw = math.atan2(R[1, 0], R[0, 0])
    return np.array([W[0, 3], W[1, 3], W[2, 3], roll, pitch, yaw])


# scanner mounted 0.1 m ahead of and 0.4 m above the base, slightly rotated
SCANNER_TF = dict(x=0.1, y=0.0, z=0.4, roll=0.0, pitch=math.radians(1.0), yaw=math.radians(2.0))


@dataclass
class Workload:
    name: str
    map: CudaSubVoxelMap
    points: np.ndarray
    rings: np.ndarray
    particles: np.ndarray
    tf: np.ndarray
    true_pose: np.ndarray
    meta: dict = field(default_factory=dict)


def room(seed=0x5EED0001, num_particles=1000, size=(12.8, 12.8, 3.2), resolution=0.05, truncation_mm=300,
         beams=16, azimuths=1024, filter_cell=0.3, objects=20):
    sx, sy, sz = size
    true_pose = np.array([sx / 2, sy / 2, 1.0, 0.0, 0.0, 0.3])
    sc = Scene()
    _s.synth_scene_build_room(sc.h, seed, sx, sy, sz, objects, true_pose[0], true_pose[1], 1.5)
    m = sc.tsdf_map((0, 0, 0), (sx, sy, sz), resolution, truncation_mm, dense=True)
    tf = tf_from_xyzrpy(**SCANNER_TF)
    sensor = compose_pose(true_pose, tf)
    raw, rings = sc.raycast(sensor, beams, azimuths, -15.0, 15.0, 30.0, 0.01, seed)
    pts, rings = voxel_filter(raw, rings, filter_cell, 1.0)
    parts = particles_normal(seed, num_particles, true_pose,
                             [0.2, 0.2, 0.2, math.radians(5), math.radians(5), math.radians(5)])
    return Workload("room", m, pts, rings, parts, tf, true_pose,
                    {"scene_objects": sc.count(), "raw_points": int(raw.shape[0])})


def building(seed=0x5EED0002, num_particles=100_000, size=(51.2, 51.2, 6.4), resolution=0.05, truncation_mm=300,
             beams=128, azimuths=2048, filter_cell=0.07, room_size=10.0, objects_per_room=4, ceiling_fraction=0.3,
             wall_h=3.0, target_points=20000):
    sx, sy, sz = size
    sc = Scene()
    _s.synth_scene_build_building(sc.h, seed, sx, sy, wall_h, room_size, objects_per_room, ceiling_fraction)
    m = sc.tsdf_map((0, 0, 0), (sx, sy, sz), resolution, truncation_mm, dense=False)
    rng = np.random.default_rng(seed)
    # true pose: a free spot near the middle of a room
    for _ in range(1000):
        x, y = rng.uniform(2, sx - 2), rng.uniform(2, sy - 2)
        if sc.sdf([[x, y, 1.0]])[0] > 1.0:
            break
    true_pose = np.array([x, y, 1.0, 0.0, 0.0, rng.uniform(-math.pi, math.pi)])
    tf = tf_from_xyzrpy(**SCANNER_TF)
    sensor = compose_pose(true_pose, tf)
    raw, rings = sc.raycast(sensor, beams, azimuths, -22.5, 22.5, 60.0, 0.01, seed)
    pts, rings = voxel_filter(raw, rings, filter_cell, 1.0)
    if target_points is not None and pts.shape[0] > target_points:
        keep = np.sort(rng.choice(pts.shape[0], target_points, replace=False))
        pts, rings = pts[keep], rings[keep]
    parts = sc.particles_free(seed, num_particles, (0.5, 0.5, 0.5), (sx - 0.5, sy - 0.5, 2.0), 0.3,
                              math.radians(3.0))
    blocks = int(np.count_nonzero(m.grid_occ >= 0))
    return Workload("building", m, pts, rings, parts, tf, true_pose,
                    {"scene_objects": sc.count(), "raw_points": int(raw.shape[0]), "allocated_blocks": blocks,
                     "block_fraction": blocks / m.grid_occ.size})


@dataclass
class Robot:
    particles: np.ndarray
    points: np.ndarray
    true_pose: np.ndarray


@dataclass
class BatchedWorkload:
    name: str
    map: CudaSubVoxelMap
    tf: np.ndarray
    robots: list
    meta: dict = field(default_factory=dict)


def batched(seed=0x5EED0003, filters=64, size=(25.6, 25.6, 4.8), resolution=0.05, truncation_mm=300,
            particles=(500, 50_000), points=(2_000, 30_000), beams=128, azimuths=2048, filter_cell=0.05,
            room_size=8.0, objects_per_room=4, ceiling_fraction=0.3, wall_h=3.0):
    """One shared map, `filters` robots at seeded free poses; robot i has
    log-uniform particle count in `particles` (around its pose: 0.2 m, 5 deg)
    and a scan of a uniform point count in `points` (its own ray cast,
    voxel-filtered, subsampled; fewer if the cast returns fewer)."""
    sx, sy, sz = size
    sc = Scene()
    _s.synth_scene_build_building(sc.h, seed, sx, sy, wall_h, room_size, objects_per_room, ceiling_fraction)
    m = sc.tsdf_map((0, 0, 0), (sx, sy, sz), resolution, truncation_mm, dense=False)
    rng = np.random.default_rng(seed)
    tf = tf_from_xyzrpy(**SCANNER_TF)
    robots = []
    lo_p, hi_p = particles
    for i in range(filters):
        for _ in range(1000):
            x, y = rng.uniform(1.5, sx - 1.5), rng.uniform(1.5, sy - 1.5)
            if sc.sdf([[x, y, 1.0]])[0] > 0.6:
                break
        true_pose = np.array([x, y, 1.0, 0.0, 0.0, rng.uniform(-math.pi, math.pi)])
        P = int(round(math.exp(rng.uniform(math.log(lo_p), math.log(hi_p)))))
        N = int(rng.integers(points[0], points[1] + 1))
        raw, rings = sc.raycast(compose_pose(true_pose, tf), beams, azimuths, -22.5, 22.5, 60.0, 0.01,
                                seed + 17 * i)
        pts, rings = voxel_filter(raw, rings, filter_cell, 1.0)
        if pts.shape[0] > N:
            pts = pts[np.sort(rng.choice(pts.shape[0], N, replace=False))]
        parts = particles_normal(seed + i, P, true_pose,
                                 [0.2, 0.2, 0.1, math.radians(5), math.radians(2), math.radians(5)])
        robots.append(Robot(parts, np.ascontiguousarray(pts, np.float32), true_pose))
    blocks = int(np.count_nonzero(m.grid_occ >= 0))
    return BatchedWorkload("batched", m, tf, robots,
                           {"allocated_blocks": blocks, "block_fraction": blocks / m.grid_occ.size,
                            "particles_total": int(sum(r.particles.shape[0] for r in robots)),
                            "evals_per_step": int(sum(r.particles.shape[0] * r.points.shape[0] for r in robots))})


@dataclass
class ServingWorkload:
    name: str
    map: CudaSubVoxelMap
    tf: np.ndarray
    robots: list          # Robot per simulated robot (its scan: one of the shared scans)
    scans: list           # distinct scans (robots reuse them)
    meta: dict = field(default_factory=dict)


def serving(seed=0x5EED0004, robots=256, particles=(2_000, 20_000), points=(5_000, 20_000), distinct_scans=32,
            beams=64, azimuths=1024):
    """Robots on the global-localisation building map: each with a particle
    set (count log-uniform in `particles`, around its pose) and a scan of
    5k-20k points; ray casts are shared between robots (distinct_scans poses,
    robot i uses scan i mod distinct_scans) to keep generation short."""
    w = building(num_particles=1)
    m = w.map
    rng = np.random.default_rng(seed)
    sc = Scene()
    # the same building as building(): rebuild the scene for ray casting
    sx, sy, sz = 51.2, 51.2, 6.4
    _s.synth_scene_build_building(sc.h, 0x5EED0002, sx, sy, 3.0, 10.0, 4, 0.3)
    tf = tf_from_xyzrpy(**SCANNER_TF)
    scans, poses = [], []
    for i in range(distinct_scans):
        N = int(rng.integers(points[0], points[1] + 1))
        for _ in range(100):  # a free pose whose scan has at least N points
            for _ in range(1000):
                x, y = rng.uniform(2, sx - 2), rng.uniform(2, sy - 2)
                if sc.sdf([[x, y, 1.0]])[0] > 0.8:
                    break
            pose = np.array([x, y, 1.0, 0.0, 0.0, rng.uniform(-math.pi, math.pi)])
            raw, rings = sc.raycast(compose_pose(pose, tf), beams, azimuths, -22.5, 22.5, 60.0, 0.01, seed + i)
            pts, rings = voxel_filter(raw, rings, 0.05, 1.0)
            if pts.shape[0] >= N:
                break
        if pts.shape[0] > N:
            pts = pts[np.sort(rng.choice(pts.shape[0], N, replace=False))]
        scans.append(np.ascontiguousarray(pts, np.float32))
        poses.append(pose)
    out = []
    for r in range(robots):
        pose = poses[r % distinct_scans]
        P = int(round(math.exp(rng.uniform(math.log(particles[0]), math.log(particles[1])))))
        parts = particles_normal(seed + 7 * r, P, pose, [0.2, 0.2, 0.05, math.radians(5), math.radians(1),
                                                         math.radians(5)])
        out.append(Robot(parts, scans[r % distinct_scans], pose))
    return ServingWorkload("serving", m, tf, out, scans,
                           {"robots": robots, "distinct_scans": distinct_scans,
                            "mean_evals_per_request": float(np.mean([r.particles.shape[0] * r.points.shape[0]
                                                                     for r in out]))})



@dataclass
class MixedWorkload:
    name: str
    maps: list            # CudaSubVoxelMap per map
    tf: np.ndarray
    tracking: list        # per map: list of Robot (particles around the robot's pose)
    global_robots: list   # per map: one Robot with particles over the whole free space
    meta: dict = field(default_factory=dict)


def mixed(seed=0x5EED0005, maps=4, tracking=9, tracking_particles=5_000, tracking_points=10_000,
          global_particles=1_000_000, global_points=4_000, beams=64, azimuths=1024, campus_tiles=(12, 11)):
    """Several maps served side by side, each with `tracking` robots
    (particle sets of tracking_particles around their poses, scans of
    tracking_points) and one global-localisation robot (global_particles
    over the free space, a scan of global_points).  Each map is a campus of
    campus_tiles building tiles (four seeded buildings per map; at 12 x 11
    tiles ~7.3e9 int16 voxels, ~14.6 GB, block offsets past 2^31: the 64-bit
    map form).  campus_tiles=None: each map is one global-localisation
    building (51.2 x 51.2 x 6.4 m) instead."""
    if campus_tiles is None:
        return _mixed_buildings(seed, maps, tracking, tracking_particles, tracking_points, global_particles,
                                global_points, beams, azimuths)
    rng = np.random.default_rng(seed)
    tf = tf_from_xyzrpy(**SCANNER_TF)
    out_maps, out_tracking, out_global = [], [], []
    for k in range(maps):
        s = seed + 101 * k
        cp = campus(s, tiles=campus_tiles)
        out_maps.append(cp.map)
        robots = []
        for i in range(tracking):
            pose, pts = campus_scan(cp, rng, tf, tracking_points, beams, azimuths)
            parts = particles_normal(s + 7 * i, tracking_particles, pose,
                                     [0.2, 0.2, 0.05, math.radians(5), math.radians(1), math.radians(5)])
            robots.append(Robot(parts, pts, pose))
        out_tracking.append(robots)
        pose, pts = campus_scan(cp, rng, tf, global_points, beams, azimuths)
        out_global.append(Robot(campus_particles(cp, s + 5, global_particles, math.radians(3.0)), pts, pose))
    return MixedWorkload("mixed", out_maps, tf, out_tracking, out_global,
                         {"maps": maps, "tracking_per_map": tracking, "campus_tiles": list(campus_tiles),
                          "map_bytes": [int(m.dataSize()) * 2 for m in out_maps]})


def _mixed_buildings(seed, maps, tracking, tracking_particles, tracking_points, global_particles, global_points,
                     beams, azimuths):
    rng = np.random.default_rng(seed)
    tf = tf_from_xyzrpy(**SCANNER_TF)
    sx, sy = 51.2, 51.2
    out_maps, out_tracking, out_global = [], [], []
    for k in range(maps):
        s = seed + 101 * k
        w = building(seed=s, num_particles=global_particles, target_points=global_points)
        out_maps.append(w.map)
        out_global.append(Robot(w.particles, np.ascontiguousarray(w.points, np.float32), w.true_pose))
        sc = Scene()
        _s.synth_scene_build_building(sc.h, s, sx, sy, 3.0, 10.0, 4, 0.3)
        robots = []
        for i in range(tracking):
            for _ in range(100):  # a free pose whose scan has enough points
                for _ in range(1000):
                    x, y = rng.uniform(2, sx - 2), rng.uniform(2, sy - 2)
                    if sc.sdf([[x, y, 1.0]])[0] > 0.8:
                        break
                pose = np.array([x, y, 1.0, 0.0, 0.0, rng.uniform(-math.pi, math.pi)])
                raw, rings = sc.raycast(compose_pose(pose, tf), beams, azimuths, -22.5, 22.5, 60.0, 0.01, s + i)
                pts, rings = voxel_filter(raw, rings, 0.05, 1.0)
                if pts.shape[0] >= tracking_points:
                    break
            if pts.shape[0] > tracking_points:
                pts = pts[np.sort(rng.choice(pts.shape[0], tracking_points, replace=False))]
            parts = particles_normal(s + 7 * i, tracking_particles, pose,
                                     [0.2, 0.2, 0.05, math.radians(5), math.radians(1), math.radians(5)])
            robots.append(Robot(parts, np.ascontiguousarray(pts, np.float32), pose))
        out_tracking.append(robots)
    return MixedWorkload("mixed", out_maps, tf, out_tracking, out_global,
                         {"maps": maps, "tracking_per_map": tracking,
                          "map_bytes": [int(m.dataSize()) * 2 for m in out_maps]})


# ---------------------------------------------------------------------------
# Campus maps: many building tiles side by side, past 2^31 voxels
# ---------------------------------------------------------------------------

TILE_BLOCKS = 52  # a building tile spans 52 one-metre blocks in x and y


@dataclass
class Campus:
    map: CudaSubVoxelMap          # 64-bit form: int64 block offsets, int16 millimetres
    tiles: tuple                  # (tiles in x, tiles in y)
    tile_of: np.ndarray           # [ty, tx] -> base building index
    scenes: list                  # the base buildings' scenes (for ray casts and free-space samples)


def campus(seed, tiles=(12, 11), bases=4, resolution=0.05, truncation_mm=300, size=(51.2, 51.2, 6.4),
           room_size=10.0, objects_per_room=4, ceiling_fraction=0.3, wall_h=3.0):
    """A sparse TSDF map of tiles[0] x tiles[1] building tiles, each one of
    `bases` seeded buildings (the global-localisation building layout),
    placed every 52 m: block (bx, by, bz) of the campus is block
    (bx mod 52, by mod 52, bz) of its tile's building, so its voxels are that
    block's voxels.  Block offsets are assigned in linear block order as
    setData does, past 2^31 voxels for large campuses (the 64-bit form).
    At 12 x 11 tiles of the ~7,200-block building: ~950,000 blocks,
    ~7.6e9 voxels, ~15 GB of int16."""
    sx, sy, sz = size
    rng = np.random.default_rng(seed)
    base_maps, scenes = [], []
    for b in range(bases):
        sc = Scene()
        _s.synth_scene_build_building(sc.h, seed + 1000 * b, sx, sy, wall_h, room_size, objects_per_room,
                                      ceiling_fraction)
        base_maps.append(sc.tsdf_map((0, 0, 0), (sx, sy, sz), resolution, truncation_mm, dense=False))
        scenes.append(sc)
    c0 = base_maps[0].coef()
    assert c0.up_dim_x == TILE_BLOCKS and c0.up_dim_y == TILE_BLOCKS
    sd3 = int(c0.sub_dim) ** 3
    nx, ny = tiles
    tile_of = rng.integers(0, bases, (ny, nx))
    init = c0.init_value
    c = coef_init((0, 0, 0), (TILE_BLOCKS * nx - 0.8, TILE_BLOCKS * ny - 0.8, sz), resolution, init)
    UX, UY, UZ = int(c.up_dim_x), int(c.up_dim_y), int(c.up_dim_z)
    assert (UX, UY, UZ) == (TILE_BLOCKS * nx, TILE_BLOCKS * ny, int(c0.up_dim_z))
    # every campus block's block in the stack of the bases' blocks (-1: empty)
    starts = np.cumsum([0] + [int(m.dataSize()) // sd3 for m in base_maps])
    grids = np.stack([m.grid_occ for m in base_maps])            # [base, 52*52*UZ]
    bz, by, bx = np.meshgrid(np.arange(UZ), np.arange(UY), np.arange(UX), indexing="ij")
    tile = tile_of[by // TILE_BLOCKS, bx // TILE_BLOCKS].ravel()
    local = ((bx % TILE_BLOCKS) + (by % TILE_BLOCKS) * TILE_BLOCKS + bz * TILE_BLOCKS * TILE_BLOCKS).ravel()
    off = grids[tile, local].astype(np.int64)
    alloc = off >= 0
    src = (starts[tile] + off // sd3)[alloc]
    grid64 = np.full(UX * UY * UZ, -1, np.int64)
    grid64[alloc] = np.arange(int(alloc.sum()), dtype=np.int64) * sd3
    stack = np.concatenate([m.tsdf_mm for m in base_maps]).reshape(-1, sd3)
    mm = stack[src].reshape(-1)
    c.data_size = mm.size
    m = CudaSubVoxelMap(c, grid64, tsdf_mm=mm, truncation_mm=truncation_mm, sigma=base_maps[0].sigma)
    return Campus(m, (nx, ny), tile_of, scenes)


def campus_pose(cp, rng, clearance=0.8):
    """A free pose somewhere on the campus: a tile, then a free spot of its
    building; returns the campus pose and the tile's base scene and offset."""
    nx, ny = cp.tiles
    for _ in range(1000):
        tx, ty = int(rng.integers(0, nx)), int(rng.integers(0, ny))
        sc = cp.scenes[cp.tile_of[ty, tx]]
        x, y = rng.uniform(2, 49.2), rng.uniform(2, 49.2)
        if sc.sdf([[x, y, 1.0]])[0] > clearance:
            break
    ox, oy = TILE_BLOCKS * tx, TILE_BLOCKS * ty
    return np.array([ox + x, oy + y, 1.0, 0.0, 0.0, rng.uniform(-math.pi, math.pi)]), sc, (ox, oy)


def campus_particles(cp, seed, n, rp):
    """n particles uniform over the campus's free space (a tile per particle,
    then a free spot of its building), yaw uniform, roll/pitch within rp."""
    rng = np.random.default_rng(seed)
    nx, ny = cp.tiles
    t = rng.integers(0, nx * ny, n)
    out = np.empty((n, 7), np.float32)
    for b, sc in enumerate(cp.scenes):
        for ti in np.unique(t):
            tx, ty = int(ti % nx), int(ti // nx)
            if cp.tile_of[ty, tx] != b:
                continue
            sel = np.nonzero(t == ti)[0]
            p = sc.particles_free(seed + int(ti), sel.size, (0.5, 0.5, 0.5), (50.7, 50.7, 2.0), 0.3, rp)
            p[:, 0] += TILE_BLOCKS * tx
            p[:, 1] += TILE_BLOCKS * ty
            out[sel] = p
    out[:, 6] = 1.0 / n
    return out


def campus_scan(cp, rng, tf, points, beams=64, azimuths=1024):
    """A robot at a free campus pose and its scan (ray cast in its tile's
    building, voxel-filtered at 0.05 m, `points` of them): (pose, points)."""
    for _ in range(100):
        pose, sc, (ox, oy) = campus_pose(cp, rng)
        local = pose.copy()
        local[0] -= ox
        local[1] -= oy
        raw, rings = sc.raycast(compose_pose(local, tf), beams, azimuths, -22.5, 22.5, 60.0, 0.01,
                                int(rng.integers(0, 1 << 30)))
        pts, rings = voxel_filter(raw, rings, 0.05, 1.0)
        if pts.shape[0] >= points:
            break
    if pts.shape[0] > points:
        pts = pts[np.sort(rng.choice(pts.shape[0], points, replace=False))]
    return pose, np.ascontiguousarray(pts, np.float32)
