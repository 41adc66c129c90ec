"""Scan reduction (TSDFEvaluator::evaluateParticles,
src/evaluation/tsdf_evaluator.cpp:304-376): the oracle against a direct
pure-Python restatement of the reference loop (dict = unordered_set with
first-insert-wins), on hand-made and random small scans.  No GPU."""
import math

import numpy as np

import oracle_binding as ob


def _python_reference(points, rings, cell):
    f = np.float32
    cell = f(cell)
    half = f(cell / f(2))
    seen = {}
    index = 0
    for j in range(points.shape[0]):
        x, y, z = (f(v) for v in points[j])
        dist = np.sqrt(f(x * x + y * y + z * z), dtype=np.float32)
        if dist < 1.0:
            continue
        ring = int(rings[index])  # the ring iterator lags behind skipped points
        c = tuple(float(f(np.floor(f(v / cell)) * cell + half)) for v in (x, y, z))
        key = (ring, c)
        if any(math.isnan(v) for v in c):
            key = (ring, c, j)  # NaN centres never compare equal
        if key not in seen:
            seen[key] = (ring, index, (x, y, z))
        index += 1
    ordered = sorted(seen.values(), key=lambda t: (t[0], t[1]))
    return np.array([t[2] for t in ordered], np.float32).reshape(-1, 3)


def test_hand_made_scan():
    pts = np.array([
        [0.5, 0.0, 0.0],      # dropped (< 1 m): the ring iterator stays
        [2.0, 0.0, 0.0],      # kept, index 0, ring[0] = 3
        [2.01, 0.01, 0.0],    # same voxel, ring[1] = 3 -> duplicate of the previous
        [2.02, 0.0, 0.0],     # same voxel, ring[2] = 1 -> new key
        [0.0, 0.3, 0.0],      # dropped
        [5.0, 5.0, 1.0],      # kept, index 3, ring[3] = 0
    ], np.float32)
    rings = np.array([3, 3, 1, 0, 9, 9], np.int16)
    got = ob.reduce_scan(pts, rings, 0.064)
    expect = np.array([[5.0, 5.0, 1.0], [2.02, 0.0, 0.0], [2.0, 0.0, 0.0]], np.float32)
    assert np.array_equal(got, expect)
    assert np.array_equal(got, _python_reference(pts, rings, 0.064))


def test_ring_out_of_range_is_reported():
    pts = np.array([[2.0, 0, 0]], np.float32)
    assert ob.reduce_scan(pts, np.array([64], np.int16)) is None
    assert ob.reduce_scan(pts, np.array([-1], np.int16)) is None


def test_random_scans_match_python_restatement():
    rng = np.random.default_rng(4)
    for trial in range(12):
        n = int(rng.integers(1, 1500))
        pts = rng.normal(0, 3, (n, 3)).astype(np.float32)
        # clusters so that deduplication happens
        pts[: n // 3] = np.round(pts[: n // 3] * 4) / 4
        rings = rng.integers(0, 64, n).astype(np.int16)
        if trial % 3 == 0:
            rings[:] = rng.integers(0, 4, n)
        cell = [0.064, 0.05, 0.2][trial % 3]
        got = ob.reduce_scan(pts, rings, cell)
        ref = _python_reference(pts, rings, cell)
        assert got.shape == ref.shape, trial
        assert np.array_equal(got, ref), trial
