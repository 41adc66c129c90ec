"""A map built from TSDF chunks (createTSDFMap, map_util.h:17-154) through the
HIP sensor update: a box room written as 2 x 2 x 1 chunks of raw TSDF words,
turned into the reference's cells and block table on the host, uploaded, and
evaluated against the oracle on the same host map.

Bar: voxel indices bit-exact; weights within 1e-5 relative of the oracle's
float64 sum."""
import numpy as np
import pytest

import oracle_binding as ob

pytestmark = pytest.mark.gpu

WEIGHT_RTOL = 1e-5
WALLS = np.array([[-3.5, 3.5], [-3.0, 3.0], [0.3, 3.0]])  # room box in metres


@pytest.fixture(scope="module")
def tsl():
    import tsdf_localization_amd as t
    assert t.device_count() >= 1, "no HIP device"
    return t


def room_chunks():
    chunks = {}
    for cx in (-1, 0):
        for cy in (-1, 0):
            i, j, k = np.meshgrid(np.arange(64), np.arange(64), np.arange(64), indexing="ij")
            xyz = np.stack([(64 * cx + i) * 0.064, (64 * cy + j) * 0.064, k * 0.064], -1)
            d = np.min(np.concatenate([xyz - WALLS[:, 0], WALLS[:, 1] - xyz], -1), -1)  # inside > 0
            mm = np.clip(np.round(d * 1000), -32767, 32767).astype(np.int64)
            weight = np.where(np.abs(mm) < 1000, 1, 0)
            chunks[f"{cx}_{cy}_0"] = ((mm & 0xFFFF) | (weight << 16)).astype(np.uint32).reshape(-1)
    return chunks


def room_scan(origin, n, rng):
    d = rng.normal(size=(n, 3))
    d[:, 2] *= 0.3
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = np.full(n, np.inf)
    for a in range(3):
        for b in range(2):
            with np.errstate(divide="ignore", invalid="ignore"):
                s = (WALLS[a, b] - origin[a]) / d[:, a]
            t = np.where((s > 0) & (s < t), s, t)
    return (d * t[:, None]).astype(np.float32)  # in the scanner frame


def test_chunk_map_update_against_oracle(tsl):
    from tsdf_localization_amd import mapfile
    m, free = mapfile.create_tsdf_map(room_chunks(), sigma=0.1)
    assert free.shape[0] > 0 and m.dataSize() > 0
    om = ob.OracleMap.from_host_map(m)
    rng = np.random.default_rng(77)
    origin = np.array([0.4, -0.2, 1.2])
    pts = room_scan(origin, 3000, rng)
    tf = np.eye(4, dtype=np.float32).reshape(-1)
    parts = np.zeros((700, 7), np.float32)
    parts[:, :3] = origin + rng.normal(scale=[0.3, 0.3, 0.05], size=(700, 3))
    parts[:, 3:5] = rng.normal(scale=0.02, size=(700, 2))
    parts[:, 5] = rng.normal(scale=0.2, size=700)
    parts[:, 6] = 1.0 / 700
    ev = tsl.TSDFEvaluator(m)
    print(f"chunk map: {m.dataSize()} voxels, trace 32 x {pts.shape[0]}", flush=True)
    assert np.array_equal(ev.trace_indices(parts[:32], pts, tf), om.trace(parts[:32], pts, tf))
    raw = ev.evaluate_raw(parts, pts, tf)
    _, w64 = om.evaluate(parts, pts, tf)
    assert np.max(np.abs(raw - w64) / np.maximum(np.abs(w64), 1e-300)) <= WEIGHT_RTOL
    # the particles nearest the true pose score best
    err = np.linalg.norm(parts[:, :2] - origin[:2], axis=1) + np.abs(parts[:, 5])
    assert np.mean(raw[err < np.quantile(err, 0.1)]) > np.mean(raw[err > np.quantile(err, 0.9)])
