"""The C++ drop-in (include/tsdf_localization/hip_evaluator.h) end to end:
a compiled C++ program builds a map with CudaSubVoxelMap::setData, runs
HipEvaluator::evaluate through an Evaluator pointer and
HipSystematicResampler::resample on a seeded generator, exactly as the
reference's mcl_3d drives CudaEvaluator and SystematicResampler.  Its output
is checked against the oracle: normalised weights within 1e-5 relative of the
oracle's float64 sum, resampled particles bit-exact for the uniform the
generator drew, the pose against the oracle's normalisation, and the
reference's error message for a zero weight sum."""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle_binding as ob

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tsdf_localization_amd", "_build", "dropin_check")
WEIGHT_RTOL = 1e-5


def _scene(seed):
    rng = np.random.default_rng(seed)
    lo, hi, res, init = (-1.0, -1.0, -1.0), (5.0, 4.0, 3.0), 0.05, 0.0
    # a few walls of TSDF-like values (LUT values of |d| < 0.3 m)
    cells, vals = [], []
    for x0 in (0.5, 3.5):
        d = rng.uniform(-0.3, 0.3, 4000).astype(np.float32)
        ys = rng.integers(0, 80, 4000)
        zs = rng.integers(0, 60, 4000)
        xs = np.round((x0 + d - lo[0]) / res).astype(int)
        ijk = np.stack([xs, ys, zs], 1)
        c = (np.array(lo, np.float32) + ijk.astype(np.float32) * np.float32(res)).astype(np.float32)
        cells.append(c)
        vals.append(np.array([ob.olib.orc_likelihood_value(int(round(v * 1000)), 0.1) for v in d], np.float32))
    cells = np.concatenate(cells)
    vals = np.concatenate(vals)
    P, N = 3000, 1500
    parts = np.zeros((P, 7), np.float32)
    parts[:, 0] = rng.uniform(1.0, 3.0, P)
    parts[:, 1] = rng.uniform(0.5, 2.5, P)
    parts[:, 2] = rng.uniform(0.5, 1.5, P)
    parts[:, 3:5] = rng.normal(0, 0.05, (P, 2))
    parts[:, 5] = rng.uniform(-np.pi, np.pi, P)
    parts[:, 6] = 1.0 / P
    ang = rng.uniform(-np.pi, np.pi, N)
    r = rng.uniform(0.5, 3.0, N)
    pts = np.stack([r * np.cos(ang), r * np.sin(ang), rng.uniform(-0.5, 1.0, N)], 1).astype(np.float32)
    tf = np.eye(4, dtype=np.float32).reshape(-1)
    tf[3], tf[11] = 0.1, 0.3
    return (lo, hi, res, init), cells, vals, parts, pts, tf


def _write_input(path, box, cells, vals, parts, pts, tf, seed, params):
    lo, hi, res, init = box
    with open(path, "wb") as f:
        f.write(np.array([*lo, *hi, res, init], np.float32).tobytes())
        f.write(struct.pack("<Q", cells.shape[0]))
        f.write(np.concatenate([cells, vals[:, None]], 1).astype(np.float32).tobytes())
        f.write(struct.pack("<Q", parts.shape[0]))
        f.write(parts.astype(np.float32).tobytes())
        f.write(struct.pack("<Q", pts.shape[0]))
        f.write(pts.astype(np.float32).tobytes())
        f.write(tf.astype(np.float32).tobytes())
        f.write(struct.pack("<I", seed))
        f.write(np.array(params, np.float32).tobytes())


def _read_output(path, P):
    b = open(path, "rb").read()
    o = 0
    pose = np.frombuffer(b, np.float64, 7, o); o += 56
    cov = np.frombuffer(b, np.float64, 36, o); o += 288
    w = np.frombuffer(b, np.float32, P, o); o += 4 * P
    U = np.frombuffer(b, np.float32, 1, o)[0]; o += 4
    n = struct.unpack_from("<Q", b, o)[0]; o += 8
    res = np.frombuffer(b, np.float32, 7 * n, o).reshape(n, 7); o += 28 * n
    err = struct.unpack_from("<i", b, o)[0]; o += 4
    wc = np.frombuffer(b, np.float32, P, o); o += 4 * P
    cpose = np.frombuffer(b, np.float64, 3, o)
    return pose, cov, w, U, res, err, wc, cpose


@pytest.mark.parametrize("seed", [7, 8])
def test_cpp_dropin_matches_oracle(tmp_path, seed):
    assert os.path.exists(BIN), "C++ drop-in program not built (__graft_entry__.build())"
    box, cells, vals, parts, pts, tf = _scene(seed)
    params = (0.9, 0.1, 0.0, 100.0)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, box, cells, vals, parts, pts, tf, seed, params)
    print(f"dropin_check seed={seed} P={parts.shape[0]} N={pts.shape[0]}", flush=True)
    r = subprocess.run([BIN, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    pose, cov, w, U, resampled, err, wc, cpose = _read_output(fout, parts.shape[0])

    c, grid, data = ob.set_data(box[0], box[1], box[2], box[3], cells, vals)
    om = ob.OracleMap(c, grid, data)
    _, w64 = om.evaluate(parts, pts, tf, *params)
    assert np.count_nonzero(w64) > parts.shape[0] // 2
    ref = w64 / w64.sum()
    assert np.max(np.abs(w.astype(np.float64) - ref) / ref) <= WEIGHT_RTOL

    _, pref = ob.normalize(parts, w64.astype(np.float32))
    assert np.allclose(pose[:3], pref[:3], rtol=0, atol=1e-4)
    assert abs(np.linalg.norm(pose[3:]) - 1.0) < 1e-6
    assert np.all(cov == 0.0)

    # resampling consumed the device's normalised weights and the drawn U
    assert 0.0 <= U < np.float32(1.0 / parts.shape[0])
    idx = ob.resample(w, U)
    expect = parts.copy()
    expect[:, 6] = w
    assert resampled.shape[0] == idx.size
    assert np.array_equal(resampled, expect[idx])

    assert err == 1, "zero weight sum must raise 'No particle is valid!'"

    # the PointCloud2 overload: the oracle's 0.064 m centre reduction of the
    # same points (rings = index % 16), evaluated on the initial particles
    centres = ob.reduce_cloud(pts, np.arange(pts.shape[0]) % 16)
    assert 0 < centres.shape[0] <= pts.shape[0]
    _, wc64 = om.evaluate(parts, centres, tf, *params)
    refc = wc64 / wc64.sum()
    assert np.max(np.abs(wc.astype(np.float64) - refc) / refc) <= WEIGHT_RTOL
    _, prefc = ob.normalize(parts, wc64.astype(np.float32))
    assert np.allclose(cpose, prefc[:3], rtol=0, atol=1e-4)
