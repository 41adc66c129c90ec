"""sin/cos of the particle angles on the device, in every binding, against
glibc and the oracle.

The device runs the library's restatement of glibc's sinf/cosf (both x86-64
builds) and the double-evaluated routine.  Each is compared with the host
build of the same code (pinned against glibc on every float by
tests/test_trig_binding.py), with this process's own sinf/cosf through
ctypes, with the fixture of the inputs that separate glibc's two builds, and,
through the pose matrices and the whole update, with the oracle.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle_binding as ob

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHT_RTOL = 1e-5


@pytest.fixture(scope="module")
def tsl():
    import tsdf_localization_amd as t
    assert t.device_count() >= 1, "no HIP device"
    return t


@pytest.fixture(scope="module")
def room():
    from tsdf_localization_amd import synth
    return synth.room()


def _libm():
    m = ctypes.CDLL("libm.so.6")
    for f in (m.sinf, m.cosf):
        f.restype = ctypes.c_float
        f.argtypes = [ctypes.c_float]
    return m


def _split():
    with open(os.path.join(HERE, "golden", "glibc_sinf_fma_split.json")) as f:
        rows = json.load(f)["inputs"]
    x = np.array([int(r["x"], 16) for r in rows], np.uint32).view(np.float32)
    fma = np.array([[int(v, 16) for v in r["fma"]] for r in rows], np.uint32)
    nofma = np.array([[int(v, 16) for v in r["nofma"]] for r in rows], np.uint32)
    return x, fma, nofma


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_device_equals_host_build_sampled_floats(tsl):
    """Every 257th float bit pattern (16.7M angles, all binades, both signs)."""
    u = np.arange(0, 1 << 32, 257, dtype=np.uint64).astype(np.uint32)
    x = u.view(np.float32)
    x = x[~np.isnan(x)]
    for mode in (tsl.TRIG_GLIBC_FMA, tsl.TRIG_GLIBC_NOFMA, tsl.TRIG_DOUBLE):
        ds, dc = tsl.trig(x, mode, on_device=True)
        hs, hc = tsl.trig(x, mode, on_device=False)
        assert np.array_equal(_bits(ds), _bits(hs)), mode
        assert np.array_equal(_bits(dc), _bits(hc)), mode


def test_device_glibc_split_inputs(tsl):
    x, fma, nofma = _split()
    for mode, want in ((tsl.TRIG_GLIBC_FMA, fma), (tsl.TRIG_GLIBC_NOFMA, nofma)):
        s, c = tsl.trig(x, mode, on_device=True)
        assert np.array_equal(_bits(s), want[:, 0]), mode
        assert np.array_equal(_bits(c), want[:, 1]), mode


def test_device_equals_this_hosts_libm(tsl):
    """Angles a particle filter sees (|a| < 8) and a spread of large ones,
    against this process's sinf/cosf (glibc's FMA build on an FMA3 CPU)."""
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-8, 8, 60000), rng.uniform(-1e6, 1e6, 20000),
                        rng.normal(0, 0.05, 20000)]).astype(np.float32)
    m = _libm()
    ls = np.array([m.sinf(float(v)) for v in x], np.float32)
    lc = np.array([m.cosf(float(v)) for v in x], np.float32)
    s, c = tsl.trig(x, tsl.TRIG_GLIBC_FMA, on_device=True)
    assert np.array_equal(_bits(s), _bits(ls))
    assert np.array_equal(_bits(c), _bits(lc))


@pytest.mark.parametrize("mode", [0, 2])
def test_pose_matrices_every_binding(tsl, room, mode):
    """Device pose matrices == the oracle's, bit for bit, in the binding's
    oracle counterpart (C library sinf/cosf, or the double routine)."""
    rng = np.random.default_rng(5)
    extra = room.particles[:512].copy()
    extra[:, 3:6] = rng.uniform(-40, 40, (512, 3)).astype(np.float32)
    parts = np.concatenate([room.particles, extra])
    got = tsl.pose_matrices(parts, room.tf, mode)
    ref = np.stack([ob.pose_matrix(p[:6], room.tf, ob.trig_mode_for(mode)) for p in parts])
    assert np.array_equal(_bits(got), _bits(ref))


def test_global_localisation_sample_under_each_binding(tsl):
    """A 4,000-particle sample of the global-localisation workload evaluated
    with the glibc binding and with the double one: each matches the oracle
    in the same binding (indices of 64 particles bit for bit, weights within
    1e-5), while the two bindings' pose matrices differ for some particles."""
    from tsdf_localization_amd import synth
    w = synth.building()
    om = ob.OracleMap.from_host_map(w.map)
    sub = w.particles[::25].copy()
    mats = {}
    for mode in (tsl.TRIG_GLIBC_FMA, tsl.TRIG_DOUBLE):
        ev = tsl.TSDFEvaluator(w.map, trig=mode)
        om_mode = ob.trig_mode_for(mode)
        assert np.array_equal(ev.trace_indices(sub[:64], w.points, w.tf),
                              om.trace(sub[:64], w.points, w.tf, trig_mode=om_mode))
        raw = ev.evaluate_raw(sub, w.points, w.tf)
        _, w64 = om.evaluate(sub, w.points, w.tf, trig_mode=om_mode)
        rel = np.abs(raw.astype(np.float64) - w64) / np.abs(w64)
        assert rel.max() <= WEIGHT_RTOL, (mode, rel.max())
        mats[mode] = tsl.pose_matrices(sub, w.tf, mode)
    assert not np.array_equal(mats[tsl.TRIG_GLIBC_FMA], mats[tsl.TRIG_DOUBLE])
