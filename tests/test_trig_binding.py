"""Which sin/cos the pose matrices use, and how much that matters.

The reference's CPU evaluator calls unqualified sin/cos on float angles
(src/evaluation/tsdf_evaluator.cpp:108-113, 206-213): std::sin(float), the C
library's sinf.  glibc (2.28+) has one sinf/cosf algorithm built twice for
x86-64, with FMA contraction (picked on CPUs with FMA3) and without; the
library restates both (TSL_TRIG_GLIBC_FMA, the default, and
TSL_TRIG_GLIBC_NOFMA) and offers a third binding, (float)sin((double)x)
(TSL_TRIG_DOUBLE).  These tests pin the restatements against glibc itself,
bit for bit on every float, and measure on the benchmark workloads how far
the bindings move voxel indices and weights.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_binding as ob
import tsdf_localization_amd as tsl
from tsdf_localization_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
SPLIT = os.path.join(HERE, "golden", "glibc_sinf_fma_split.json")
TUNABLES = "glibc.cpu.hwcaps=-FMA,-AVX2"


def _host_has_fma():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " fma " in f" {line.split(':', 1)[1].strip()} "
    except OSError:
        pass
    return False


def _split():
    with open(SPLIT) as f:
        rows = json.load(f)["inputs"]
    x = np.array([int(r["x"], 16) for r in rows], np.uint32).view(np.float32)
    fma = np.array([[int(v, 16) for v in r["fma"]] for r in rows], np.uint32)
    nofma = np.array([[int(v, 16) for v in r["nofma"]] for r in rows], np.uint32)
    return x, fma, nofma


def test_split_inputs_separate_the_builds():
    """The fixture's 34 inputs are where glibc's two builds disagree; the
    library's host build of each variant gives that build's glibc outputs."""
    x, fma, nofma = _split()
    assert x.size == 34
    assert np.all((fma != nofma).any(axis=1))
    for mode, want in ((tsl.TRIG_GLIBC_FMA, fma), (tsl.TRIG_GLIBC_NOFMA, nofma)):
        s, c = tsl.trig(x, mode, on_device=False)
        assert np.array_equal(s.view(np.uint32), want[:, 0])
        assert np.array_equal(c.view(np.uint32), want[:, 1])


@pytest.mark.skipif(not _host_has_fma(), reason="glibc selects its FMA build only on CPUs with FMA3")
def test_glibc_fma_build_every_float():
    """TSL_TRIG_GLIBC_FMA == this process's sinf/cosf on all 4,278,190,082
    non-NaN floats (glibc selected its FMA build on this CPU)."""
    bad, n = tsl.check_trig_libm(tsl.TRIG_GLIBC_FMA, 1)
    assert n == 4278190082
    assert bad == 0


def test_glibc_nofma_build_every_float():
    """TSL_TRIG_GLIBC_NOFMA == sinf/cosf of a process in which glibc was told
    (GLIBC_TUNABLES) not to use FMA, on every non-NaN float; that process's
    sinf/cosf give the fixture's no-FMA outputs, so the tunable took effect."""
    x, _, nofma = _split()
    code = ("import ctypes, json, sys, numpy as np, tsdf_localization_amd as t\n"
            "m = ctypes.CDLL('libm.so.6')\n"
            "m.sinf.restype = m.cosf.restype = ctypes.c_float\n"
            "m.sinf.argtypes = m.cosf.argtypes = [ctypes.c_float]\n"
            "x = np.array(json.loads(sys.argv[1]), np.uint32).view(np.float32)\n"
            "sc = [[int(np.float32(f(float(v))).view(np.uint32)) for f in (m.sinf, m.cosf)] for v in x]\n"
            "print(json.dumps([sc, t.check_trig_libm(t.TRIG_GLIBC_NOFMA, 1)]))\n")
    env = dict(os.environ, GLIBC_TUNABLES=TUNABLES, PYTHONPATH=os.path.dirname(HERE))
    r = subprocess.run([sys.executable, "-c", code, json.dumps([int(v) for v in x.view(np.uint32)])], env=env,
                       capture_output=True, text=True, timeout=600, check=True)
    sc, (bad, n) = json.loads(r.stdout.strip().splitlines()[-1])
    assert np.array_equal(np.array(sc, np.uint32), nofma)
    assert n == 4278190082
    assert bad == 0


def test_double_binding_is_the_shared_routine():
    """TSL_TRIG_DOUBLE is the oracle's pinned double routine (orc_sincosf)."""
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-7, 7, 20000), rng.uniform(-1e3, 1e3, 2000)]).astype(np.float32)
    s, c = tsl.trig(x, tsl.TRIG_DOUBLE, on_device=False)
    for i in range(0, x.size, 37):
        rs, rc = ob.sincosf(x[i])
        assert s[i] == rs and c[i] == rc


def _binding_diff(w, particles):
    om = ob.OracleMap.from_host_map(w.map)
    return om.trig_binding_diff(particles, w.points, w.tf, ob.TRIG_LIBM, ob.TRIG_DOUBLE)


def test_binding_mismatch_room():
    """Plain update of the room (1,000 particles x 2,056 points): the two
    bindings give different pose matrices for 44 particles but the same voxel
    index for every pair, hence the same weights."""
    w = synth.room()
    bad, badm, worst = _binding_diff(w, w.particles)
    pairs = w.particles.shape[0] * w.points.shape[0]
    print(f"room: {badm} matrices differ, {bad}/{pairs} indices differ, max weight rel diff {worst:.2e}")
    assert bad / pairs <= 1e-6
    assert badm > 0


def test_binding_mismatch_global_localisation():
    """Global localisation (all 100,000 particles x 20,000 points): the
    index mismatch fraction between the bindings is within 1e-6 (206 of 2e9
    pairs), but the weights of the particles concerned move by up to 4.4e-4
    relative, past the 1e-5 bar -- so the binding has to be the reference's."""
    w = synth.building()
    bad, badm, worst = _binding_diff(w, w.particles)
    pairs = w.particles.shape[0] * w.points.shape[0]
    print(f"global localisation: {badm} matrices differ, {bad}/{pairs} indices differ "
          f"({bad / pairs:.2e}), max weight rel diff {worst:.2e}")
    assert bad / pairs <= 1e-6
    assert worst > 1e-5


def test_motion_inverse_trig_every_float():
    """The motion model's getAngleFromMat (src/util/util.cpp:86-111) takes
    floats, so with libstdc++ its asin/atan2/cos are asinf/atan2f/cosf.  The
    library's restatements of glibc's asinf and atanf equal this process's on
    every non-NaN float, its atan2f on 2e7 seeded random pairs (half of them
    with exponents in [2^-27, 2^28])."""
    bad, n = tsl.evaluator.check_invtrig_libm(1, 20_000_000)
    assert n == 2 * 4278190082 + 20_000_000
    assert bad == 0


def test_motion_angle_binding_matters():
    """The float binding (the reference's) and the C binding of
    getAngleFromMat give different angles for a large share of particles
    after one motion step of the room's particle set, so the motion model
    has to follow the float binding too."""
    w = synth.room()
    mean = [0.4, 0.05, 0.01, 0.01, -0.02, 0.15]
    sd = [0.02, 0.02, 0.01, 0.01, 0.01, 0.01]
    a = ob.motion(w.particles, mean, sd, 1234, trig_mode=ob.TRIG_LIBM)
    b = ob.motion(w.particles, mean, sd, 1234, trig_mode=ob.TRIG_DOUBLE)
    differ = np.mean(np.any(a[:, 3:6] != b[:, 3:6], axis=1))
    print(f"particles whose angles differ between the bindings: {differ:.3f}")
    assert differ > 0.05
    assert np.max(np.abs(a[:, 3:6] - b[:, 3:6])) < 1e-5
