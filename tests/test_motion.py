"""Motion model (ParticleCloud::apply_model, src/particle_cloud.cpp:499-617):
the oracle is pinned by closed forms, the counter-based draws by their
statistics, and the host helpers by the reference's distribution formulas.
No GPU."""
import math

import numpy as np
import pytest

import oracle_binding as ob
from tsdf_localization_amd import evaluator as tev


def _rot(roll, pitch, yaw):
    cr, sr, cp, sp, cy, sy = math.cos(roll), math.sin(roll), math.cos(pitch), math.sin(pitch), math.cos(yaw), math.sin(yaw)
    Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    Ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def test_zero_noise_pure_translation_follows_heading():
    """A forward step dx moves each particle along its own heading."""
    rng = np.random.default_rng(1)
    P = 200
    parts = np.zeros((P, 7), np.float32)
    parts[:, :3] = rng.uniform(-10, 10, (P, 3))
    parts[:, 5] = rng.uniform(-3, 3, P)
    parts[:, 6] = rng.uniform(0, 1, P)
    out = ob.motion(parts, [0.5, 0, 0, 0, 0, 0], [0] * 6, 7)
    yaw = parts[:, 5].astype(np.float64)
    assert np.allclose(out[:, 0], parts[:, 0] + 0.5 * np.cos(yaw), atol=2e-6)
    assert np.allclose(out[:, 1], parts[:, 1] + 0.5 * np.sin(yaw), atol=2e-6)
    assert np.array_equal(out[:, 2], parts[:, 2])
    assert np.allclose(out[:, 5], yaw, atol=1e-6)
    assert np.array_equal(out[:, 6], parts[:, 6])  # weights carried over


def test_zero_noise_yaw_step_adds_and_wraps():
    parts = np.zeros((3, 7), np.float32)
    parts[:, 5] = [0.1, 3.0, -3.1]
    out = ob.motion(parts, [0, 0, 0, 0, 0, 0.3], [0] * 6, 1)
    expect = np.angle(np.exp(1j * (parts[:, 5].astype(np.float64) + 0.3)))
    assert np.allclose(out[:, 5], expect, atol=1e-6)
    assert np.array_equal(out[:, :3], parts[:, :3])


def test_zero_noise_rotation_composition():
    """New orientation = R(particle) * R(step) read back as roll/pitch/yaw;
    new position = old + R(particle) * t(step)."""
    rng = np.random.default_rng(2)
    P = 300
    parts = np.zeros((P, 7), np.float32)
    parts[:, :3] = rng.uniform(-5, 5, (P, 3))
    parts[:, 3:6] = rng.uniform(-0.6, 0.6, (P, 3))
    step = [0.3, -0.1, 0.05, 0.02, -0.03, 0.2]
    out = ob.motion(parts, step, [0] * 6, 3)
    Rs = _rot(*step[3:])
    for i in range(P):
        Rp = _rot(*parts[i, 3:6].astype(np.float64))
        R = Rp @ Rs
        assert np.allclose(_rot(*out[i, 3:6].astype(np.float64)), R, atol=3e-6)
        assert np.allclose(out[i, :3], parts[i, :3] + Rp @ np.array(step[:3]), atol=3e-6)


def test_gimbal_branch_matches_get_angle_from_mat():
    """|R[2,0]| reaching 1 takes the branch with yaw = 0 (util.cpp:88-103)."""
    parts = np.zeros((1, 7), np.float32)
    out = ob.motion(parts, [0, 0, 0, 0, math.pi / 2, 0], [0] * 6, 1)
    assert out[0, 5] == 0.0
    assert abs(out[0, 4] - math.pi / 2) < 1e-6


def test_draws_are_standard_normal_and_independent():
    n = 40000
    z = np.array([[ob.olib.orc_motion_normal(99, i, k) for k in range(6)] for i in range(n)])
    assert np.all(np.abs(z.mean(0)) < 0.02)
    assert np.all(np.abs(z.std(0) - 1.0) < 0.02)
    c = np.corrcoef(z.T)
    assert np.max(np.abs(c - np.eye(6))) < 0.02
    # tails of a normal: P(|z| > 2) = 0.0455
    assert abs(np.mean(np.abs(z) > 2.0) - 0.0455) < 0.004
    # pure function of (seed, particle, draw)
    assert ob.olib.orc_motion_normal(99, 5, 3) == ob.olib.orc_motion_normal(99, 5, 3)
    assert ob.olib.orc_motion_normal(99, 5, 3) != ob.olib.orc_motion_normal(100, 5, 3)


def test_noise_moves_particles_with_stated_spread():
    P = 20000
    parts = np.zeros((P, 7), np.float32)
    out = ob.motion(parts, [0, 0, 0, 0, 0, 0], [0.1, 0.2, 0.05, 0, 0, 0.01], 5)
    assert abs(out[:, 0].std() - 0.1) < 0.004
    assert abs(out[:, 1].std() - 0.2) < 0.008
    assert abs(out[:, 2].std() - 0.05) < 0.002
    assert abs(out[:, 5].std() - 0.01) < 0.0004


def test_motion_from_delta_formulas():
    """particle_cloud.cpp:471-492: float squares and spreads."""
    a = np.arange(1, 13, dtype=np.float32) * np.float32(0.01)
    m = tev.motion_from_delta(0.5, 0.1, 0.02, 0.01, -0.02, 0.3, 0.05, 0.01, a, seed=4)
    f = np.float32
    dx2, dy2, dz2 = f(0.5) * f(0.5), f(0.1) * f(0.1), f(0.02) * f(0.02)
    r2, p2, y2 = f(0.01) * f(0.01), f(-0.02) * f(-0.02), f(0.3) * f(0.3)
    expect = [a[0] * dx2 + a[1] * y2 + f(0.05), a[2] * dy2 + a[3] * y2 + f(0.05), a[4] * dz2 + a[5] * y2 + f(0.05),
              a[6] * dx2 + a[7] * r2 + f(0.01), a[8] * dx2 + a[9] * p2 + f(0.01), a[10] * dx2 + a[11] * y2 + f(0.01)]
    assert list(m.stddev) == [float(np.float32(e)) for e in expect]
    assert list(m.mean) == [float(np.float32(v)) for v in (0.5, 0.1, 0.02, 0.01, -0.02, 0.3)]
    assert m.seed == 4


def test_motion_from_odom_formulas():
    """particle_cloud.cpp:181-200."""
    a = np.full(12, 0.1, np.float32)
    m = tev.motion_from_odom(1.5, 0.4, 0.1, a)
    f = np.float32
    d, th = f(1.5) * f(0.1), f(0.4) * f(0.1)
    assert m.mean[0] == float(d) and m.mean[5] == float(th)
    assert all(m.mean[k] == 0.0 for k in (1, 2, 3, 4))
    sd = float(f(0.1) * (d * d) + f(0.1) * (th * th))
    assert all(abs(m.stddev[k] - sd) == 0.0 for k in range(6))
