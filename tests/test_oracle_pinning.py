"""Pins the CPU oracle to things that do not depend on the oracle itself:
the native gcc/x86-64 behaviour it restates, closed forms, brute force on
tiny hand-built inputs, and textbook algorithms in exact arithmetic.

The reference ships no tests or fixtures for this path, so these checks are
what fixes the oracle; each is chosen so that a plausible slip (a dropped
term, a swapped index or sign, a transposed operand) fails it.
"""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_binding as ob

libm = ctypes.CDLL("libm.so.6")
libm.sinf.restype = ctypes.c_float
libm.sinf.argtypes = [ctypes.c_float]
libm.cosf.restype = ctypes.c_float
libm.cosf.argtypes = [ctypes.c_float]


# ----------------------------------------------------------------- conversions

def test_float_to_size_t_matches_gcc_x86():
    """orc_cvt_f32_u64 spells out (size_t)f; it must equal what this gcc
    emits for x86-64 on every input, including the range C++ leaves undefined."""
    rng = np.random.default_rng(1)
    bits = rng.integers(0, 2**32, 300_000, dtype=np.uint64).astype(np.uint32)
    special = np.array([0.0, -0.0, 0.5, -0.5, -0.99999994, -1.0, -1.5, -5.3, 1e19, 9.3e18, 1.8446744e19, 2e19,
                        -9.3e18, -1e30, 3.4e38, np.inf, -np.inf, np.nan, 4294967296.0, 8388607.5],
                       np.float32).view(np.uint32)
    for b in np.concatenate([bits, special]):
        x = float(np.uint32(b).view(np.float32))
        assert ob.olib.orc_cvt_f32_u64(x) == ob.olib.orc_native_cvt_f32_u64(x), hex(int(b))


def test_size_t_to_float_rounds_to_nearest_even():
    for v in [0, 1, 2**24 + 1, 2**24 + 3, 2**53 + 1, 2**63, 2**64 - 1, 2**64 - 2**40, 12345678901234567]:
        got = np.float32(ob.olib.orc_cvt_u64_f32(v))
        exact = Fraction(v)
        # nearest float by construction from exact rational comparison
        lo = np.float32(float(v))
        cands = [lo, np.nextafter(lo, np.float32(np.inf)), np.nextafter(lo, np.float32(0))]
        best = min(cands, key=lambda c: (abs(Fraction(float(c)) - exact), int(np.float32(c).view(np.uint32)) & 1))
        assert got == best, v


# ----------------------------------------------------------------- sin / cos

def test_sincos_special_values():
    assert ob.sincosf(0.0) == (np.float32(0.0), np.float32(1.0))
    s, c = ob.sincosf(-0.0)
    assert s == 0.0 and np.signbit(s) and c == 1.0
    s, c = ob.sincosf(np.float32(math.pi / 2))
    assert c == np.float32(math.cos(float(np.float32(math.pi / 2))))
    assert s == np.float32(1.0)
    s, c = ob.sincosf(np.float32(math.pi / 6))
    assert abs(s - 0.5) < 1e-7
    s, c = ob.sincosf(np.float32(2e6))
    assert np.isnan(s) and np.isnan(c)


def test_sincos_agrees_with_double_libm_rounded():
    """The shared sincos equals (float)sin((double)x) (the reference's sin
    with the double overload) on this sample; its mismatch fraction against
    glibc's own sinf is measured and bounded."""
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(-7, 7, 60_000), rng.normal(0, 0.1, 20_000),
                         rng.uniform(-3000, 3000, 10_000)]).astype(np.float32)
    bad_d = bad_f = 0
    for x in xs:
        s, c = ob.sincosf(x)
        if s != np.float32(math.sin(float(x))) or c != np.float32(math.cos(float(x))):
            bad_d += 1
        if s != np.float32(libm.sinf(float(x))) or c != np.float32(libm.cosf(float(x))):
            bad_f += 1
    assert bad_d / xs.size <= 1e-5
    assert bad_f / xs.size < 0.05  # glibc sinf is not correctly rounded (~1.4% here)


# ----------------------------------------------------------------- pose matrix

def _closed_form_L(pose, tf):
    x, y, z, r, p, yw = [float(v) for v in pose]
    Rx = np.array([[1, 0, 0], [0, math.cos(r), -math.sin(r)], [0, math.sin(r), math.cos(r)]])
    Ry = np.array([[math.cos(p), 0, math.sin(p)], [0, 1, 0], [-math.sin(p), 0, math.cos(p)]])
    Rz = np.array([[math.cos(yw), -math.sin(yw), 0], [math.sin(yw), math.cos(yw), 0], [0, 0, 1]])
    M = np.eye(4)
    M[:3, :3] = Rz @ Ry @ Rx
    M[:3, 3] = (x, y, z)
    return (M @ np.asarray(tf, np.float64).reshape(4, 4))[:3].reshape(-1)


def test_pose_matrix_identity_particle_gives_tf():
    tf = np.array([0.9, -0.1, 0.2, 0.5, 0.1, 0.95, 0.0, -0.2, -0.2, 0.0, 0.97, 0.4, 0, 0, 0, 1], np.float32)
    L = ob.pose_matrix(np.zeros(6, np.float32), tf)
    assert np.array_equal(L, tf[:12])


def test_pose_matrix_closed_form():
    rng = np.random.default_rng(3)
    tf = np.eye(4, dtype=np.float32).reshape(-1)
    tf[3], tf[7], tf[11] = 0.1, -0.05, 0.4
    for _ in range(300):
        pose = np.concatenate([rng.uniform(-50, 50, 3), rng.uniform(-math.pi, math.pi, 3)]).astype(np.float32)
        L = ob.pose_matrix(pose, tf).astype(np.float64)
        ref = _closed_form_L(pose, tf)
        assert np.allclose(L[[0, 1, 2, 4, 5, 6, 8, 9, 10]], ref[[0, 1, 2, 4, 5, 6, 8, 9, 10]], atol=2e-7)
        assert np.allclose(L[[3, 7, 11]], ref[[3, 7, 11]], atol=1e-5)
        R = L.reshape(3, 4)[:, :3]
        assert abs(np.linalg.det(R) - 1.0) < 1e-5
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-6)


def test_pose_matrix_rejects_wrong_rotation_order():
    """Guard: Rx*Ry*Rz (the other common convention) must not match."""
    pose = np.array([0, 0, 0, 0.4, -0.3, 1.1], np.float32)
    tf = np.eye(4, dtype=np.float32).reshape(-1)
    L = ob.pose_matrix(pose, tf).astype(np.float64).reshape(3, 4)[:, :3]
    r, p, y = 0.4, -0.3, 1.1
    Rx = np.array([[1, 0, 0], [0, math.cos(r), -math.sin(r)], [0, math.sin(r), math.cos(r)]])
    Ry = np.array([[math.cos(p), 0, math.sin(p)], [0, 1, 0], [-math.sin(p), 0, math.cos(p)]])
    Rz = np.array([[math.cos(y), -math.sin(y), 0], [math.sin(y), math.cos(y), 0], [0, 0, 1]])
    assert not np.allclose(L, Rx @ Ry @ Rz, atol=1e-3)
    assert np.allclose(L, Rz @ Ry @ Rx, atol=1e-6)


# ----------------------------------------------------------------- map lookup

def test_coef_init_hand_values():
    c = ob.coef_init((0, 0, 0), (12.8, 12.8, 3.2), 0.05, 0.0)
    assert (c.dim_x, c.dim_y, c.dim_z) == (256, 256, 64)
    assert (c.up_dim_x, c.up_dim_y, c.up_dim_z) == (13, 13, 4)
    assert c.up_dim_2 == 169 and c.sub_dim == 20 and c.sub_dim_2 == 400
    assert c.grid_occ_size == 13 * 13 * 4 and c.data_size == 0
    c = ob.coef_init((-1.024, 0, 0), (3.072, 2.048, 1.024), 0.064, 0.0)
    assert (c.dim_x, c.dim_y, c.dim_z) == (64, 32, 16)
    assert (c.up_dim_x, c.up_dim_y, c.up_dim_z) == (5, 3, 2)
    assert c.sub_dim == 16  # ceil(1 / 0.064) = ceil(15.625)


def _tiny_map():
    """2x2x2 one-metre blocks at 0.25 m (4^3 voxels), blocks 0, 3 and 6
    allocated in that linear order, every voxel value distinct."""
    c = ob.coef_init((0, 0, 0), (2, 2, 2), 0.25, -7.0)
    grid = np.full(c.grid_occ_size, -1, np.int32)
    alloc = [0, 3, 6]
    for n, b in enumerate(alloc):
        grid[b] = n * 64
    c.data_size = 3 * 64
    data = (np.arange(c.data_size, dtype=np.float32) + 1) * 0.5
    return c, grid, data, alloc


def test_lookup_brute_force_voxel_centres():
    c, grid, data, alloc = _tiny_map()
    om = ob.OracleMap(c, grid, data)
    for b in range(8):
        bx, by, bz = b % 2, (b // 2) % 2, b // 4
        for v in range(64):
            vx, vy, vz = v % 4, (v // 4) % 4, v // 16
            p = (bx + (vx + 0.5) * 0.25, by + (vy + 0.5) * 0.25, bz + (vz + 0.5) * 0.25)
            idx = om.get_index(*p)
            if b in alloc:
                expect = alloc.index(b) * 64 + vx + 4 * vy + 16 * vz
                assert idx == expect
                assert om.get_entry(*p) == data[expect]
            else:
                assert idx == c.data_size
                assert om.get_entry(*p) == np.float32(-7.0)


def test_lookup_outside_high_side_is_rejected():
    c, grid, data, _ = _tiny_map()
    om = ob.OracleMap(c, grid, data)
    for p in [(2.01, 0.5, 0.5), (0.5, 2.2, 0.5), (0.5, 0.5, 9.0), (1e6, 0.1, 0.1)]:
        assert om.get_index(*p) == c.data_size
        assert om.get_entry(*p) == np.float32(-7.0)


def _x86_u64(f):
    """(size_t)f on x86-64 for finite float32 f, independent restatement."""
    f = float(np.float32(f))
    if f >= 2.0**63:
        g = float(np.float32(np.float32(f) - np.float32(2.0**63)))
        return (int(g) if -2.0**63 <= g < 2.0**63 else -2**63) % 2**64 ^ 2**63
    return (int(f) if -2.0**63 <= f < 2.0**63 else -2**63) % 2**64


def test_lookup_negative_band_wraps_like_x86():
    """Offsets in (-1, 0) and below -1: the reference converts negative floats
    to size_t; on x86-64 that wraps.  Hand-derived from the reference's
    arithmetic with exact float32 steps in numpy."""
    c, grid, data, _ = _tiny_map()
    om = ob.OracleMap(c, grid, data)
    res = np.float32(0.25)
    for (x, y, z) in [(-0.3, 0.6, 0.1), (0.6, -0.6, 0.3), (0.1, 0.2, -0.9), (1.2, -1.5, 0.4), (-2.5, 1.4, 1.4),
                      (0.4, 1.6, -3.2), (-0.01, -0.02, -0.03)]:
        o = np.array([x, y, z], np.float32)
        g = [_x86_u64(v) for v in o]
        q = [_x86_u64(np.float32(np.float32(gg) / res)) for gg in g]
        if q[0] >= c.dim_x or q[1] >= c.dim_y or q[2] >= c.dim_z:
            expect = c.data_size
        else:
            sp = [np.float32(o[a] - np.float32(g[a])) for a in range(3)]
            up = (g[0] + g[1] * c.up_dim_x + g[2] * c.up_dim_2) % 2**64
            if up >= c.grid_occ_size or grid[up] < 0:
                expect = c.data_size
            else:
                s = [_x86_u64(np.float32(sp[a] / res)) for a in range(3)]
                expect = (int(grid[up]) + s[0] + s[1] * c.sub_dim + s[2] * c.sub_dim_2) % 2**64
        assert om.get_index(x, y, z) == expect, (x, y, z)


def test_set_data_places_cells_like_the_reference():
    """Cells at exact voxel corners (res 0.25) land in their own voxel; block
    offsets follow linear block order (cuda_sub_voxel_map.tcc:193-211)."""
    cells = np.array([[0.0, 0.0, 0.0], [1.75, 0.25, 0.5], [0.5, 1.5, 1.25], [1.0, 1.0, 1.0]], np.float32)
    vals = np.array([1.0, 2.0, 3.0, 4.0], np.float32)
    c, grid, data = ob.set_data((0, 0, 0), (2, 2, 2), 0.25, -1.0, cells, vals)
    # blocks touched: (0,0,0)=0, (1,0,0)=1, (0,1,1)=6, (1,1,1)=7 -> offsets in that order
    expect_grid = np.full(8, -1, np.int32)
    for n, b in enumerate([0, 1, 6, 7]):
        expect_grid[b] = n * 64
    assert np.array_equal(grid, expect_grid)
    assert c.data_size == 256
    assert data[0] == 1.0
    assert data[64 + 3 + 4 * 1 + 16 * 2] == 2.0
    assert data[128 + 2 + 4 * 2 + 16 * 1] == 3.0
    assert data[192] == 4.0
    assert np.count_nonzero(data == -1.0) == 256 - 4


# ----------------------------------------------------------------- likelihood

def test_likelihood_closed_form():
    sigma = 0.1
    peak = (1.0 / math.sqrt(2 * math.pi * sigma * sigma)) ** 3
    v0 = ob.olib.orc_likelihood_value(0, sigma)
    assert abs(v0 - peak) / peak < 1e-5 and abs(v0 - 63.4936) < 1e-3
    for mm in (1, 37, 150, 299):
        d = mm * 1e-3
        exact = (math.exp(-d * d / (sigma * sigma) / 2) / math.sqrt(2 * math.pi * sigma * sigma)) ** 3
        got = ob.olib.orc_likelihood_value(mm, sigma)
        assert abs(got - exact) / exact < 1e-5
        assert got == ob.olib.orc_likelihood_value(-mm, sigma)
    vals = [ob.olib.orc_likelihood_value(mm, sigma) for mm in range(0, 600)]
    assert all(a >= b for a, b in zip(vals, vals[1:]))
    assert ob.olib.orc_likelihood_init(sigma) == 0.0


# ----------------------------------------------------------------- score and sum

def test_score_closed_forms():
    c, grid, data, alloc = _tiny_map()
    om = ob.OracleMap(c, grid, data)
    tf = np.eye(4, dtype=np.float32).reshape(-1)
    part = np.zeros((1, 7), np.float32)
    # points at voxel centres of block 3 (x in [1,2), y in [1,2), z in [0,1))
    rng = np.random.default_rng(4)
    vox = rng.integers(0, 64, 50)
    pts = np.stack([1 + (vox % 4 + 0.5) * 0.25, 1 + ((vox // 4) % 4 + 0.5) * 0.25, ((vox // 16) + 0.5) * 0.25],
                   1).astype(np.float32)
    _, w64 = om.evaluate(part, pts, tf, a_hit=1.0, a_range=0.0, a_max=0.0, max_range=100.0)
    assert w64[0] == float(np.sum(data[64 + vox].astype(np.float64)))
    # empty space far away, inside max range: every term is a_range / max_range
    far = np.full((40, 3), 50.0, np.float32)
    _, w64 = om.evaluate(part, far, tf, a_hit=0.9, a_range=0.1, a_max=0.0, max_range=100.0)
    term = float(np.float32(np.float32(0.9) * np.float32(-7.0)) + np.float32(np.float32(0.1) * np.float32(1.0 / 100.0)))
    assert w64[0] == 40 * term
    # beyond max range the a_max term is used instead
    _, w64 = om.evaluate(part, far, tf, a_hit=0.0, a_range=0.1, a_max=0.25, max_range=10.0)
    assert w64[0] == 40 * 0.25


# ----------------------------------------------------------------- normalisation

def test_normalise_one_hot_and_sum():
    rng = np.random.default_rng(5)
    parts = rng.uniform(-3, 3, (50, 7)).astype(np.float32)
    w = np.zeros(50, np.float32)
    w[17] = 3.5
    wn, pose = ob.normalize(parts, w)
    assert wn[17] == 1.0 and np.count_nonzero(wn) == 1
    assert np.allclose(pose[:3], parts[17, :3], atol=1e-6)
    assert np.allclose(pose[3:], parts[17, 3:6], atol=1e-6)
    w = rng.uniform(0, 1, 50).astype(np.float32)
    wn, _ = ob.normalize(parts, w)
    assert abs(float(np.sum(wn.astype(np.float64))) - 1.0) < 1e-6
    with pytest.raises(RuntimeError, match="No particle is valid!"):
        ob.normalize(parts, np.zeros(50, np.float32))


# ----------------------------------------------------------------- resampling

def _textbook_systematic(w, u0):
    """Kitagawa (1996) systematic resampling in exact arithmetic:
    U_j = u0 + j/M, output j = min{m : sum_{i<=m} w_i > U_j}."""
    M = len(w)
    cs = np.cumsum([Fraction(float(x)) for x in w])
    out, j = [], 0
    u = Fraction(float(u0))
    total = cs[-1]
    while True:
        uj = u + Fraction(j, M)
        if not (total > uj):
            break
        m = next(i for i in range(M) if cs[i] > uj)
        out.append(m)
        j += 1
    return out


@pytest.mark.parametrize("M", [1, 5, 37, 200])
def test_resample_matches_textbook_for_small_m(M):
    rng = np.random.default_rng(M)
    compared = 0
    for trial in range(20):
        w = rng.gamma(0.7, 1.0, M)
        w = (w / w.sum()).astype(np.float32)
        u0 = np.float32(rng.uniform(0, 1.0 / M))
        got = ob.resample(w, u0).tolist()
        ref = _textbook_systematic(w, u0)
        # the float U recurrence drifts by far less than the gap to any
        # cumulative-sum boundary in these draws; if not, skip the draw
        cs = np.cumsum(w.astype(np.float64))
        uj = float(u0) + np.arange(len(ref) + 2) / M
        gap = np.min(np.abs(cs[:, None] - uj[None, :]))
        if gap < 1e-6:
            continue
        assert got == ref
        compared += 1
    assert compared >= 15


def test_resample_uniform_and_one_hot():
    M = 64
    w = np.full(M, 1.0 / M, np.float32)
    got = ob.resample(w, np.float32(0.3 / M))
    assert got.tolist() == list(range(M))
    w = np.zeros(M, np.float32)
    w[9] = 1.0
    got = ob.resample(w, np.float32(0.5 / M))
    assert set(got.tolist()) == {9} and got.size == M


def test_resample_large_m_float_recurrence_drifts():
    """With M = 1e5 the float U recurrence drifts by many steps: the count
    emitted differs from M, as in the reference; indices stay sorted."""
    M = 100_000
    rng = np.random.default_rng(9)
    w = rng.uniform(0, 1, M)
    w = (w / w.sum()).astype(np.float32)
    out = ob.resample(w, np.float32(0.5 / M))
    assert out.size != M
    assert np.all(np.diff(out) >= 0)
    U = ob.systematic_u(M, np.float32(0.5 / M), out.size + 1)
    assert np.all(np.diff(U.astype(np.float64)) > 0)
    s_final = ob.cumsum(w)[-1]
    assert U[out.size - 1] < s_final <= U[out.size]
