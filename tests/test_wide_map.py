"""The 64-bit map form (int64 block offsets, int16 millimetres), on the CPU:
the oracle's 64-bit lookup pinned by brute force on a small map whose block
offsets lie past 2^31 and 2^32, the 64-bit form of a map equal to its int32
form, the campus generator's blocks equal to its buildings' blocks, and the
library's refusal of what the device form cannot hold."""
import ctypes

import numpy as np
import pytest

import oracle_binding as ob
from tsdf_localization_amd import _lib, synth
from tsdf_localization_amd import evaluator as tev


def test_get_index64_brute_force_past_2_32():
    """2 x 2 x 2 blocks of 1 m at 0.25 m (sub_dim 4): block offsets of
    3.2e9, 4.4e9 and 8.6e9 voxels.  Every voxel centre returns offset + sx +
    4 sy + 16 sz computed with Python integers; empty blocks, points past
    max and below min return data_size."""
    c = ob.coef_init((0, 0, 0), (2, 2, 2), 0.25, 0.0)
    assert (c.sub_dim, c.grid_occ_size) == (4, 8)
    offs = {0: 50_000_000 * 64, 3: 68_750_000 * 64, 6: 134_375_000 * 64}
    grid = np.full(8, -1, np.int64)
    for b, o in offs.items():
        grid[b] = o
    c.data_size = max(offs.values()) + 64
    checked = 0
    for bz in range(2):
        for by in range(2):
            for bx in range(2):
                b = bx + 2 * by + 4 * bz
                for sz in range(4):
                    for sy in range(4):
                        for sx in range(4):
                            x, y, z = bx + 0.125 + 0.25 * sx, by + 0.125 + 0.25 * sy, bz + 0.125 + 0.25 * sz
                            got = ob.olib.orc_get_index64(ctypes.byref(c), ob._ptr(grid), x, y, z)
                            want = offs[b] + sx + 4 * sy + 16 * sz if b in offs else int(c.data_size)
                            assert got == want, (x, y, z)
                            checked += 1
    assert checked == 512
    for p in ((2.1, 0.5, 0.5), (0.5, 0.5, 2.5), (-1.5, 0.5, 0.5)):
        assert ob.olib.orc_get_index64(ctypes.byref(c), ob._ptr(grid), *p) == c.data_size


def _wide_form(m):
    g = m.grid_occ.astype(np.int64)
    return tev.CudaSubVoxelMap(m.coef(), g, tsdf_mm=m.tsdf_mm, truncation_mm=m.truncation_mm, sigma=m.sigma)


def test_wide_form_of_a_map_equals_its_reference_form():
    """The room map with its block table widened to int64: the oracle's
    64-bit form (voxels read through the likelihood values) gives the same
    indices and bit-identical weights as its reference form (float voxels)."""
    w = synth.room()
    om = ob.OracleMap.from_host_map(w.map)
    om64 = ob.OracleMap64.from_host_map(_wide_form(w.map))
    assert np.array_equal(om.trace(w.particles[:50], w.points, w.tf), om64.trace(w.particles[:50], w.points, w.tf))
    a32, a64 = om.evaluate(w.particles, w.points, w.tf)
    b32, b64 = om64.evaluate(w.particles, w.points, w.tf)
    assert np.array_equal(a32.view(np.uint32), b32.view(np.uint32))
    assert np.array_equal(a64.view(np.uint64), b64.view(np.uint64))


def test_campus_blocks_are_its_buildings_blocks():
    """A 3 x 2 campus of two buildings: campus block (bx, by, bz) is
    allocated exactly when block (bx mod 52, by mod 52, bz) of its tile's
    building is, with the same 8,000 millimetres, and offsets follow linear
    block order (setData's), past nothing here but the same rule as the
    full-size campus."""
    cp = synth.campus(11, tiles=(3, 2), bases=2)
    c = cp.map.coef()
    sd3 = int(c.sub_dim) ** 3
    bases = [sc.tsdf_map((0, 0, 0), (51.2, 51.2, 6.4), 0.05, 300, dense=False) for sc in cp.scenes]
    UX, UY, UZ = int(c.up_dim_x), int(c.up_dim_y), int(c.up_dim_z)
    g = cp.map.grid_occ
    alloc = g >= 0
    assert np.array_equal(g[alloc], np.arange(int(alloc.sum())) * sd3)
    blocks = cp.map.tsdf_mm.reshape(-1, sd3)
    for bz in range(UZ):
        for by in range(UY):
            for bx in range(UX):
                up = bx + by * UX + bz * UX * UY
                base = bases[cp.tile_of[by // 52, bx // 52]]
                bo = base.grid_occ[(bx % 52) + (by % 52) * 52 + bz * 52 * 52]
                assert (g[up] >= 0) == (bo >= 0), (bx, by, bz)
                if bo >= 0:
                    assert np.array_equal(blocks[g[up] // sd3], base.tsdf_mm[bo: bo + sd3]), (bx, by, bz)


def test_wide_create_refuses_what_the_device_form_cannot_hold():
    c = tev.coef_init((0, 0, 0), (2, 2, 2), 0.25, 0.0)
    grid = np.full(8, -1, np.int64)
    grid[1] = 65  # not a multiple of sub_dim^3 = 64
    c.data_size = 192
    mm = np.zeros(192, np.int16)
    h = ctypes.c_void_p()
    rc = _lib.lib.tsl_map_create_tsdf_wide(ctypes.byref(c), _lib.ptr(grid), _lib.ptr(mm), 300, 0.1, ctypes.byref(h))
    assert rc == _lib.TSL_ERR_UNSUPPORTED and b"multiple" in _lib.lib.tsl_last_error()
    grid[1] = 192  # block past the voxel array
    rc = _lib.lib.tsl_map_create_tsdf_wide(ctypes.byref(c), _lib.ptr(grid), _lib.ptr(mm), 300, 0.1, ctypes.byref(h))
    assert rc == _lib.TSL_ERR_UNSUPPORTED
    c2 = tev.coef_init((0, 0, 0), (2, 2, 2), 0.2, 0.0)  # sub_dim 5
    g2 = np.full(int(c2.grid_occ_size), -1, np.int64)
    rc = _lib.lib.tsl_map_create_tsdf_wide(ctypes.byref(c2), _lib.ptr(g2), _lib.ptr(mm), 300, 0.1, ctypes.byref(h))
    assert rc == _lib.TSL_ERR_UNSUPPORTED and b"multiple of 4" in _lib.lib.tsl_last_error()
