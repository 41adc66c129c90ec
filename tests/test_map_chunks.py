"""Map chunks -> cells (createTSDFMap, include/tsdf_localization/map/map_util.h:17-154).

The oracle's restatement is pinned by hand-computed bounds and positions, a
crafted chunk for the value / weight split and the 600 mm truncation edge,
and a single-voxel chunk for the x-major layout; the product's host function
(tsl_map_chunks_to_cells, then tsl_map_set_data) must equal the oracle bit
for bit.  CPU only: no kernel runs here.
"""
import ctypes

import numpy as np
import pytest

import oracle_binding as ob

VOX = 64 ** 3


def word(value, weight):
    return np.uint32((int(value) & 0xFFFF) | ((int(weight) & 0xFFFF) << 16))


def empty_chunks(n):
    return np.zeros((n, VOX), np.uint32)


def f32(x):
    return np.float32(x)


def test_oracle_bounds_hand_computed():
    # chunk coordinates (-1, 0, 2) and (3, -2, 0): one chunk spans 64 voxels of
    # 64 mm = 4.096 m; min is a chunk corner, max the far side of the last chunk
    pos = np.array([[-1, 0, 2], [3, -2, 0]], np.int32)
    c, cells, vals, free = ob.chunks_to_cells(pos, empty_chunks(2))
    assert (c.min_x, c.min_y, c.min_z) == (f32(-4.096), f32(-8.192), f32(0.0))
    assert (c.max_x, c.max_y, c.max_z) == (f32(16.384), f32(4.096), f32(12.288))
    assert c.resolution == f32(0.064)
    assert c.init_value == ob.olib.orc_likelihood_init(0.1)
    assert cells.shape == (0, 3) and free.shape == (0, 3)
    # every box holds the origin: the bounds start at 0 (map_util.h:22-23)
    c, *_ = ob.chunks_to_cells(np.array([[2, 3, 1]], np.int32), empty_chunks(1))
    assert (c.min_x, c.min_y, c.min_z) == (0.0, 0.0, 0.0)
    assert (c.max_x, c.max_y, c.max_z) == (f32(12.288), f32(16.384), f32(8.192))


def test_oracle_value_weight_split_and_truncation():
    raw = empty_chunks(1)
    cases = [(599, 1, "cell"), (600, 1, "free"), (-599, -3, "cell"), (-600, 5, "free"), (0, 0, None),
             (100, 0, None), (-32768, 1, "free"), (0, 32767, "cell"), (32767, -32768, "free")]
    for z, (v, w, _) in enumerate(cases):
        raw[0, z] = word(v, w)
    c, cells, vals, free = ob.chunks_to_cells(np.zeros((1, 3), np.int32), raw, sigma=0.2)
    want_cells = [z for z, (_, _, k) in enumerate(cases) if k == "cell"]
    want_free = [z for z, (_, _, k) in enumerate(cases) if k == "free"]
    assert cells.shape[0] == len(want_cells) and free.shape[0] == len(want_free)
    # chunk (0, 0, 0), flat index z -> voxel (0, 0, z) at z * 0.064 m
    assert np.array_equal(cells, np.array([[0, 0, f32(z * 64 / 1000)] for z in want_cells], np.float32))
    assert np.array_equal(free, np.array([[0, 0, f32(z * 64 / 1000)] for z in want_free], np.float32))
    want_vals = [ob.olib.orc_likelihood_value(cases[z][0], 0.2) for z in want_cells]
    assert np.array_equal(vals, np.array(want_vals, np.float32))
    # the likelihood peaks at tsdf 0 and is symmetric
    assert vals[2] > vals[0] and vals[0] == vals[1]


def test_oracle_layout_is_x_major_with_chunk_offset():
    raw = empty_chunks(2)
    raw[1, 4096 * 1 + 64 * 2 + 3] = word(10, 1)
    c, cells, vals, free = ob.chunks_to_cells(np.array([[0, 0, 0], [5, -1, 0]], np.int32), raw)
    # voxel (64*5 + 1, -64 + 2, 3)
    assert np.array_equal(cells, np.array([[f32(321 * 64 / 1000), f32(-62 * 64 / 1000), f32(3 * 64 / 1000)]]))
    assert free.shape == (0, 3)


def test_oracle_positions_closed_form():
    # (float)x * 64 * 0.001 for a row of voxels along x, both signs
    raw = empty_chunks(4)
    for k in range(4):
        raw[k, np.arange(64) * 4096] = word(0, 1)
    pos = np.array([[-3, 0, 0], [-1, 0, 0], [0, 0, 0], [7, 0, 0]], np.int32)
    _, cells, _, _ = ob.chunks_to_cells(pos, raw)
    xs = np.concatenate([64 * p + np.arange(64) for p in (-3, -1, 0, 7)])
    assert np.array_equal(cells[:, 0], (xs.astype(np.float64) * 64 / 1000).astype(np.float32))


def random_chunks(seed, n):
    rng = np.random.default_rng(seed)
    pos = rng.integers(-4, 5, size=(n, 3)).astype(np.int32)
    value = rng.integers(-1200, 1200, size=(n, VOX))
    value[:, ::97] = rng.choice([-600, -599, 599, 600], size=value[:, ::97].shape)
    weight = np.where(rng.random((n, VOX)) < 0.4, rng.integers(-5, 30, size=(n, VOX)), 0)
    raw = ((value & 0xFFFF) | ((weight & 0xFFFF) << 16)).astype(np.uint32)
    return pos, raw


COEF_FIELDS = [f for f, _ in ob.OrcCoef._fields_ if f != "data_size"]


@pytest.mark.parametrize("seed,n", [(1, 1), (2, 3)])
def test_product_matches_oracle(seed, n):
    from tsdf_localization_amd import mapfile
    pos, raw = random_chunks(seed, n)
    oc, ocells, ovals, ofree = ob.chunks_to_cells(pos, raw, sigma=0.15)
    pc, pcells, pvals, pfree = mapfile.chunks_to_cells(pos, raw, sigma=0.15)
    for f in COEF_FIELDS:
        assert getattr(pc, f) == getattr(oc, f), f
    assert ocells.shape[0] > 0 and ofree.shape[0] > 0
    assert np.array_equal(pcells.view(np.uint32), ocells.view(np.uint32))
    assert np.array_equal(pvals.view(np.uint32), ovals.view(np.uint32))
    assert np.array_equal(pfree.view(np.uint32), ofree.view(np.uint32))


def test_create_tsdf_map_matches_oracle_set_data():
    from tsdf_localization_amd import mapfile
    pos, raw = random_chunks(5, 2)
    pos[1] = pos[0] + np.array([1, 0, 0])  # neighbouring chunks share a block face
    # names out of order, one with blanks and a plus sign: the map follows
    # the group's name order whatever order the mapping has
    tags = [f"{p[0]}_{p[1]}_{p[2]}" for p in pos]
    tags[1] = f" {pos[1][0]}_+{pos[1][1]}_{pos[1][2]}" if pos[1][1] >= 0 else tags[1]
    chunks = {tags[1]: raw[1], tags[0]: raw[0]}
    m, free = mapfile.create_tsdf_map(chunks, sigma=0.1)
    order = [tags.index(t) for t in mapfile.chunk_order(tags)]
    oc, ocells, ovals, ofree = ob.chunks_to_cells(pos[order], raw[order], sigma=0.1)
    c2, grid, data = ob.set_data((oc.min_x, oc.min_y, oc.min_z), (oc.max_x, oc.max_y, oc.max_z), oc.resolution,
                                 oc.init_value, ocells, ovals)
    assert m.dataSize() == int(c2.data_size) > 0
    assert np.array_equal(m.rawGridOcc(), grid)
    assert np.array_equal(m.rawData().view(np.uint32), data.view(np.uint32))
    assert np.array_equal(free.view(np.uint32), ofree.view(np.uint32))


def test_capacity_and_null_arguments():
    from tsdf_localization_amd import _lib
    from tsdf_localization_amd._lib import MapCoef, lib, ptr
    pos, raw = random_chunks(3, 1)
    c = MapCoef()
    nc, nf = ctypes.c_size_t(0), ctypes.c_size_t(0)
    rc = lib.tsl_map_chunks_to_cells(ptr(pos), ptr(raw), 1, 0.1, ctypes.byref(c), None, None, 0, ctypes.byref(nc),
                                     None, 0, ctypes.byref(nf))
    assert rc == _lib.TSL_ERR_CAPACITY
    _, ocells, _, ofree = ob.chunks_to_cells(pos, raw)
    assert (nc.value, nf.value) == (ocells.shape[0], ofree.shape[0])
    rc = lib.tsl_map_chunks_to_cells(None, None, 1, 0.1, ctypes.byref(c), None, None, 0, ctypes.byref(nc),
                                     None, 0, ctypes.byref(nf))
    assert rc == _lib.TSL_ERR_INVALID_ARGUMENT
    # no chunks: the box is the origin alone, no cells
    rc = lib.tsl_map_chunks_to_cells(None, None, 0, 0.1, ctypes.byref(c), None, None, 0, ctypes.byref(nc),
                                     None, 0, ctypes.byref(nf))
    assert rc == _lib.TSL_OK and (nc.value, nf.value) == (0, 0)


def test_chunk_names():
    from tsdf_localization_amd import mapfile
    assert mapfile.parse_chunk_tag("-3_5_0") == (-3, 5, 0)
    assert mapfile.parse_chunk_tag(" 7_+2_-1") == (7, 2, -1)
    assert mapfile.parse_chunk_tag("3abc_1_2") == (3, 1, 2)  # std::stoi stops at the first non-digit
    for bad in ("a_1_2", "1_2", "1_2_3_4", "_1_2", "99999999999_0_0"):
        with pytest.raises(ValueError):
            mapfile.parse_chunk_tag(bad)
    assert mapfile.chunk_order(["10_0_0", "-1_0_0", "2_0_0", "1_0_0"]) == ["-1_0_0", "10_0_0", "1_0_0", "2_0_0"]


def test_chunk_size_checked():
    from tsdf_localization_amd import mapfile
    with pytest.raises(ValueError):
        mapfile.create_tsdf_map({"0_0_0": np.zeros(100, np.uint32)})


def test_h5_reader_absent_is_reported():
    from tsdf_localization_amd import mapfile
    try:
        import h5py  # noqa: F401
        pytest.skip("h5py is installed")
    except ImportError:
        pass
    with pytest.raises(ImportError, match="h5py"):
        mapfile.load_h5_map("/nonexistent.h5")
