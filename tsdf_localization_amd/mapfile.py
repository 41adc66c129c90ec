"""Map files: the TSDF chunks of a map file to a CudaSubVoxelMap.

createTSDFMap (include/tsdf_localization/map/map_util.h:17-154) reads the
/map group of an HDF5 file: one dataset per chunk, named "x_y_z" by its chunk
coordinates, holding CHUNK_SIZE^3 = 64^3 raw TSDFValue words.  It builds the
map's bounding box from the chunk names, turns every voxel with weight != 0
and |tsdf| < 600 mm into a likelihood cell, collects the voxels with weight !=
0 and |tsdf| >= 600 mm as free-space points, and places the cells with
CudaSubVoxelMap::setData.

The arithmetic runs in libtsl.so (tsl_map_chunks_to_cells, then
tsl_map_set_data).  Reading the HDF5 file is left to h5py where it is
installed (load_h5_map); any other reader can pass the chunks to
create_tsdf_map as a mapping from dataset name to array.
"""
import ctypes
import re

import numpy as np

from . import _lib
from ._lib import MapCoef, check, lib, ptr
from .evaluator import CudaSubVoxelMap

CHUNK_SIZE = 64  # grid_map.h:18-19
CHUNK_VOXELS = CHUNK_SIZE ** 3

_STOI = re.compile(r"\s*([+-]?\d+)")


def parse_chunk_tag(tag):
    """Chunk coordinates from a dataset name, as map_util.h:27-44 parses it:
    split at "_", each part through std::stoi (leading blanks and a sign
    accepted, trailing characters ignored).  Raises ValueError where stoi
    throws and for a name without exactly three parts (the reference indexes
    its three-element bounds with every part)."""
    parts = tag.split("_")
    if len(parts) != 3:
        raise ValueError(f"chunk name {tag!r} does not hold three coordinates")
    out = []
    for p in parts:
        m = _STOI.match(p)
        if m is None:
            raise ValueError(f"chunk name {tag!r}: {p!r} is not an integer")
        v = int(m.group(1))
        if not -2 ** 31 <= v < 2 ** 31:
            raise ValueError(f"chunk name {tag!r}: {p!r} is out of int range")
        out.append(v)
    return tuple(out)


def chunk_order(tags):
    """The order HighFive's listObjectNames gives the datasets of a group: the
    HDF5 name index, ascending by byte string."""
    return sorted(tags, key=lambda t: t.encode())


def chunks_to_cells(positions, raw, sigma=0.1):
    """tsl_map_chunks_to_cells: (coef, cells_xyz (C,3), values (C,), free (F,3)).

    positions: int32 (K, 3) chunk coordinates; raw: uint32 (K, 64^3) TSDFValue
    words, x-major, in the order the reference would list the chunks."""
    pos = np.ascontiguousarray(positions, np.int32).reshape(-1, 3)
    words = np.ascontiguousarray(raw).reshape(pos.shape[0], CHUNK_VOXELS).view(np.uint32)
    c = MapCoef()
    nc, nf = ctypes.c_size_t(0), ctypes.c_size_t(0)
    rc = lib.tsl_map_chunks_to_cells(ptr(pos), ptr(words), pos.shape[0], float(sigma), ctypes.byref(c),
                                     None, None, 0, ctypes.byref(nc), None, 0, ctypes.byref(nf))
    if rc not in (_lib.TSL_OK, _lib.TSL_ERR_CAPACITY):
        check(rc)
    cells = np.empty((nc.value, 3), np.float32)
    values = np.empty(nc.value, np.float32)
    free = np.empty((nf.value, 3), np.float32)
    check(lib.tsl_map_chunks_to_cells(ptr(pos), ptr(words), pos.shape[0], float(sigma), ctypes.byref(c),
                                      ptr(cells), ptr(values), nc.value, ctypes.byref(nc),
                                      ptr(free), nf.value, ctypes.byref(nf)))
    return c, cells, values, free


def create_tsdf_map(chunks, sigma=0.1):
    """createTSDFMap<CudaSubVoxelMap<float, float>, float, float>
    (map_util.h:17-154) over chunks given as a mapping from dataset name
    ("x_y_z") to its 64^3 raw words.  Returns (map, free_map) with free_map a
    float32 (F, 3) array of free-space points (the reference's
    std::vector<CudaPoint>&)."""
    tags = chunk_order(list(chunks.keys()))
    pos = np.array([parse_chunk_tag(t) for t in tags], np.int32).reshape(-1, 3)
    raw = np.empty((len(tags), CHUNK_VOXELS), np.uint32)
    for i, t in enumerate(tags):
        a = np.asarray(chunks[t]).reshape(-1)
        if a.size != CHUNK_VOXELS:
            raise ValueError(f"chunk {t!r} holds {a.size} values, not {CHUNK_VOXELS}")
        raw[i] = a.astype(np.int64) & 0xFFFFFFFF if a.dtype.kind == "i" else a.astype(np.uint32)
    c, cells, values, free = chunks_to_cells(pos, raw, sigma)
    m = CudaSubVoxelMap.from_cells((c.min_x, c.min_y, c.min_z), (c.max_x, c.max_y, c.max_z), c.resolution,
                                   c.init_value, cells, values)
    return m, free


def load_h5_map(path, sigma=0.1):
    """createTSDFMap over an HDF5 map file, read with h5py (not part of this
    package; ImportError where it is not installed)."""
    try:
        import h5py
    except ImportError as e:
        raise ImportError("load_h5_map reads HDF5 map files with h5py, which is not installed; read the /map "
                          "group with another reader and pass its datasets to create_tsdf_map") from e
    with h5py.File(path, "r") as f:
        g = f["/map"]
        chunks = {name: np.asarray(g[name][()]) for name in g.keys()}
    return create_tsdf_map(chunks, sigma)
