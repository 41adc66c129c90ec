"""The .mcl snapshot format (MCLFile, src/util/mcl_file.cpp:14-113): the
writer's text equals the reference's stream formatting (default precision,
6 significant digits, %g), the reader parses it back, and a hand-written
file reads as expected.  No GPU."""
import numpy as np
import pytest

from tsdf_localization_amd import _lib, mcl


def _g(v):
    return "%g" % float(np.float32(v))


def test_writer_text_matches_stream_formatting(tmp_path):
    rng = np.random.default_rng(1)
    pts = rng.normal(0, 10, (5, 3)).astype(np.float32)
    pts[0] = [1e-5, 123456789.0, -0.0]
    rings = np.array([0, 15, 3, 63, 7], np.int32)
    parts = rng.normal(0, 2, (3, 7)).astype(np.float32)
    tf = np.eye(4, dtype=np.float32).reshape(-1)
    tf[3] = 0.1
    pose = np.array([1.5, -2.25, 0.0, 0.0, 0.0, 0.38268343, 0.9238795], np.float32)
    path = tmp_path / "a.mcl"
    mcl.write(path, pts, rings, parts, tf, pose)
    text = path.read_text()
    expect = f"{len(pts)}\n\n"
    expect += "".join(" ".join(_g(v) for v in p) + "\n" for p in pts)
    expect += "".join(f"{r}\n" for r in rings)
    expect += f"{len(parts)}\n"
    expect += "".join(" ".join(_g(v) for v in p[:6]) + "  " + _g(p[6]) + "\n" for p in parts)
    expect += "\n" + "".join(_g(v) + " " for v in tf) + "\n"
    expect += " ".join(_g(v) for v in pose)
    assert text == expect


def test_round_trip_to_six_digits(tmp_path):
    rng = np.random.default_rng(2)
    pts = rng.uniform(-30, 30, (200, 3)).astype(np.float32)
    rings = rng.integers(0, 64, 200).astype(np.int32)
    parts = rng.uniform(-3, 3, (50, 7)).astype(np.float32)
    tf = rng.normal(0, 1, 16).astype(np.float32)
    pose = rng.normal(0, 1, 7).astype(np.float32)
    path = tmp_path / "b.mcl"
    mcl.write(path, pts, rings, parts, tf, pose)
    s = mcl.read(path)
    assert np.array_equal(s.rings, rings)
    for a, b in ((s.points, pts), (s.particles, parts), (s.tf, tf), (s.pose, pose)):
        assert np.allclose(a, b, rtol=1e-5, atol=1e-6)
        # exactly what %g text parses to
        assert np.array_equal(a, np.array([float(_g(v)) for v in b.reshape(-1)], np.float32).reshape(b.shape))


def test_hand_written_file(tmp_path):
    path = tmp_path / "c.mcl"
    path.write_text("2\n\n1 2 3\n4.5 -6 7e-1\n5\n9\n1\n0.5 0.25 1 0 0 3.14159  1\n\n"
                    + "1 0 0 0 0 1 0 0 0 0 1 0.4 0 0 0 1 \n" + "1 2 3 0 0 0 1")
    s = mcl.read(path)
    assert s.points.tolist() == [[1, 2, 3], [4.5, -6, np.float32(0.7)]]
    assert s.rings.tolist() == [5, 9]
    assert s.particles.shape == (1, 7) and s.particles[0, 5] == np.float32(3.14159)
    assert s.tf[11] == np.float32(0.4) and s.pose.tolist() == [1, 2, 3, 0, 0, 0, 1]


def test_truncated_file_raises(tmp_path):
    path = tmp_path / "d.mcl"
    path.write_text("3\n\n1 2 3\n")
    with pytest.raises(_lib.TslError, match="Could not read mcl data"):
        mcl.read(path)
