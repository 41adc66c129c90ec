"""bench.py's host-side contract, without a GPU: the one-GPU refusal, and
that a profile summary's figures are read only for the tile they were taken
at (the roofline's traffic and the VALU view must describe the kernel the
line times)."""
import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_profile_values_only_for_their_tile(bench):
    pinned = dict(bench.PINNED_TILE)
    traffic = bench._profile_value(bench.TRAFFIC_FILE, "traffic_bytes_per_launch", pinned)
    assert traffic is not None and traffic > 1e9
    other = dict(pinned, k=3 - pinned["k"])
    assert bench._profile_value(bench.TRAFFIC_FILE, "traffic_bytes_per_launch", other) is None
    other = dict(pinned, order=0)
    assert bench._profile_value(bench.TRAFFIC_FILE, "traffic_bytes_per_launch", other) is None
    assert bench._profile_value(os.path.join(ROOT, "profiles", "absent.md"), "x", pinned) is None


def test_every_tile_line_is_a_tile_key(bench):
    """Every committed summary that names a tile names it in the form the
    bench compares against, and the keyed figures parse as numbers."""
    prof = os.path.join(ROOT, "profiles")
    keys = {bench.tile_key(dict(bench.PINNED_TILE)), bench.tile_key(dict(bench.CAMPUS_TILE))}
    seen = 0
    for name in sorted(os.listdir(prof)):
        with open(os.path.join(prof, name)) as fh:
            lines = fh.read().splitlines()
        tiles = [l.split(":", 1)[1].strip() for l in lines if l.startswith("tile:")]
        assert len(tiles) <= 1, name
        if not tiles:
            continue
        seen += 1
        assert tiles[0] in keys, (name, tiles[0])
        for l in lines:
            if l.endswith("_per_launch:") or (l.split(":", 1)[0].endswith("_per_launch") and ":" in l):
                float(l.split(":", 1)[1])
    assert seen >= 5
