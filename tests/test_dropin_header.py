"""The C++ drop-in header compiles where the reference would include it.

tests/cpp/refshape holds declarations with the shape of the reference's
headers (Evaluator with both evaluate overloads and no virtual destructor,
cuda_evaluator.h:94-107; Resampler without one, resampler.h:16-31; the
MapCoef layout, cuda_sub_voxel_map.h:22-50; the ROS message members read).
tests/cpp/refshape_check.cpp holds the integration snippets (TSDFEvaluator's
construction of its evaluator, both evaluate calls through the base pointer,
mcl_3d's resampler) and is compiled with TSL_WITH_REFERENCE_HEADERS.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


def _compile(src, *extra):
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter", "-fsyntax-only",
           "-I", os.path.join(ROOT, "include"), *extra, os.path.join(CPP, src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_reference_shaped_headers():
    _compile("refshape_check.cpp", "-I", os.path.join(CPP, "refshape"))


def test_standin_headers():
    _compile("dropin_check.cpp")
