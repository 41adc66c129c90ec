"""Build recipe for the native parts of tsdf_localization_amd.

libtsl.so     HIP kernels + C ABI (include/tsl.h), gfx950 only, hipcc.
libsynth.so   seeded synthetic scenes, TSDF maps, ray-cast scans and particle
              sets (host C++/OpenMP); input generators for tests and bench.

Floating-point flags are part of the contract: -ffp-contract=off keeps
a*b+c as two roundings (the reference is compiled for baseline x86-64,
which has no FMA), fp32 division is IEEE, fp32 denormals are kept, and
packed fp32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]

HIP_SOURCES = ["api.hip", "eval.hip", "resample.hip", "tables.hip", "scan.hip", "batch.hip", "motion.hip", "mcl.hip", "server.hip", "chunks.hip",
               "noalloc.hip", "probe.hip"]
# eval_tiles.hip is compiled once per evaluation tile shape (particles per
# wavefront, points per lane), so that the kernel instantiations build in
# parallel
TILE_SHAPES = [(1, 1), (1, 2), (1, 4), (2, 1), (2, 2), (2, 4)]
SYNTH_SOURCES = ["synth.cpp"]

LIBTSL = os.path.join(HERE, "libtsl.so")
LIBSYNTH = os.path.join(HERE, "libsynth.so")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out, flush=True)


def build_tsl(verbose=False, jobs=None):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for k, p in TILE_SHAPES:
        obj = os.path.join(objdir, f"eval_tiles_k{k}_p{p}.o")
        procs.append([HIPCC, *HIP_FLAGS, f"-DTSL_TILE_K={k}", f"-DTSL_TILE_P={p}", "-c",
                      os.path.join(CSRC, "eval_tiles.hip"), "-o", obj])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        objs.append(obj)
        procs.append(cmd)
    if jobs is None:
        jobs = max(1, min(8, os.cpu_count() or 1))
    # compile in parallel
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), procs))
    tmp = LIBTSL + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"], verbose)
    os.replace(tmp, LIBTSL)
    return LIBTSL


def build_synth(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SYNTH_SOURCES]
    tmp = LIBSYNTH + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          "-fvisibility=hidden", "-Wall", "-o", tmp, *srcs], verbose)
    os.replace(tmp, LIBSYNTH)
    return LIBSYNTH


DROPIN_SRC = os.path.join(ROOT, "tests", "cpp", "dropin_check.cpp")
DROPIN_BIN = os.path.join(HERE, "_build", "dropin_check")


def build_dropin_check(verbose=False):
    """The C++ drop-in test program (include/tsdf_localization/hip_evaluator.h
    over libtsl.so), a plain host build: g++ links the HIP library like any
    C++ application of the reference would."""
    os.makedirs(os.path.dirname(DROPIN_BIN), exist_ok=True)
    tmp = DROPIN_BIN + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), "-o", tmp,
          DROPIN_SRC, "-L", HERE, "-ltsl", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/.."], verbose)
    os.replace(tmp, DROPIN_BIN)
    return DROPIN_BIN


def build_all(verbose=False):
    build_synth(verbose)
    build_tsl(verbose)
    build_dropin_check(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIBTSL, LIBSYNTH, DROPIN_BIN)
