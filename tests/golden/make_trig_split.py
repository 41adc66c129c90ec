"""Writes glibc_sinf_fma_split.json: the float inputs on which glibc's two
x86-64 builds of sinf/cosf (with and without FMA contraction) differ, with
the outputs of each build.

The inputs are found with the library's host build of both variants
(tsl_trig, host side) over every non-NaN float; the outputs are then taken
from glibc itself, in this process (glibc selects its FMA build on a CPU
with FMA3) and in a child process started with
GLIBC_TUNABLES=glibc.cpu.hwcaps=-FMA,-AVX2 (glibc then selects the build
without FMA).  Run from the repository root after building libtsl.so:
    python tests/golden/make_trig_split.py
"""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "glibc_sinf_fma_split.json")
TUNABLES = "glibc.cpu.hwcaps=-FMA,-AVX2"


def glibc_outputs(inputs):
    libm = ctypes.CDLL("libm.so.6")
    libm.sinf.restype = ctypes.c_float
    libm.sinf.argtypes = [ctypes.c_float]
    libm.cosf.restype = ctypes.c_float
    libm.cosf.argtypes = [ctypes.c_float]
    out = []
    for u in inputs:
        x = np.array([u], np.uint32).view(np.float32)[0]
        s = np.array([libm.sinf(float(x))], np.float32).view(np.uint32)[0]
        c = np.array([libm.cosf(float(x))], np.float32).view(np.uint32)[0]
        out.append([f"{int(s):08x}", f"{int(c):08x}"])
    return out


def find_split():
    sys.path.insert(0, ROOT)
    import tsdf_localization_amd as tsl
    found = []
    step = 1 << 24
    for start in range(0, 1 << 32, step):
        u = np.arange(start, start + step, dtype=np.uint64).astype(np.uint32)
        x = u.view(np.float32)
        s1, c1 = tsl.trig(x, tsl.TRIG_GLIBC_FMA, on_device=False)
        s0, c0 = tsl.trig(x, tsl.TRIG_GLIBC_NOFMA, on_device=False)
        d = (s1.view(np.uint32) != s0.view(np.uint32)) | (c1.view(np.uint32) != c0.view(np.uint32))
        d &= ~np.isnan(x)
        found.extend(int(v) for v in u[d])
    return found


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--glibc":
        inputs = [int(v, 16) for v in sys.argv[2:]]
        print(json.dumps(glibc_outputs(inputs)))
        return
    inputs = find_split()
    fma = glibc_outputs(inputs)
    env = dict(os.environ, GLIBC_TUNABLES=TUNABLES)
    r = subprocess.run([sys.executable, __file__, "--glibc", *[f"{u:08x}" for u in inputs]], env=env,
                       check=True, capture_output=True, text=True)
    nofma = json.loads(r.stdout)
    rows = [{"x": f"{u:08x}", "fma": a, "nofma": b} for u, a, b in zip(inputs, fma, nofma)]
    with open(OUT, "w") as f:
        json.dump({"tunables_without_fma": TUNABLES, "inputs": rows}, f, indent=1)
    print(f"{len(rows)} inputs written to {OUT}")


if __name__ == "__main__":
    main()
