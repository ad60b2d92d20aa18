"""Build the gfx950 library libcc_gpu.so in-tree with hipcc.

    python parallel-critical-collapse_amd/build.py [-v] [-j JOBS]

The library holds the HIP kernels and the C-ABI of include/cc_gpu.h; the
executable cc_newton is the C++ NewtonSolver driver over it, cc_layers the
driver of the SpectralTransformer / ODEStepper / ShootingSolver classes
(host/).  The
length-N kernels are split over one translation unit per supported Ntau
(csrc/cc_ops_<N>.hip) so that they compile in parallel; objects go to
_build/ next to this file, the library next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libcc_gpu.so")
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = [os.path.join(CSRC, "cc_capi.hip")] + \
          [os.path.join(CSRC, f"cc_ops_{n}.hip") for n in (2048, 1024, 512, 256, 128)]
HEADERS = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith(".hpp")] + \
          [os.path.join(ROOT, "include", "cc_gpu.h")]

HOST_SOURCES = [os.path.join(HERE, "host", f) for f in ("newton.cpp", "suite.cpp", "cc_newton.cpp")]
CLI = os.path.join(HERE, "cc_newton")
LAYER_SOURCES = [os.path.join(HERE, "host", f) for f in ("newton.cpp", "layers.cpp", "cc_layers.cpp")]
LAYER_CLI = os.path.join(HERE, "cc_layers")
HOST_HEADERS = [os.path.join(HERE, "host", f) for f in ("newton.hpp", "suite.hpp", "layers.hpp", "json.hpp")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 1)))
    objs = [os.path.join(OBJDIR, os.path.basename(s).replace(".hip", ".o")) for s in SOURCES]
    cmds = [[HIPCC, *FLAGS, "-c", "-o", o, s] for s, o in zip(SOURCES, objs)]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd, out in zip(cmds, ex.map(_run, cmds)):
            if verbose:
                print(" ".join(cmd), flush=True)
                if out.strip():
                    print(out)
    tmp = LIB + ".tmp"
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    out = _run(link)
    if verbose:
        print(" ".join(link))
        if out.strip():
            print(out)
    os.replace(tmp, LIB)
    # host-side drivers (plain C++ over the C-ABI): NewtonSolver and the lower layers
    for srcs, exe in ((HOST_SOURCES, CLI), (LAYER_SOURCES, LAYER_CLI)):
        host = [CXX, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), *srcs,
                "-L", HERE, "-lcc_gpu", "-L/opt/rocm/lib", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib",
                "-o", exe + ".tmp"]
        out = _run(host)
        if verbose:
            print(" ".join(host))
            if out.strip():
                print(out)
        os.replace(exe + ".tmp", exe)
    return LIB


def stale() -> bool:
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    inputs = SOURCES + HEADERS + HOST_SOURCES + LAYER_SOURCES + HOST_HEADERS
    return any(os.path.getmtime(f) > t for f in inputs) or not all(os.path.exists(e) for e in (CLI, LAYER_CLI))


if __name__ == "__main__":
    j = None
    if "-j" in sys.argv:
        j = int(sys.argv[sys.argv.index("-j") + 1])
    print(build(verbose="-v" in sys.argv, jobs=j))
