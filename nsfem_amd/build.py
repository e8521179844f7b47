"""Builds libnsfem_hip.so in-tree: HIP sources with hipcc for gfx950, host
set-up sources with g++, linked by hipcc.  Run as ``python nsfem_amd/build.py``
(a plain script: importing the package would load the library first)."""

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnsfem_hip.so")

HIP_SOURCES = ["ops.hip", "solver.hip", "amg_pcg.hip", "chorin.hip", "capi.hip", "assembly.hip"]
# The GPU assembly evaluates the element integrals the host assembly shares
# with it; without FP contraction both run the same IEEE operations.
HIP_EXTRA_FLAGS = {"assembly.hip": ["-ffp-contract=off"]}
CPP_SOURCES = ["host/mesh.cpp", "host/csr.cpp", "host/fem.cpp", "host/context.cpp", "host/amg.cpp", "host/fsai.cpp"]
APPS = ["navier_stokes", "solver_test"]
HEADERS = ["common.hpp", "kernels.hpp", "ops.hpp", "solver.hpp", "chorin.hpp", "host/mesh.hpp", "host/csr.hpp",
           "host/fem.hpp", "host/context.hpp", "host/element_math.hpp", "assembly.hpp"]

ARCH = os.environ.get("NSFEM_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}",
    # Packed FP32 instructions stay off: the compiler must not form
    # v_pk_{add,mul,fma}_f32 from scalar or float2 code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-variable", "-Wno-unused-function",
    "-I" + CSRC, "-I" + os.path.join(CSRC, "host"),
]
CXX_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", "-I" + os.path.join(CSRC, "host")]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return res


def build(verbose: bool = False, jobs: int = 0) -> str:
    """Compiles every source (always) and links the shared library."""
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    jobs = jobs or min(8, os.cpu_count() or 4)
    tasks = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
        tasks.append([hipcc, *HIPCC_FLAGS, *HIP_EXTRA_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CPP_SOURCES:
        obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
        tasks.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for fut in [ex.submit(_run, t, verbose) for t in tasks]:
            fut.result()
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lpthread"], verbose)
    os.replace(tmp, LIB)
    # C++ host tools over the C ABI (the reference's host language)
    for app in APPS:
        out = os.path.join(HERE, app)
        _run(["g++", "-O2", "-std=c++20", "-Wall", "-Wextra", "-o", out + ".tmp", os.path.join(CSRC, "app", app + ".cpp"),
              "-L" + HERE, "-lnsfem_hip", "-Wl,-rpath,$ORIGIN"], verbose)
        os.replace(out + ".tmp", out)
    return LIB


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
