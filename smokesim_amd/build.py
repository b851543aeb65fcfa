"""Build recipe of libsmoke_hip.so (gfx950) and the C host driver.

Run ``python -m smokesim_amd.build`` or call ``build()``.  Everything is
compiled in-tree into ``smokesim_amd/_build`` and ``smokesim_amd/libsmoke_hip.so``.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsmoke_hip.so")
DRIVER = os.path.join(BUILD, "smk_driver")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: every a*b+c rounds twice, as the reference expressions
# are written; the kernels are bandwidth bound so FMA contraction buys
# nothing.  Packed fp32 instructions are disabled for device code.
DEVICE_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []  # (source, object, extra flags)
    for src, extra in (("smk_kernels.hip", []), ("smk_pressure.hip", []), ("smk_adjoint.hip", []),
                       ("smk_noalloc.hip", []),
                       # the llvmpipe-compatible arithmetic (namespace smk::gl)
                       ("smk_kernels.hip", ["-DSMK_GL_ARITH=1"]),
                       ("smk_pressure.hip", ["-DSMK_GL_ARITH=1"])):
        obj = os.path.join(BUILD, src.replace(".hip", "_gl.o" if extra else ".o"))
        jobs.append([HIPCC, *DEVICE_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
    robj = os.path.join(BUILD, "smk_runtime.o")
    jobs.append([HIPCC, *HOST_FLAGS, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c",
                 os.path.join(CSRC, "smk_runtime.cpp"), "-o", robj])
    # the translation units compile independently: run them side by side
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as ex:
        list(ex.map(_run, jobs))
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    shutil.move(tmp, LIB)
    # C host driver (the reference's main.c step loop, no GUI)
    drv = os.path.join(CSRC, "smk_driver.c")
    if os.path.exists(drv):
        _run(["gcc", "-O2", "-std=c11", "-Wall", "-I", os.path.join(ROOT, "include"), drv,
              "-o", DRIVER, "-L", HERE, "-lsmoke_hip", f"-Wl,-rpath,{HERE}", "-lm"])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
