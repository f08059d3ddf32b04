"""Builds libpannde_rsg.so in-tree for gfx950 (hipcc cross-compiles without a GPU)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpannde_rsg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off keeps every multiply and add separate, as in the reference's x86-64
# build; packed FP32 VALU ops are disabled (they are formed automatically otherwise).
DEVICE_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-x", "c++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
              "-I/opt/rocm/include", "-Wall", "-Wno-unused-function"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    if verbose and out.strip():
        print(out)


def build(verbose=False):
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    # the host runtime and the four parts of the kernel library (FP64 / FP32, single-instance /
    # batched; see RSG_PART in rsg_kernels.hip) compile in parallel
    jobs = [([HIPCC] + HOST_FLAGS + ["-c", os.path.join(CSRC, "rsg_host.cpp"), "-o",
                                     os.path.join(bdir, "rsg_host.cpp.o")])]
    for part in (1, 2, 3, 4):
        jobs.append([HIPCC] + DEVICE_FLAGS + ["-DRSG_PART=%d" % part, "-c", os.path.join(CSRC, "rsg_kernels.hip"),
                                              "-o", os.path.join(bdir, "rsg_kernels_part%d.o" % part)])
    # the energy reduction
    jobs.append([HIPCC] + DEVICE_FLAGS + ["-c", os.path.join(CSRC, "rsg_energy.hip"), "-o",
                                          os.path.join(bdir, "rsg_energy.o")])
    objs = [cmd[-1] for cmd in jobs]
    procs = []
    for cmd in jobs:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for cmd, pr in procs:
        out = "\n".join(l for l in pr.communicate()[0].splitlines() if "packed-fp32-ops" not in l)
        if pr.returncode != 0:
            failed.append("build failed: %s\n%s" % (" ".join(cmd), out))
        elif verbose and out.strip():
            print(out)
    if failed:
        raise RuntimeError("\n".join(failed))
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950"] + objs +
         ["-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib", "-o", tmp], verbose)
    os.replace(tmp, OUT)
    # the native mesh builder of the case generator (host code only)
    _run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-shared", "-fPIC",
          os.path.join(CSRC, "case_gen.cpp"), "-o", os.path.join(HERE, "libpannde_cases.so")], verbose)
    # the C++ drop-in driver (PanNDE::Model over the C ABI; VTK XML jobs through vtk_job.hpp,
    # zlib for compressed arrays), linked against the library
    inc = os.path.join(ROOT, "include")
    _run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-I" + inc, "-I" + os.path.join(inc, "PanNDE"),
          os.path.join(CSRC, "rsg_driver.cpp"), "-o", os.path.join(HERE, "rsg_driver"), "-L" + HERE,
          "-lpannde_rsg", "-lz", "-Wl,-rpath,$ORIGIN"], verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
