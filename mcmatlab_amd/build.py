"""Build libmcm.so (the gfx950 HIP library behind include/mcm_abi.h) in-tree.

    python -m mcmatlab_amd.build            # library only
    python -m mcmatlab_amd.build --all      # library, oracle, test harnesses

hipcc cross-compiles for gfx950 without a GPU present.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libmcm.so")
BUILD = os.path.join(PKG, "_build")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("MCM_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
          "-I", INCLUDE, "-I", CSRC, "-Wall", "-Wno-unused-function", "-Wno-unused-lambda-capture"]

SOURCES = [
    # (source, extra flags)
    ("mcm_common.cpp", []),
    # contraction only within a source expression, so every kernel
    # instantiation (plain, batched) evaluates the same operations and a
    # packet's trajectory does not depend on the launch it runs in
    ("mc_kernel.hip", ["-ffp-contract=on"]),
    # both passes of a packet must retrace the same path bit for bit
    ("mc_grad.hip", ["-ffp-contract=off"]),
    # the heat stencils are compared with the FP64 oracle operation for
    # operation, so they are built without FMA contraction
    ("heat_kernel.hip", ["-ffp-contract=off"]),
    # the binning distances are compared with the reference bit for bit
    ("binning_kernel.hip", ["-ffp-contract=off"]),
]


def _run(cmd):
    t = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    noisy = [ln for ln in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noisy:
        sys.stderr.write("\n".join(noisy) + "\n")
    return time.time() - t


def build_library(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, extra in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        dt = _run([HIPCC, *COMMON, *DEVICE_FLAGS, *extra, *lang, "-c", path, "-o", obj])
        if verbose:
            print(f"  compiled {src} ({dt:.1f}s)")
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print(f"  linked {os.path.relpath(LIB, ROOT)}")
    return LIB


def build_oracle(verbose=True):
    """The CPU oracle (test infrastructure): the restatement always; the
    reference's Lambert W source only where the reference sources exist."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")
    if verbose and r.stdout.strip():
        print(r.stdout.strip())


def build_test_harnesses(verbose=True):
    """Host build of the photon physics (tests/cpu_harness) and this
    project's MEX layer against the mx-API shim (test infrastructure)."""
    out = os.path.join(ROOT, "tests", "_build")
    os.makedirs(out, exist_ok=True)
    h = os.path.join(ROOT, "tests", "cpu_harness", "mc_host_harness.hip")
    if os.path.exists(h):
        _run([HIPCC, "-O2", "-fPIC", "-shared", "-std=c++17", f"--offload-arch={ARCH}", *DEVICE_FLAGS, "-I", INCLUDE, "-I", CSRC,
              "-x", "hip", h, "-o", os.path.join(out, "libmc_host.so")])
        if verbose:
            print("  built tests/_build/libmc_host.so")
    mexdir = os.path.join(PKG, "mex")
    shim = os.path.join(ROOT, "tests", "mexshim")
    for name in ("MCmatlab_mex.c", "finiteElementHeatPropagator_mex.c", "getDownsampledParamVals_mex.c"):
        src = os.path.join(mexdir, name)
        if not os.path.exists(src):
            continue
        so = os.path.join(out, "lib" + name.replace(".c", "_shim.so"))
        _run(["gcc", "-std=gnu11", "-O2", "-fPIC", "-shared", "-Wall", "-I", shim, "-I", INCLUDE, src,
              os.path.join(shim, "mexshim.c"), "-o", so, "-L", PKG, "-lmcm", f"-Wl,-rpath,{PKG}", "-lm"])
        if verbose:
            print(f"  built {os.path.relpath(so, ROOT)}")


def main(argv):
    t = time.time()
    build_library()
    if "--all" in argv:
        build_oracle()
        build_test_harnesses()
    print(f"build done in {time.time() - t:.1f}s")


if __name__ == "__main__":
    main(sys.argv[1:])
