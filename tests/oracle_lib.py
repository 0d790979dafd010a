"""Python access to the CPU oracle (oracle/_build/liboracle*.so) and to the
reference's Lambert W source compiled into oracle/_ref.  Test infrastructure
only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from mcmatlab_amd import abi

ROOT = abi.repo_root()
ORACLE_DIR = os.path.join(ROOT, "oracle")
PORT_LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
PORT_OMP_LIB = os.path.join(ORACLE_DIR, "_build", "liboracle_omp.so")
PORT_FAST_LIB = os.path.join(ORACLE_DIR, "_build", "liboracle_fast.so")
REF_LAMBERT_LIB = os.path.join(ORACLE_DIR, "_ref", "libref_lambert.so")

RNG_DSFMT = 0
RNG_PHILOX = 1
HEAT_FP32_DGADI = 0
HEAT_FP64_DGADI = 1
HEAT_FP64_EXPLICIT = 2


class orc_mc_stats(C.Structure):
    _fields_ = [("launched", C.c_double), ("substeps", C.c_double), ("voxelChanges", C.c_double),
                ("riCrossings", C.c_double), ("scatterings", C.c_double)]


_libs = {}


def load_oracle(omp=False):
    path = PORT_OMP_LIB if omp else PORT_LIB
    if path not in _libs:
        lib = C.CDLL(path)
        lib.oracle_mc_run.argtypes = [C.POINTER(abi.mcm_mc_input), C.POINTER(abi.mcm_mc_output), C.c_int, C.c_int,
                                      C.c_uint64, C.POINTER(orc_mc_stats)]
        lib.oracle_heat_run.argtypes = [C.POINTER(abi.mcm_heat_input), C.POINTER(abi.mcm_heat_output), C.c_int, C.c_int]
        lib.oracle_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        lib.oracle_dsfmt_stream.argtypes = [C.c_uint32, C.c_int, C.POINTER(C.c_double)]
        lib.oracle_last_error.restype = C.c_char_p
        lib.oracle_downsample_params.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p]
        _libs[path] = lib
    return _libs[path]


def host_threads(cap=16):
    """Threads for OpenMP oracle runs: the cores this process may use, at
    most `cap`."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(cap, n))


_libc = C.CDLL(None)


def oracle_mc(model, simType=1, seed=0, rng=RNG_DSFMT, nthreads=1, stats=False, omp=False):
    lib = load_oracle(omp or nthreads > 1)
    inp = abi.MCInputs(model, simType=simType, seed=seed)
    out = abi.MCOutputs(inp)
    st = orc_mc_stats()
    rc = lib.oracle_mc_run(C.byref(inp.struct), C.byref(out.struct), nthreads, rng, seed, C.byref(st))
    if rc != 0:
        raise RuntimeError("oracle_mc_run failed (%d): %s" % (rc, lib.oracle_last_error().decode()))
    res = out.result(free_fn=lambda p: _libc.free(C.cast(p, C.c_void_p)))
    if stats:
        res["stats"] = {k: getattr(st, k) for k, _ in orc_mc_stats._fields_}
    return res


def oracle_heat(T, Omega, P, variant=HEAT_FP64_DGADI, want_f64=True, T_f64=None, Omega_f64=None, nthreads=1):
    """nthreads > 1 runs the OpenMP build; every value is independent of the
    thread count (lines and voxels are independent, maxima are exact)."""
    lib = load_oracle(nthreads > 1)
    inp = abi.HeatInputs(T, Omega, P, T_f64=T_f64, Omega_f64=Omega_f64)
    out = abi.HeatOutputs(inp, want_f64=want_f64)
    rc = lib.oracle_heat_run(C.byref(inp.struct), C.byref(out.struct), variant, int(nthreads))
    if rc != 0:
        raise RuntimeError("oracle_heat_run failed (%d)" % rc)
    return out.result()


def oracle_binning(paramVals, n2):
    """[downsampledParamVals, binidx] = getDownsampledParamVals(paramVals, n2)
    through the C restatement; paramVals is 3 x N1 single."""
    pv = np.asfortranarray(np.asarray(paramVals, dtype=np.float32).reshape(3, -1))
    n1 = pv.shape[1]
    flat = np.ascontiguousarray(pv.ravel(order="F"))
    out = np.zeros(3 * n2, np.float32)
    bins = np.zeros(n1, np.uint8)
    rc = load_oracle().oracle_downsample_params(flat.ctypes.data, n1, n2, out.ctypes.data, bins.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle_downsample_params failed (%d)" % rc)
    return out.reshape(3, n2, order="F"), bins


def philox(ctr, key):
    lib = load_oracle()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.oracle_philox4x32_10(c, k, o)
    return list(o)


def dsfmt_stream(seed, n):
    lib = load_oracle()
    buf = (C.c_double * n)()
    lib.oracle_dsfmt_stream(seed, n, buf)
    return np.array(buf[:])
