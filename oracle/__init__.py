"""FP64 CPU oracle of the reference's finite-volume path (TEST INFRASTRUCTURE ONLY).

Imported only by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, as
the checker.  `backend()` binds oracle/_build/liboracle.so (same C ABI as the HIP library,
prefix `pref_`) to the prism_amd Python mirror so one scenario runs on both.
"""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liboracle.so")
OMP_LIB = os.path.join(_HERE, "_build", "liboracle_omp.so")


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def backend():
    from prism_amd._capi import Lib
    from prism_amd.core import Backend
    if not os.path.exists(LIB):
        build()
    return Backend(Lib(LIB, "pref_"))


def omp_bicgstab(rp, col, val, b, maxit, threads):
    """Unpreconditioned FP64 BiCGSTAB from x = 0 over a CSR system on `threads` OpenMP
    threads (oracle/omp_bicgstab.cpp), exactly `maxit` iterations.  Returns
    (x, iterations, seconds) with the seconds timed around the iteration loop."""
    import ctypes

    import numpy as np

    if not os.path.exists(OMP_LIB):
        build()
    lib = ctypes.CDLL(OMP_LIB)
    f = lib.oref_omp_bicgstab
    P = ctypes.c_void_p
    f.argtypes = [ctypes.c_int64, P, P, P, P, P, ctypes.c_int64, ctypes.c_int32, P, P]
    f.restype = ctypes.c_int
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    n = len(b)
    x = np.zeros(n)
    it = ctypes.c_int64(0)
    sec = ctypes.c_double(0.0)
    rc = f(n, rp.ctypes.data, col.ctypes.data, val.ctypes.data, b.ctypes.data, x.ctypes.data, int(maxit),
           int(threads), ctypes.addressof(it), ctypes.addressof(sec))
    if rc != 0:
        raise ValueError("omp_bicgstab: empty system or no threads")
    return x, it.value, sec.value
