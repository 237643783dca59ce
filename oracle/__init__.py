"""FP64 CPU oracle of the reference's finite-volume path (TEST INFRASTRUCTURE ONLY).

Imported only by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, as
the checker.  `backend()` binds oracle/_build/liboracle.so (same C ABI as the HIP library,
prefix `pref_`) to the prism_amd Python mirror so one scenario runs on both.
"""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liboracle.so")


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def backend():
    from prism_amd._capi import Lib
    from prism_amd.core import Backend
    if not os.path.exists(LIB):
        build()
    return Backend(Lib(LIB, "pref_"))
