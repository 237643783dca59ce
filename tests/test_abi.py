"""The C ABI: every function include/prism_hip.h declares is exported by libprism_hip.so (the
product) and, under the pref_ prefix, by the oracle library (the checker), so one Python
scenario binds to either.  Loading needs no GPU; nothing here calls into the device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "prism_hip.h")
HIP_LIB = os.path.join(ROOT, "prism_amd", "libprism_hip.so")

# entry points that only a device context has (timing, device info): the oracle has no device
DEVICE_ONLY = {"ctx_set_timing", "ctx_reset_timing", "ctx_kernel_stats", "ctx_info", "eqn_step_noalloc"}
# mesh generation and UNV reading produce the input descriptor both libraries consume
MESH_INPUT = {"meshgen_box", "meshgen_unv", "meshgen_desc", "meshgen_patch_name", "meshgen_destroy"}


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:int|void|prism_\w+\s*\*|const\s+char\s*\*)\s*(prism_\w+)\s*\(", text, re.M)
    assert len(names) > 60
    return sorted(set(names))


def test_header_declares_the_path():
    names = declared()
    for must in ("prism_mesh_create", "prism_eqn_update_coeffs", "prism_solve", "prism_eqn_step",
                 "prism_csr_bicgstab", "prism_simple_step", "prism_piso_step", "prism_prime_step",
                 "prism_adjoint_gradient"):
        assert must in names


def test_hip_library_exports_every_symbol():
    if not os.path.exists(HIP_LIB):
        pytest.fail("libprism_hip.so is not built")
    lib = ctypes.CDLL(HIP_LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert missing == []


def test_oracle_exports_the_same_abi():
    import oracle
    oracle.build()
    lib = ctypes.CDLL(oracle.LIB)
    missing = [n for n in declared() if n[len("prism_"):] not in DEVICE_ONLY | MESH_INPUT
               and not hasattr(lib, "pref_" + n[6:])]
    assert missing == []


def test_python_prototypes_cover_the_header():
    from prism_amd import _capi
    missing = [n for n in declared() if n[len("prism_"):] not in _capi.PROTOS]
    assert missing == []
