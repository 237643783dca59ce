"""The C++20 mirror of the reference's API (include/prism_hip.hpp): the reference's Poisson,
1-D advection and SIMPLE pipe tests and the examples' VTU export written against it as the reference's users write them,
run against the FP64 oracle library (CPU) and against the HIP library (GPU)."""
import gzip
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "cases")
HIP_LIB = os.path.join(ROOT, "prism_amd", "libprism_hip.so")
CASES = ("poisson", "versteeg_advection_1d", "coarsePipeHex")


def binary():
    import __graft_entry__ as g
    if not os.path.exists(g.MIRROR_BIN):
        g.build_cpp_mirror_test()
    return g.MIRROR_BIN


@pytest.fixture(scope="module")
def cases_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("cases")
    for c in CASES:
        os.makedirs(d / c)
        for f in os.listdir(os.path.join(GOLDEN, c)):
            src = os.path.join(GOLDEN, c, f)
            if f.endswith(".unv.gz"):
                with gzip.open(src, "rb") as a, open(d / c / f[:-3], "wb") as b:
                    shutil.copyfileobj(a, b)
            elif f.endswith(".json"):
                shutil.copy(src, d / c / f)
    return str(d)


def run(lib, prefix, cases_dir):
    r = subprocess.run([binary(), lib, prefix, HIP_LIB, cases_dir], capture_output=True, text=True, timeout=600)
    print(r.stdout, r.stderr)
    return r


def test_cpp_mirror_on_oracle(cases_dir):
    import oracle
    oracle.build()
    r = run(oracle.LIB, "pref_", cases_dir)
    assert r.returncode == 0
    assert r.stdout.count(" ok") == 5


@pytest.mark.gpu
def test_cpp_mirror_on_gpu(cases_dir):
    r = run(HIP_LIB, "prism_", cases_dir)
    assert r.returncode == 0
    assert r.stdout.count(" ok") == 5
