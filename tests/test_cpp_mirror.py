"""The C++20 mirror of the reference's API (include/prism_hip.hpp): the reference's Poisson,
1-D advection and SIMPLE pipe tests, the examples' VTU export and a k-epsilon run written against it as the reference's users write them,
run against the FP64 oracle library (CPU) and against the HIP library (GPU)."""
import gzip
import json
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
    # the coarse pipe with k, epsilon and nu fields for the k-epsilon case
    kp = d / "kepsilon_pipe"
    os.makedirs(kp)
    shutil.copy(d / "coarsePipeHex" / "mesh.unv", kp / "mesh.unv")
    shutil.copy(d / "coarsePipeHex" / "U.json", kp / "U.json")
    fields = [{"name": "U", "type": "velocity", "gradScheme": "greenGauss"}, {"name": "k", "type": "scalar"},
              {"name": "epsilon", "type": "scalar"}, {"name": "nu", "type": "scalar"}]
    with open(kp / "fields.json", "w") as f:
        json.dump({"fields": fields}, f)
    for name, inlet in (("k", 1e-4), ("epsilon", 1e-5), ("nu", 1e-3)):
        patches = [{"name": "Front", "type": "empty"}, {"name": "Back", "type": "empty"},
                   {"name": "UpperWall", "type": "zero-gradient"}, {"name": "LowerWall", "type": "zero-gradient"},
                   {"name": "Inlet", "type": "fixed", "value": inlet}, {"name": "Outlet", "type": "zero-gradient"}]
        with open(kp / f"{name}.json", "w") as f:
            json.dump({"patches": patches}, f)
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
    assert r.stdout.count(" ok") == 6


@pytest.mark.gpu
def test_cpp_mirror_on_gpu(cases_dir):
    r = run(HIP_LIB, "prism_", cases_dir)
    assert r.returncode == 0
    assert r.stdout.count(" ok") == 6
