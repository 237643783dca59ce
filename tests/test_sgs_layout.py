"""Host-side matrix layouts checked on the CPU by programs under tests/cpp (built here with g++
against prism_amd/csrc/host_mesh.cpp): the multicolour SGS preconditioner's layout
(buildSgsLayout) with the sweep arithmetic of its kernels, and the sliced-ELL column
encodings (buildSell, sellRelative)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_sgs_layout(tmp_path):
    exe = str(tmp_path / "sgs_layout_check")
    cmd = ["g++", "-O2", "-std=c++20", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           os.path.join(ROOT, "tests", "cpp", "sgs_layout_check.cpp"),
           os.path.join(ROOT, "prism_amd", "csrc", "host_mesh.cpp"), "-o", exe, "-lz"]
    subprocess.run(cmd, check=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert r.stdout.strip().endswith("ok")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_sell_offsets(tmp_path):
    """The sliced-ELL 16-bit column encoding (per-slot base + 16-bit offset) reproduces every
    32-bit column, also where columns lie more than 2^15 rows away (tetrahedral box meshes),
    and refuses slots whose offsets span more than 2^16 (tests/cpp/sell_offsets_check.cpp)."""
    exe = str(tmp_path / "sell_offsets_check")
    cmd = ["g++", "-O2", "-std=c++20", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           os.path.join(ROOT, "tests", "cpp", "sell_offsets_check.cpp"),
           os.path.join(ROOT, "prism_amd", "csrc", "host_mesh.cpp"), "-o", exe, "-lz"]
    subprocess.run(cmd, check=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert r.stdout.strip().endswith("ok")
