#!/usr/bin/env python3
"""Copies the case fixtures of the reference's tests (tests/cases/*: UNV meshes, fields.json,
per-field boundary files, OpenFOAM reference fields) into tests/golden/cases, gzipping the
meshes.  Data only: no source file is copied.  Run once where the reference is mounted:

    python tests/golden/make_fixtures.py [/root/reference]
"""
import gzip
import os
import shutil
import sys

CASES = ["poisson", "versteeg_advection_1d", "versteeg_trans_1d", "channel1d_coarse", "coarsePipeHex",
         "coarsePipeHexSymmetrical", "ductSIMPLE", "cavity_hex", "boundary"]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src_root = os.path.join(ref, "tests", "cases")
    dst_root = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cases")
    for case in CASES:
        for dirpath, _, files in os.walk(os.path.join(src_root, case)):
            rel = os.path.relpath(dirpath, src_root)
            out = os.path.join(dst_root, rel)
            os.makedirs(out, exist_ok=True)
            for f in sorted(files):
                src = os.path.join(dirpath, f)
                if f.endswith(".unv"):
                    with open(src, "rb") as a, gzip.GzipFile(os.path.join(out, f + ".gz"), "wb", 9, mtime=0) as b:
                        shutil.copyfileobj(a, b)
                elif f.endswith(".json"):
                    shutil.copyfile(src, os.path.join(out, f))
    print("fixtures written to", dst_root)


if __name__ == "__main__":
    main()
