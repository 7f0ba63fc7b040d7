"""Copies the reference's own input data files for this path into tests/golden/.

Sources (FwiFlow.jl docs/data, used by docs/src/tutorials/fwi.md:35-43):
  Model_Cp_true.bin, Model_Cp_init_1D.bin : 224 x 448 float32, z fastest
  sourceF_4p5_2_high.bin                  : 2000 float32 source wavelet
These are data (model and wavelet inputs), copied byte for byte; the
reference's expected-output file test/test_data.mat is absent upstream.
Run in a container where the reference is mounted:
    python tests/golden/make_fixtures.py /root/reference
"""
import os
import shutil
import sys

FILES = ["Model_Cp_true.bin", "Model_Cp_init_1D.bin", "sourceF_4p5_2_high.bin"]


def main(ref_root):
    here = os.path.dirname(os.path.abspath(__file__))
    for f in FILES:
        shutil.copyfile(os.path.join(ref_root, "docs", "data", f), os.path.join(here, f))
        print("copied", f)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
