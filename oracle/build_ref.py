"""ORACLE: build the reference-compiled pieces of the oracle into oracle/_ref.

Only the reference's PRAXIS (CUDA/praxis.cpp, plain C++ with no external
dependencies) compiles from its own sources here; the CUDA transport needs
nvcc/cuRAND/thrust, which this image does not have, so it is not built.
Outputs go to oracle/_ref/ (git-ignored); the sources are read in place
from the reference checkout and never copied.

    python -m oracle.build_ref [--ref /root/reference]
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_OUT = os.path.join(HERE, "_ref")
PRAXIS_LIB = os.path.join(REF_OUT, "libpraxis_ref.so")


def build_if_available(ref="/root/reference"):
    src = os.path.join(ref, "CUDA", "praxis.cpp")
    if not os.path.exists(src):
        return None
    os.makedirs(REF_OUT, exist_ok=True)
    tmp = PRAXIS_LIB + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-w", "-o", tmp,
           os.path.join(HERE, "praxis_ref_harness.cpp"), src]
    subprocess.check_call(cmd)
    os.replace(tmp, PRAXIS_LIB)
    return PRAXIS_LIB


if __name__ == "__main__":
    ref = sys.argv[sys.argv.index("--ref") + 1] if "--ref" in sys.argv else "/root/reference"
    print(build_if_available(ref))
