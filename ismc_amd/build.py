"""Build libismc.so (the product: csrc/*.hip) and libismc_diag.so (the
measurement probes: csrc/diag/*.hip) in-tree with hipcc for gfx950.

    python -m ismc_amd.build
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libismc.so")
DIAG_OUT = os.path.join(HERE, "libismc_diag.so")
ARCH = os.environ.get("ISMC_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "--offload-arch=" + ARCH, "-std=c++17", "-fPIC", "-shared",
    # no packed-FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def sources(directory=CSRC):
    return sorted(glob.glob(os.path.join(directory, "*.hip")) + glob.glob(os.path.join(directory, "*.cpp")))


def build(verbose=False):
    out = _link(sources(), OUT, verbose)
    _link(sources(os.path.join(CSRC, "diag")), DIAG_OUT, verbose)
    return out


def _link(srcs, out, verbose):
    hipcc = os.environ.get("HIPCC", "hipcc")
    tmp = out + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-I", os.path.join(HERE, "..", "include"), "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    noise = "is not a recognized feature for this target"
    msgs = "\n".join(l for l in p.stdout.splitlines() if noise not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + msgs)
    if msgs and verbose:
        print(msgs)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
