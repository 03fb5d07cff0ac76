"""Build the gfx950 decoder library in-tree.

``python -m brkga_fft2_amd.build`` compiles ``csrc/brkga_decoder.hip`` into
``_lib/libbrkga_mi355x.so`` with hipcc (cross-compiles without a GPU).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libbrkga_mi355x.so")
SOURCES = [os.path.join(HERE, "csrc", n) for n in ("brkga_decoder.hip", "brkga_ga.hip", "brkga_surrogate.hip")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BRKGA_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    # bookkeeping in float64 must keep the reference's add order; no contraction
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the machine scheduler's extra stage that reschedules high-register-
    # pressure regions unclustered costs the decode kernel 1-1.5% (measured
    # on C2/C3/C5, identical results)
    "-mllvm", "-amdgpu-disable-unclustered-high-rp-reschedule",
    "-I" + os.path.join(ROOT, "include"),
]


def build(verbose=False):
    """Compile the library; raises CalledProcessError with hipcc's output on failure."""
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + ["-o", tmp] + SOURCES
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise subprocess.CalledProcessError(res.returncode, cmd, output=out)
    if verbose and out.strip():
        print(out)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
