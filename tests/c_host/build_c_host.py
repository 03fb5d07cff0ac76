"""Compile the C host program ``decode_c.c`` against the built decoder library
(gcc, C99).  Called by the repository's build step after the library."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB_DIR = os.path.join(ROOT, "brkga_fft2_amd", "_lib")
EXE = os.path.join(HERE, "decode_c")
ROCM_LIB = os.environ.get("ROCM_PATH", "/opt/rocm") + "/lib"


def build():
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror",
           "-I" + os.path.join(ROOT, "include"), os.path.join(HERE, "decode_c.c"),
           "-L" + LIB_DIR, "-lbrkga_mi355x",
           "-Wl,-rpath,$ORIGIN/../../brkga_fft2_amd/_lib", "-Wl,-rpath-link," + ROCM_LIB,
           "-o", EXE]
    subprocess.run(cmd, check=True)
    return EXE


if __name__ == "__main__":
    print(build())
