"""Build recipe of libsdegpu.so (hipcc, gfx950), in-tree.

    python -m sde_amd.build          # or sde_amd.build.build()

The library is written next to this file (sde_amd/libsdegpu.so) so that it
travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = [os.path.join(HERE, "csrc", f) for f in ("abi.hip", "grid.hip", "fft.hip", "degrid.hip", "sort.hip")]
OUT = os.path.join(HERE, "libsdegpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # keep the reference's operation order: no fused multiply-add unless written
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    # every symbol resolved at link time
    "-Wl,--no-undefined",
]


def build(verbose=False):
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + SRC
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libsdegpu.so")
    warn = [l for l in (r.stdout + r.stderr).splitlines() if "warning" in l and "packed-fp32-ops" not in l]
    if verbose and warn:
        print("\n".join(warn))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
