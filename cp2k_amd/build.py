"""Build the MI355X grid library in-tree: cp2k_amd/libcp2k_grid_mi355x.so.

Plain hipcc, gfx950 only.  Flags:
  -ffp-contract=off    task geometry and truncation bounds must round like the
                       reference; hot loops use explicit fma()
  -munsafe-fp-atomics  hardware FP64 atomics (global and LDS)
  -packed-fp32-ops     disabled: no packed FP32 instructions are generated
"""
import os
import pathlib
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libcp2k_grid_mi355x.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = [CSRC / "grid_api.cpp", CSRC / "grid_kernels.hip", CSRC / "grid_wave.hip",
           CSRC / "grid_general.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def build(verbose=False, extra=()):
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, *extra, "-o", str(tmp), *map(str, SOURCES)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building " + LIB.name)
    warn = [ln for ln in (res.stdout + res.stderr).splitlines()
            if "warning" in ln and "packed-fp32-ops" not in ln]
    if verbose and warn:
        print("\n".join(warn))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
