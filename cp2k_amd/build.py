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
           CSRC / "grid_general.hip", CSRC / "grid_noalloc.hip", CSRC / "grid_buffers.hip",
           CSRC / "grid_rs_pw.hip", CSRC / "grid_task_gen.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed: " + " ".join(cmd[-3:]))
    return [ln for ln in (res.stdout + res.stderr).splitlines()
            if "warning" in ln and "packed-fp32-ops" not in ln]


def build(verbose=False, extra=()):
    """Compiles every source to an object (in parallel), then links."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = [objdir / (src.stem + ".o") for src in SOURCES]
    cmds = [[HIPCC, *FLAGS, *extra, "-c", "-o", str(o), str(src)]
            for src, o in zip(SOURCES, objs)]
    workers = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(workers) as ex:
        warn = [w for ws in ex.map(lambda c: _run(c, verbose), cmds) for w in ws]
    tmp = LIB.with_suffix(".so.tmp")
    rocm_lib = str(pathlib.Path(HIPCC).resolve().parent.parent / "lib")
    warn += _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
                  *map(str, objs), "-L" + rocm_lib, "-Wl,-rpath," + rocm_lib, "-lhipfft",
                  "-Wl,-z,defs"],
                 verbose)
    if verbose and warn:
        print("\n".join(warn))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
