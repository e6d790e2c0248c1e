"""C-ABI of libcp2k_grid_mi355x.so without a GPU: the library loads, exports
every entry point declared in include/*.h, and the host-only entry points
behave like the reference's (basis sets, library state, statistics).
Compute entry points need a device and must fail loudly without one."""
import ctypes as C
import os
import pathlib
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def L():
    from cp2k_amd import build, grid
    if not build.LIB.exists():
        build.build()
    return grid.lib()


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b((?:grid|offload)_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_reference_api():
    names = set(declared_functions())
    for fn in ("grid_create_task_list", "grid_free_task_list", "grid_collocate_task_list",
               "grid_integrate_task_list", "grid_create_basis_set", "grid_free_basis_set",
               "grid_library_init", "grid_library_finalize", "grid_library_set_config",
               "grid_library_get_config", "grid_library_print_stats",
               "grid_library_counter_add", "offload_create_buffer", "offload_free_buffer",
               "offload_get_buffer_host_pointer"):
        assert fn in names


def test_every_declared_symbol_is_exported(L):
    from cp2k_amd import build
    out = subprocess.run(["nm", "-D", "--defined-only", str(build.LIB)],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_product_does_not_use_the_oracle():
    pat = re.compile(r"(import\s+oracle|from\s+oracle|liboracle|grid_oracle|_ref/)")
    for p in (ROOT / "cp2k_amd").rglob("*"):
        if p.suffix in (".py", ".cpp", ".hip", ".h"):
            assert not pat.search(p.read_text()), p


def test_basis_set_roundtrip(L):
    from cp2k_amd import grid, synth
    rng = np.random.default_rng(0)
    spec = synth.make_basis(rng, [(0, 2, 3, 2), (3, 3, 1, 1)], 0.1, 10.0)
    b = grid.Basis(spec)
    s = b.h.contents
    assert (s.nset, s.nsgf, s.maxco, s.maxpgf) == (spec.nset, spec.nsgf, spec.maxco, spec.maxpgf)
    sphi = np.ctypeslib.as_array(s.sphi, shape=(spec.nsgf * spec.maxco,))
    assert np.array_equal(sphi, spec.sphi.ravel())
    zet = np.ctypeslib.as_array(s.zet, shape=(spec.nset * spec.maxpgf,))
    assert np.array_equal(zet, spec.zet.ravel())
    assert [s.first_sgf[i] for i in range(spec.nset)] == list(spec.first_sgf)


def test_library_config_and_stats(L):
    from cp2k_amd import grid
    L.grid_library_init()
    try:
        L.grid_library_set_config(C.c_int(14), C.c_bool(False), C.c_bool(True))
        cfg = L.grid_library_get_config()
        assert (cfg.backend, cfg.validate, cfg.apply_cutoff) == (14, False, True)
        lines = []
        CB = C.CFUNCTYPE(None, C.c_char_p, C.c_int, C.c_int)
        cb = CB(lambda msg, n, unit: lines.append(msg[:n].decode()))
        L.grid_library_print_stats(C.c_int(0), cb, C.c_int(6))
        assert lines == []  # nothing counted yet
        L.grid_library_counter_add(C.c_int(4), C.c_int(14), C.c_int(0), C.c_int(7))
        L.grid_library_counter_add(C.c_int(2), C.c_int(14), C.c_int(3), C.c_int(3))
        L.grid_library_print_stats(C.c_int(0), cb, C.c_int(6))
        text = "".join(lines)
        assert "GRID STATISTICS" in text
        assert re.search(r"4\s+collocate ortho\s+GPU\s+7\s+70.00%", text)
        assert re.search(r"2\s+integrate general\s+GPU\s+3\s+30.00%", text)
    finally:
        L.grid_library_set_config(C.c_int(10), C.c_bool(False), C.c_bool(False))
        L.grid_library_finalize()


def _run(code):
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1",
               PYTHONPATH=str(ROOT))
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                          env=env, timeout=300)


def test_create_without_device_fails_loudly(L):
    code = (
        "from cp2k_amd import synth, grid\n"
        "a, pab, V = synth.config('ref4', seed=1)\n"
        "grid.GridTaskList(a)\n"
        "print('NOT REACHED')\n")
    r = _run(code)
    assert r.returncode != 0
    assert "NOT REACHED" not in r.stdout
    assert "grid_mi355x error" in r.stderr


def test_host_backends_are_refused(L):
    code = (
        "import ctypes as C\n"
        "from cp2k_amd import synth, grid\n"
        "grid.lib().grid_library_set_config(C.c_int(11), C.c_bool(False), C.c_bool(False))\n"
        "a, pab, V = synth.config('ref4', seed=1)\n"
        "grid.GridTaskList(a)\n")
    r = _run(code)
    assert r.returncode != 0
    assert "not provided by this library" in r.stderr


def test_empty_task_list_needs_no_device(L):
    """ntasks == 0: create succeeds and collocate zeroes the grids on the host
    (src/grid/grid_task_list.c:65-70, :185-189), without touching a GPU."""
    code = (
        "import ctypes as C, numpy as np\n"
        "from cp2k_amd import grid\n"
        "L = grid.lib()\n"
        "h = C.c_void_p()\n"
        "npts = np.array([[4,4,4]], dtype=np.int32)\n"
        "z = np.zeros(9)\n"
        "L.grid_create_task_list(C.c_bool(True), 0, 1, 0, 0, 0, None, None, None, None,\n"
        "  None, None, None, None, None, None, None, None, None, None, None,\n"
        "  npts.ctypes.data_as(C.c_void_p), npts.ctypes.data_as(C.c_void_p),\n"
        "  np.zeros(3,np.int32).ctypes.data_as(C.c_void_p), np.zeros(3,np.int32).ctypes.data_as(C.c_void_p),\n"
        "  z.ctypes.data_as(C.c_void_p), z.ctypes.data_as(C.c_void_p), C.byref(h))\n"
        "buf = np.full(64, 7.0)\n"
        "ob = grid.OffloadBuffer(64*8, buf.ctypes.data_as(C.POINTER(C.c_double)), None)\n"
        "arr = (C.POINTER(grid.OffloadBuffer)*1)(C.pointer(ob))\n"
        "L.grid_collocate_task_list(h, 100, 1, npts.ctypes.data_as(C.c_void_p), None, arr)\n"
        "assert not buf.any()\n"
        "L.grid_free_task_list(h)\n"
        "print('OK')\n")
    r = _run(code)
    assert r.returncode == 0, r.stderr
    assert "OK" in r.stdout


@pytest.mark.parametrize("compiler", [["gcc", "-std=c11", "-x", "c"],
                                      ["g++", "-std=c++17", "-x", "c++"]])
def test_headers_compile_and_link_as_a_caller(L, tmp_path, compiler):
    """A C (and C++) translation unit that includes every header and takes
    the address of every declared function compiles warning-free and links
    against the library: the declarations are usable as written."""
    from cp2k_amd import build
    names = declared_functions()
    headers = sorted(h.name for h in (ROOT / "include").glob("*.h"))
    src = "\n".join(f'#include "{h}"' for h in headers)
    src += "\n#include <stdio.h>\nint main(void) {\n  void *fns[] = {\n"
    src += ",\n".join(f"    (void *)&{n}" for n in names)
    src += "\n  };\n  printf(\"%d\\n\", (int)(sizeof fns / sizeof fns[0]));\n  return 0;\n}\n"
    f = tmp_path / "caller.src"
    f.write_text(src)
    exe = tmp_path / "caller"
    cmd = compiler + [str(f), "-x", "none", "-Wall", "-Wextra", "-Werror",
                      f"-I{ROOT / 'include'}", "-o", str(exe), str(build.LIB),
                      f"-Wl,-rpath,{build.LIB.parent}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert run.returncode == 0 and int(run.stdout) == len(names), run.stderr
