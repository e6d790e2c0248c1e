"""ctypes binding of the CPU oracle (liboracle.so) and of the reference's
single-task kernels compiled into oracle/_ref (when that build exists).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- never by the product package.
"""
import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libgridref.so"
REF_LIB_O3 = HERE / "_ref" / "libgridref_o3.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(ref=None):
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    targets = ["all"]
    if ref is None:
        ref = pathlib.Path("/root/reference/src/grid").is_dir()
    if ref:
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", str(HERE)] + targets, check=True)


def _ptr(a, ctype=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


class OracleBasis(C.Structure):
    _fields_ = [("nset", C.c_int), ("nsgf", C.c_int), ("maxco", C.c_int),
                ("maxpgf", C.c_int), ("lmin", _ip), ("lmax", _ip),
                ("npgf", _ip), ("nsgf_set", _ip), ("first_sgf", _ip),
                ("sphi", _dp), ("zet", _dp)]


def make_basis_struct(b):
    """b: object with nset, nsgf, maxco, maxpgf, lmin, lmax, npgf, nsgf_set,
    first_sgf (int32 arrays) and sphi [nsgf][maxco], zet [nset][maxpgf]."""
    keep = []

    def iarr(x):
        a = np.ascontiguousarray(x, dtype=np.int32)
        keep.append(a)
        return _ptr(a, C.c_int)

    def darr(x):
        a = np.ascontiguousarray(x, dtype=np.float64)
        keep.append(a)
        return _ptr(a)

    s = OracleBasis(int(b.nset), int(b.nsgf), int(b.maxco), int(b.maxpgf),
                    iarr(b.lmin), iarr(b.lmax), iarr(b.npgf), iarr(b.nsgf_set),
                    iarr(b.first_sgf), darr(b.sphi), darr(b.zet))
    s._keep = keep
    return s


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build(ref=False)
        L = C.CDLL(str(LIB))
        L.oracle_collocate_pgf_product.restype = None
        L.oracle_integrate_pgf_product.restype = None
        L.oracle_integrate_pgf_product_full.restype = None
        L.oracle_task_points.restype = C.c_long
        L.oracle_create_task_list.restype = C.c_void_p
        L.oracle_free_task_list.argtypes = [C.c_void_p]
        L.oracle_collocate_task_list.restype = None
        L.oracle_integrate_task_list.restype = None
        L.oracle_task_list_points.restype = C.c_long
        L.oracle_task_list_points.argtypes = [C.c_void_p, C.c_int]
        _lib = L
    return _lib


def _d(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def _i(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def _single_args(t):
    return [_d(t.dh), _d(t.dh_inv), _d(t.ra), _d(t.rab), _i(t.npts_global),
            _i(t.npts_local), _i(t.shift_local), _i(t.border_width)]


def collocate_pgf(t, pab=None, func=None, rscale=None, grid=None, libh=None,
                  prefix="oracle"):
    """Single-task collocation of fixture-like object t (see tests/taskfile)."""
    L = libh or lib()
    fn = getattr(L, f"{prefix}_collocate_pgf_product")
    pab = _d(t.pab if pab is None else pab)
    nx, ny, nz = [int(v) for v in t.npts_local]
    if grid is None:
        grid = np.zeros((nz, ny, nx))
    dh, dh_inv, ra, rab, ng, nl, sh, bw = _single_args(t)
    fn(C.c_bool(t.orthorhombic), C.c_int(t.border_mask),
       C.c_int(t.func if func is None else func), C.c_int(t.la_max),
       C.c_int(t.la_min), C.c_int(t.lb_max), C.c_int(t.lb_min),
       C.c_double(t.zeta), C.c_double(t.zetb),
       C.c_double(t.rscale if rscale is None else rscale), _ptr(dh),
       _ptr(dh_inv), _ptr(ra), _ptr(rab), _ptr(ng, C.c_int),
       _ptr(nl, C.c_int), _ptr(sh, C.c_int), _ptr(bw, C.c_int),
       C.c_double(t.radius), C.c_int(t.o1), C.c_int(t.o2), C.c_int(t.n1),
       C.c_int(t.n2), _ptr(pab), _ptr(grid))
    return grid


def integrate_pgf(t, grid=None, compute_tau=None, forces=True, virial=True,
                  libh=None, prefix="oracle", derivs=False):
    """Single-task integration; returns (hab, forces[2,3], virials[2,3,3]),
    with derivs also (hdab [n2,n1,3], hadb [n2,n1,3], a_hdab [n2,n1,3,3])
    (the oracle's full form or the reference's own signature)."""
    L = libh or lib()
    full = derivs and prefix == "oracle"
    fn = getattr(L, f"{prefix}_integrate_pgf_product" + ("_full" if full else ""))
    grid = _d(t.grid if grid is None else grid)
    hab = np.zeros((t.n2, t.n1))
    f = np.zeros((2, 3)) if forces else None
    v = np.zeros((2, 3, 3)) if virial else None
    d3 = [np.zeros((t.n2, t.n1, 3)), np.zeros((t.n2, t.n1, 3)),
          np.zeros((t.n2, t.n1, 3, 3))] if derivs else [None, None, None]
    pab = _d(t.pab)
    dh, dh_inv, ra, rab, ng, nl, sh, bw = _single_args(t)
    tau = t.compute_tau if compute_tau is None else compute_tau
    args = [C.c_bool(t.orthorhombic), C.c_bool(tau), C.c_int(t.border_mask),
            C.c_int(t.la_max), C.c_int(t.la_min), C.c_int(t.lb_max),
            C.c_int(t.lb_min), C.c_double(t.zeta), C.c_double(t.zetb),
            _ptr(dh), _ptr(dh_inv), _ptr(ra), _ptr(rab), _ptr(ng, C.c_int),
            _ptr(nl, C.c_int), _ptr(sh, C.c_int), _ptr(bw, C.c_int),
            C.c_double(t.radius), C.c_int(t.o1), C.c_int(t.o2),
            C.c_int(t.n1), C.c_int(t.n2), _ptr(grid), _ptr(hab), _ptr(pab),
            _ptr(f), _ptr(v)]
    if prefix != "oracle" or full:  # reference signature adds hdab, hadb, a_hdab
        args += [_ptr(x) for x in d3]
    fn(*args)
    if derivs:
        return hab, f, v, d3
    return hab, f, v


_ref = {}


def ref_lib(fast=False):
    """The reference's own single-task kernels (None if not built)."""
    path = REF_LIB_O3 if fast else REF_LIB
    if path not in _ref:
        if not path.exists():
            return None
        L = C.CDLL(str(path))
        L.grid_library_init()
        _ref[path] = L
    return _ref[path]


class TaskList:
    """Oracle task list built from the grid_create_task_list argument set."""

    def __init__(self, a):
        L = lib()
        self.a = a
        self._basis = [make_basis_struct(b) for b in a.basis_sets]
        arr = (C.POINTER(OracleBasis) * len(self._basis))(
            *[C.pointer(s) for s in self._basis])
        self._arr = arr
        self.h = L.oracle_create_task_list(
            C.c_bool(a.orthorhombic), C.c_int(a.ntasks), C.c_int(a.nlevels),
            C.c_int(a.natoms), C.c_int(a.nkinds), C.c_int(a.nblocks),
            _ptr(a.block_offsets, C.c_int), _ptr(a.atom_positions),
            _ptr(a.atom_kinds, C.c_int), arr, _ptr(a.level_list, C.c_int),
            _ptr(a.iatom_list, C.c_int), _ptr(a.jatom_list, C.c_int),
            _ptr(a.iset_list, C.c_int), _ptr(a.jset_list, C.c_int),
            _ptr(a.ipgf_list, C.c_int), _ptr(a.jpgf_list, C.c_int),
            _ptr(a.border_mask_list, C.c_int), _ptr(a.block_num_list, C.c_int),
            _ptr(a.radius_list), _ptr(a.rab_list), _ptr(a.npts_global, C.c_int),
            _ptr(a.npts_local, C.c_int), _ptr(a.shift_local, C.c_int),
            _ptr(a.border_width, C.c_int), _ptr(a.dh), _ptr(a.dh_inv))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_free_task_list(C.c_void_p(self.h))
            self.h = None

    def collocate(self, func, pab_blocks, kernel=None):
        """kernel: None (the oracle's own) or "ref" (the reference's cpu-backend
        kernel grid_cpu_collocate_pgf_product from oracle/_ref)."""
        a = self.a
        grids = [np.zeros((int(n[2]), int(n[1]), int(n[0])))
                 for n in a.npts_local]
        ptrs = (_dp * a.nlevels)(*[_ptr(g) for g in grids])
        args = [C.c_void_p(self.h), C.c_int(func), C.c_int(a.nlevels),
                _ptr(_d(pab_blocks)), ptrs]
        if kernel is None:
            lib().oracle_collocate_task_list(*args)
        else:
            fn = C.cast(ref_lib(fast=True).grid_cpu_collocate_pgf_product, C.c_void_p)
            lib().oracle_collocate_task_list_with(*args, fn)
        return grids

    def integrate(self, grids, pab_blocks=None, compute_tau=False,
                  forces=False, virial=False, hab_len=None, kernel=None):
        a = self.a
        hab_len = int(a.buffer_size if hab_len is None else hab_len)
        hab = np.zeros(hab_len)
        f = np.zeros((a.natoms, 3)) if forces else None
        v = np.zeros((3, 3)) if virial else None
        gs = [_d(g) for g in grids]
        ptrs = (_dp * a.nlevels)(*[_ptr(g) for g in gs])
        pab = None if pab_blocks is None else _d(pab_blocks)
        args = [C.c_void_p(self.h), C.c_bool(compute_tau), C.c_int(a.natoms),
                C.c_int(a.nlevels), _ptr(pab), ptrs, _ptr(hab), C.c_long(hab_len),
                _ptr(f), _ptr(v)]
        if kernel is None:
            lib().oracle_integrate_task_list(*args)
        else:
            fn = C.cast(ref_lib(fast=True).grid_cpu_integrate_pgf_product, C.c_void_p)
            lib().oracle_integrate_task_list_with(*args, fn)
        return hab, f, v

    def task_points(self):
        return int(lib().oracle_task_list_points(C.c_void_p(self.h), 0))


def threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


# ---- task-list generation (oracle/task_gen.c) ----
def generate_task_list(nl):
    """Oracle restatement of the device task-list generator for
    cp2k_amd.tasklist.NeighbourList nl (same output dict)."""
    from cp2k_amd.tasklist import call_generate
    structs = [make_basis_struct(b) for b in nl.basis_sets]
    arr = (C.POINTER(OracleBasis) * len(structs))(*[C.pointer(s) for s in structs])
    return call_generate(lib().oracle_generate_task_list, nl, arr)


def exp_radius(l, alpha, threshold, prefactor, epsabs=1.0e-2, rlow=0.0, bracket=False):
    """exp_radius (src/aobasis/ao_util.F:95-190); bracket=True returns
    (lower, upper) ends of the final bracket."""
    L = lib()
    L.oracle_exp_radius_bracket.restype = C.c_double
    lo = C.c_double(0.0)
    hi = L.oracle_exp_radius_bracket(C.c_int(l), C.c_double(alpha), C.c_double(threshold),
                                     C.c_double(prefactor), C.c_double(epsabs),
                                     C.c_double(rlow), C.byref(lo))
    return (lo.value, hi) if bracket else hi


def product_polycoef(la_max, lb_max, rad_a, rad_b):
    poly = np.zeros(la_max + lb_max + 1)
    lib().oracle_product_polycoef(C.c_int(la_max), C.c_int(lb_max), C.c_double(rad_a),
                                  C.c_double(rad_b), _ptr(poly))
    return poly


def gridlevel(cutoff, rel_cutoff, exponent):
    c = _d(cutoff)
    lib().oracle_gridlevel.restype = C.c_int
    return lib().oracle_gridlevel(C.c_int(len(c)), _ptr(c), C.c_double(rel_cutoff),
                                  C.c_double(exponent))
