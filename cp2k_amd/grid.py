"""Python binding of libcp2k_grid_mi355x.so, shaped like CP2K's grid_api.F.

`GridTaskList` mirrors the Fortran module's task-list calls
(src/grid/grid_api.F:721-1129): create from the per-task lists, collocate a
density (per grid level), integrate a potential into Hamiltonian blocks,
forces and virial.  Calls go through the library's C-ABI exactly as CP2K's
ISO_C_BINDING interfaces do.  There is no fallback: if the HIP library is not
built, or no GPU is visible, the calls fail.
"""
import ctypes as C
import pathlib

import numpy as np

from .build import LIB as LIB_PATH

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

# exported by the library; checked by tests/test_abi.py against include/*.h
EXPORTS = [
    "grid_create_task_list", "grid_free_task_list", "grid_collocate_task_list",
    "grid_integrate_task_list", "grid_create_basis_set", "grid_free_basis_set",
    "grid_library_init", "grid_library_finalize", "grid_library_set_config",
    "grid_library_get_config", "grid_library_print_stats",
    "grid_library_counter_add", "offload_create_buffer", "offload_free_buffer",
    "offload_get_buffer_host_pointer", "grid_gpu_create_task_list",
    "grid_gpu_free_task_list", "grid_gpu_collocate_task_list",
    "grid_gpu_integrate_task_list", "grid_mi355x_set_precision",
    "grid_mi355x_get_precision", "grid_mi355x_collocate_task_list_dev",
    "grid_mi355x_integrate_task_list_dev", "grid_mi355x_device_malloc",
    "grid_mi355x_device_free", "grid_mi355x_memcpy", "grid_mi355x_set_stream",
    "grid_mi355x_get_stream", "grid_mi355x_synchronize",
    "grid_mi355x_set_timing", "grid_mi355x_reset_timing",
    "grid_mi355x_kernel_time_ms", "grid_mi355x_task_points",
    "grid_mi355x_task_points_by_lp", "grid_mi355x_task_counts",
    "grid_mi355x_create_batch", "grid_mi355x_free_batch", "grid_mi355x_batch_buffers",
    "grid_mi355x_batch_grid", "grid_mi355x_batch_run", "grid_mi355x_batch_task_points",
    "grid_mi355x_autotune", "grid_mi355x_tuned_shape", "grid_mi355x_point_launches",
    "grid_mi355x_collocate_task_list_noalloc",
    "grid_mi355x_collocate_pgf_product", "grid_mi355x_integrate_pgf_product",
    "grid_mi355x_pack_buffer", "grid_mi355x_unpack_buffer", "grid_mi355x_buffer_set_resident",
    "grid_mi355x_buffer_sync_host", "grid_mi355x_buffer_host_written",
    "grid_mi355x_buffer_device_current",
    "grid_mi355x_density_rs2pw", "grid_mi355x_potential_pw2rs", "grid_mi355x_transfer_rs2pw",
    "grid_mi355x_transfer_pw2rs", "grid_mi355x_batch_set_work_queue",
    "grid_mi355x_batch_work_queue", "grid_mi355x_autotune_batch",
    "grid_mi355x_generate_task_list", "grid_mi355x_generate_task_list_ms",
]

GRID_FUNC_AB = 100
GRID_FUNC_DADB = 200
FP64 = 0
FP32_MIXED = 1

KERNEL_COLLOCATE, KERNEL_INTEGRATE, KERNEL_DECONTRACT, KERNEL_CONTRACT, \
    KERNEL_FILL, KERNEL_REDUCE, KERNEL_CXYZ, KERNEL_TILES, KERNEL_DIRECT, \
    KERNEL_INTEGRATE_POINTS, KERNEL_FINISH, KERNEL_PER_TASK, KERNEL_TO_FLOAT, \
    KERNEL_GENERAL_POINTS, KERNEL_BATCH = range(15)
KERNEL_NAMES = ["collocate", "integrate", "decontract", "contract", "fill", "reduce",
                "cxyz", "collocate_tiles", "collocate_direct", "integrate_points",
                "finish", "per_task", "to_float", "general_points", "batch"]


class GridBasisSet(C.Structure):
    """grid_basis_set, src/grid/common/grid_basis_set.h:14-26."""
    _fields_ = [("nset", C.c_int), ("nsgf", C.c_int), ("maxco", C.c_int),
                ("maxpgf", C.c_int), ("lmin", _ip), ("lmax", _ip),
                ("npgf", _ip), ("nsgf_set", _ip), ("first_sgf", _ip),
                ("sphi", _dp), ("zet", _dp)]


class OffloadBuffer(C.Structure):
    """offload_buffer, src/offload/offload_buffer.h:16-20."""
    _fields_ = [("size", C.c_size_t), ("host_buffer", _dp),
                ("device_buffer", _dp)]


class LibraryConfig(C.Structure):
    _fields_ = [("backend", C.c_int), ("validate", C.c_bool),
                ("apply_cutoff", C.c_bool)]


_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not pathlib.Path(LIB_PATH).exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run "
                               "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(str(LIB_PATH))
        L.grid_create_basis_set.restype = None
        L.grid_free_basis_set.argtypes = [C.c_void_p]
        L.grid_create_task_list.restype = None
        L.grid_free_task_list.argtypes = [C.c_void_p]
        L.grid_collocate_task_list.restype = None
        L.grid_integrate_task_list.restype = None
        L.grid_library_get_config.restype = LibraryConfig
        L.offload_create_buffer.argtypes = [C.c_int, C.POINTER(C.POINTER(OffloadBuffer))]
        L.offload_free_buffer.argtypes = [C.POINTER(OffloadBuffer)]
        L.grid_mi355x_kernel_time_ms.restype = C.c_double
        L.grid_mi355x_kernel_time_ms.argtypes = [C.c_int, C.POINTER(C.c_long)]
        L.grid_mi355x_point_launches.argtypes = [C.c_int, C.POINTER(C.c_long),
                                                  C.POINTER(C.c_long)]
        L.grid_mi355x_task_points.restype = C.c_longlong
        L.grid_mi355x_task_points.argtypes = [C.c_void_p]
        L.grid_mi355x_task_points_by_lp.restype = None
        L.grid_mi355x_task_points_by_lp.argtypes = [C.c_void_p, C.POINTER(C.c_longlong), C.c_int]
        L.grid_mi355x_get_stream.restype = C.c_void_p
        L.grid_mi355x_set_stream.argtypes = [C.c_void_p]
        L.grid_mi355x_set_timing.argtypes = [C.c_bool]
        L.grid_mi355x_create_batch.argtypes = [C.c_int, C.POINTER(C.c_void_p),
                                               C.POINTER(C.c_void_p)]
        L.grid_mi355x_free_batch.argtypes = [C.c_void_p]
        L.grid_mi355x_batch_buffers.argtypes = [C.c_void_p, C.c_int] + [C.POINTER(C.c_void_p)] * 2 + [
            C.POINTER(C.c_long)] + [C.POINTER(C.c_void_p)] * 2
        L.grid_mi355x_batch_grid.restype = C.c_void_p
        L.grid_mi355x_batch_grid.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.grid_mi355x_batch_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_bool, C.c_bool]
        L.grid_mi355x_batch_task_points.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
        L.grid_mi355x_batch_set_work_queue.argtypes = [C.c_void_p, C.c_bool]
        L.grid_mi355x_batch_work_queue.argtypes = [C.c_void_p]
        L.grid_mi355x_autotune_batch.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_bool, C.c_bool]
        L.grid_mi355x_device_malloc.restype = C.c_void_p
        L.grid_mi355x_device_malloc.argtypes = [C.c_size_t]
        L.grid_mi355x_device_free.argtypes = [C.c_void_p]
        L.grid_mi355x_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.grid_mi355x_collocate_pgf_product.restype = None
        L.grid_mi355x_density_rs2pw.argtypes = [C.c_int, _ip, C.POINTER(C.c_void_p), C.c_void_p,
                                                C.c_void_p]
        L.grid_mi355x_potential_pw2rs.argtypes = [C.c_int, _ip, C.c_void_p, C.POINTER(C.c_void_p)]
        L.grid_mi355x_transfer_rs2pw.argtypes = [_ip, _ip, C.c_void_p, C.c_void_p]
        L.grid_mi355x_transfer_pw2rs.argtypes = [_ip, _ip, C.c_void_p, C.c_void_p]
        L.grid_mi355x_pack_buffer.argtypes = [C.c_int, C.POINTER(C.c_void_p), _ip, _ip,
                                              C.POINTER(OffloadBuffer)]
        L.grid_mi355x_unpack_buffer.argtypes = [C.c_int, C.POINTER(C.c_void_p), _ip, _ip,
                                                C.POINTER(OffloadBuffer)]
        L.grid_mi355x_buffer_set_resident.argtypes = [C.POINTER(OffloadBuffer), C.c_bool]
        L.grid_mi355x_buffer_sync_host.argtypes = [C.POINTER(OffloadBuffer)]
        L.grid_mi355x_buffer_host_written.argtypes = [C.POINTER(OffloadBuffer)]
        L.grid_mi355x_buffer_device_current.argtypes = [C.POINTER(OffloadBuffer)]
        L.grid_mi355x_integrate_pgf_product.restype = None
        L.grid_mi355x_generate_task_list.restype = C.c_longlong
        L.grid_mi355x_generate_task_list_ms.restype = C.c_double
        _lib = L
    return _lib


def _ptr(a, ctype=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


class Basis:
    """A grid_basis_set created through the library."""

    def __init__(self, spec):
        L = lib()
        self.spec = spec
        self._arrays = [np.ascontiguousarray(x, dtype=np.int32) for x in
                        (spec.lmin, spec.lmax, spec.npgf, spec.nsgf_set, spec.first_sgf)]
        self._sphi = np.ascontiguousarray(spec.sphi, dtype=np.float64)
        self._zet = np.ascontiguousarray(spec.zet, dtype=np.float64)
        self.h = C.POINTER(GridBasisSet)()
        L.grid_create_basis_set(
            C.c_int(spec.nset), C.c_int(spec.nsgf), C.c_int(spec.maxco),
            C.c_int(spec.maxpgf), *[_ptr(a, C.c_int) for a in self._arrays],
            _ptr(self._sphi), _ptr(self._zet), C.byref(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().grid_free_basis_set(self.h)
            self.h = None


class OffloadBuf:
    """An offload_buffer (pinned host + HBM mirror) created by the library."""

    def __init__(self, length):
        self.h = C.POINTER(OffloadBuffer)()
        lib().offload_create_buffer(C.c_int(int(length)), C.byref(self.h))
        self.length = int(length)

    def host(self):
        if self.length == 0:
            return np.zeros(0)
        return np.ctypeslib.as_array(self.h.contents.host_buffer, shape=(self.length,))

    # ---- residency (include/grid_buffers.h) ----
    def set_resident(self, resident=True):
        lib().grid_mi355x_buffer_set_resident(self.h, C.c_bool(resident))

    def sync_host(self):
        lib().grid_mi355x_buffer_sync_host(self.h)

    def host_written(self):
        lib().grid_mi355x_buffer_host_written(self.h)

    def device_current(self):
        return bool(lib().grid_mi355x_buffer_device_current(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().offload_free_buffer(self.h)
            self.h = None


class DeviceArray:
    """A float64 array in HBM allocated through the library (no torch in the
    process: torch ships its own HIP runtime, which cannot share a process
    with the system ROCm runtime this library links)."""

    def __init__(self, shape):
        self.shape = tuple(int(x) for x in np.atleast_1d(shape)) if np.ndim(shape) else (int(shape),)
        self.size = int(np.prod(self.shape))
        self.nbytes = 8 * self.size
        self.ptr = lib().grid_mi355x_device_malloc(C.c_size_t(self.nbytes))

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.shape)
        lib().grid_mi355x_memcpy(C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p),
                                 C.c_size_t(d.nbytes), C.c_int(1))
        return d

    def numpy(self):
        a = np.empty(self.shape)
        lib().grid_mi355x_memcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                 C.c_size_t(self.nbytes), C.c_int(2))
        return a

    def copy_from(self, other):
        """Device-to-device copy (HIP's blit kernel) of another DeviceArray."""
        if other.nbytes != self.nbytes:
            raise ValueError(f"copy of {other.nbytes} bytes into {self.nbytes}")
        lib().grid_mi355x_memcpy(C.c_void_p(self.ptr), C.c_void_p(other.ptr),
                                 C.c_size_t(self.nbytes), C.c_int(3))

    def data_ptr(self):
        return self.ptr

    def numel(self):
        return self.size

    def free(self):
        if getattr(self, "ptr", None):
            lib().grid_mi355x_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        self.free()


def set_precision(p):
    lib().grid_mi355x_set_precision(C.c_int(int(p)))


def get_precision():
    return int(lib().grid_mi355x_get_precision())


class GridTaskList:
    """A task list on the GPU, created from a TaskListArgs."""

    def __init__(self, args):
        L = lib()
        self.args = a = args
        self._bases = [Basis(b) for b in a.basis_sets]
        self._basis_arr = (C.POINTER(GridBasisSet) * len(self._bases))(
            *[b.h for b in self._bases])
        self.h = C.c_void_p()
        L.grid_create_task_list(
            C.c_bool(a.orthorhombic), C.c_int(a.ntasks), C.c_int(a.nlevels),
            C.c_int(a.natoms), C.c_int(a.nkinds), C.c_int(a.nblocks),
            _ptr(a.block_offsets, C.c_int), _ptr(a.atom_positions),
            _ptr(a.atom_kinds, C.c_int), self._basis_arr,
            _ptr(a.level_list, C.c_int), _ptr(a.iatom_list, C.c_int),
            _ptr(a.jatom_list, C.c_int), _ptr(a.iset_list, C.c_int),
            _ptr(a.jset_list, C.c_int), _ptr(a.ipgf_list, C.c_int),
            _ptr(a.jpgf_list, C.c_int), _ptr(a.border_mask_list, C.c_int),
            _ptr(a.block_num_list, C.c_int), _ptr(a.radius_list),
            _ptr(a.rab_list), _ptr(a.npts_global, C.c_int),
            _ptr(a.npts_local, C.c_int), _ptr(a.shift_local, C.c_int),
            _ptr(a.border_width, C.c_int), _ptr(a.dh), _ptr(a.dh_inv),
            C.byref(self.h))
        self._npts = np.ascontiguousarray(a.npts_local, dtype=np.int32)

    def free(self):
        if getattr(self, "h", None) and self.h.value:
            lib().grid_free_task_list(self.h)
        self.h = None

    def __del__(self):
        self.free()

    # ---- host-buffer API (the CP2K ABI) ----
    def collocate(self, func, pab_blocks):
        """grid_collocate_task_list: returns one [nz, ny, nx] array per level."""
        a = self.args
        pab = OffloadBuf(a.buffer_size)
        pab.host()[:] = pab_blocks
        grids = [OffloadBuf(int(np.prod(s))) for s in a.grid_shapes()]
        arr = (C.POINTER(OffloadBuffer) * a.nlevels)(*[g.h for g in grids])
        lib().grid_collocate_task_list(self.h, C.c_int(func), C.c_int(a.nlevels),
                                       _ptr(self._npts, C.c_int), pab.h, arr)
        return [g.host().reshape(s).copy() for g, s in zip(grids, a.grid_shapes())]

    def integrate(self, grids, pab_blocks=None, compute_tau=False, forces=False,
                  virial=False):
        """grid_integrate_task_list: returns (hab_blocks, forces, virial)."""
        a = self.args
        gbufs = []
        for g, s in zip(grids, a.grid_shapes()):
            b = OffloadBuf(int(np.prod(s)))
            b.host()[:] = np.ascontiguousarray(g, dtype=np.float64).ravel()
            gbufs.append(b)
        arr = (C.POINTER(OffloadBuffer) * a.nlevels)(*[g.h for g in gbufs])
        pab = None
        if pab_blocks is not None:
            pab = OffloadBuf(a.buffer_size)
            pab.host()[:] = pab_blocks
        hab = OffloadBuf(a.buffer_size)
        f = np.zeros((a.natoms, 3)) if forces else None
        v = np.zeros((3, 3)) if virial else None
        lib().grid_integrate_task_list(
            self.h, C.c_bool(compute_tau), C.c_int(a.natoms), C.c_int(a.nlevels),
            _ptr(self._npts, C.c_int), pab.h if pab is not None else None, arr,
            hab.h, _ptr(f), _ptr(v))
        return hab.host().copy(), f, v

    # ---- HBM-resident API (extension) on DeviceArray buffers ----
    def collocate_dev(self, func, pab, grids):
        """pab, grids[level]: DeviceArray (anything with data_ptr())."""
        a = self.args
        gp = (C.c_void_p * a.nlevels)(*[g.data_ptr() for g in grids])
        lib().grid_mi355x_collocate_task_list_dev(
            self.h, C.c_int(func), C.c_int(a.nlevels), C.c_void_p(pab.data_ptr()), gp)

    def collocate_noalloc(self, func, pab, grids):
        """grid_mi355x_collocate_task_list_noalloc on device buffers."""
        a = self.args
        gp = (C.c_void_p * a.nlevels)(*[g.data_ptr() for g in grids])
        lib().grid_mi355x_collocate_task_list_noalloc(
            self.h, C.c_int(func), C.c_int(a.nlevels), C.c_void_p(pab.data_ptr()), gp)

    def integrate_dev(self, grids, hab, pab=None, forces=None, virial=None,
                      compute_tau=False):
        a = self.args
        gp = (C.c_void_p * a.nlevels)(*[g.data_ptr() for g in grids])
        lib().grid_mi355x_integrate_task_list_dev(
            self.h, C.c_bool(compute_tau), C.c_int(a.natoms), C.c_int(a.nlevels),
            C.c_void_p(pab.data_ptr() if pab is not None else 0), gp,
            C.c_void_p(hab.data_ptr()), C.c_long(hab.numel()),
            C.c_void_p(forces.data_ptr() if forces is not None else 0),
            C.c_void_p(virial.data_ptr() if virial is not None else 0))

    def task_points(self):
        return int(lib().grid_mi355x_task_points(self.h))

    def autotune(self, func, pab, grids):
        """grid_mi355x_autotune on device buffers (DeviceArray or views)."""
        a = self.args
        gp = (C.c_void_p * a.nlevels)(*[g.data_ptr() for g in grids])
        lib().grid_mi355x_autotune(self.h, C.c_int(func), C.c_void_p(pab.data_ptr()), gp)

    def tuned_shape(self, level, general=False):
        r, w = C.c_int(), C.c_int()
        lib().grid_mi355x_tuned_shape(self.h, C.c_int(level), C.c_int(1 if general else 0),
                                      C.byref(r), C.byref(w))
        return r.value, w.value

    def task_points_by_lp(self, n=11):
        """Task points per la_max + lb_max of the tasks (index = lp)."""
        out = (C.c_longlong * n)()
        lib().grid_mi355x_task_points_by_lp(self.h, out, C.c_int(n))
        return [int(x) for x in out]

    def task_counts(self):
        o, g, s = C.c_long(), C.c_long(), C.c_long()
        lib().grid_mi355x_task_counts(self.h, C.byref(o), C.byref(g), C.byref(s))
        return o.value, g.value, s.value


class AbiBuffers:
    """Persistent offload_buffers of one task list for repeated calls of the
    host-buffer C-ABI (grid_collocate_task_list / grid_integrate_task_list),
    as CP2K keeps its buffers between SCF steps: every call copies pab to the
    device and grids / hab back to the host."""

    def __init__(self, gt):
        a = gt.args
        self.gt = gt
        self.pab = OffloadBuf(a.buffer_size)
        self.hab = OffloadBuf(a.buffer_size)
        self.rho = [OffloadBuf(int(np.prod(s))) for s in a.grid_shapes()]
        self.pot = [OffloadBuf(int(np.prod(s))) for s in a.grid_shapes()]
        self._rho = (C.POINTER(OffloadBuffer) * a.nlevels)(*[g.h for g in self.rho])
        self._pot = (C.POINTER(OffloadBuffer) * a.nlevels)(*[g.h for g in self.pot])
        self.forces = np.zeros((a.natoms, 3))
        self.virial = np.zeros((3, 3))

    def set_inputs(self, pab_blocks, V):
        self.pab.host()[:] = pab_blocks
        for b, v in zip(self.pot, V):
            b.host()[:] = np.ascontiguousarray(v, dtype=np.float64).ravel()

    def collocate(self, func):
        a = self.gt.args
        lib().grid_collocate_task_list(self.gt.h, C.c_int(func), C.c_int(a.nlevels),
                                       _ptr(self.gt._npts, C.c_int), self.pab.h, self._rho)

    def integrate(self, forces=False, virial=False, compute_tau=False):
        a = self.gt.args
        lib().grid_integrate_task_list(
            self.gt.h, C.c_bool(compute_tau), C.c_int(a.natoms), C.c_int(a.nlevels),
            _ptr(self.gt._npts, C.c_int), self.pab.h if (forces or virial) else None,
            self._pot, self.hab.h, _ptr(self.forces) if forces else None,
            _ptr(self.virial) if virial else None)

    def free(self):
        for b in [self.pab, self.hab] + self.rho + self.pot:
            b.__del__()


# ---- matrix blocks <-> buffers (include/grid_buffers.h) ----
def _block_args(blocks, pair_offsets):
    ptrs = (C.c_void_p * len(blocks))(*[
        b.ptr if isinstance(b, DeviceArray) else b.ctypes.data for b in blocks])
    sizes = np.array([b.size for b in blocks], dtype=np.int32)
    offs = np.ascontiguousarray(pair_offsets, dtype=np.int32)
    return ptrs, sizes, offs


def pack_buffer(blocks, pair_offsets, dest):
    """grid_mi355x_pack_buffer: blocks (all DeviceArray or all float64 numpy
    arrays, column-major data) into OffloadBuf dest at pair_offsets."""
    ptrs, sizes, offs = _block_args(blocks, pair_offsets)
    lib().grid_mi355x_pack_buffer(C.c_int(len(blocks)), ptrs, _ptr(sizes, C.c_int),
                                  _ptr(offs, C.c_int), dest.h)


def unpack_buffer(blocks, pair_offsets, src):
    """grid_mi355x_unpack_buffer: blocks[i] += src at pair_offsets[i]."""
    ptrs, sizes, offs = _block_args(blocks, pair_offsets)
    lib().grid_mi355x_unpack_buffer(C.c_int(len(blocks)), ptrs, _ptr(sizes, C.c_int),
                                    _ptr(offs, C.c_int), src.h)


# ---- real-space multigrids <-> plane-wave grids (include/grid_rs_pw.h) ----
def _npts(shapes):
    """{nx, ny, nz} per level from [nz, ny, nx] shapes."""
    return np.ascontiguousarray([list(s)[::-1] for s in shapes], dtype=np.int32)


def density_rs2pw(rs_grids, want_g=True):
    """grid_mi355x_density_rs2pw on DeviceArray level grids ([nz, ny, nx],
    level 0 the finest): returns (rho_r, rho_g) as DeviceArrays, rho_g of
    shape [nz, ny, nx, 2] (real, imaginary) or None."""
    shapes = [g.shape for g in rs_grids]
    npts = _npts(shapes)
    rho_r = DeviceArray(shapes[0])
    rho_g = DeviceArray(tuple(shapes[0]) + (2,)) if want_g else None
    ptrs = (C.c_void_p * len(rs_grids))(*[g.ptr for g in rs_grids])
    lib().grid_mi355x_density_rs2pw(C.c_int(len(rs_grids)), _ptr(npts, C.c_int), ptrs,
                                    C.c_void_p(rho_r.ptr),
                                    C.c_void_p(rho_g.ptr if rho_g is not None else 0))
    return rho_r, rho_g


def potential_pw2rs(v_r, shapes):
    """grid_mi355x_potential_pw2rs: DeviceArray level grids of the given
    [nz, ny, nx] shapes (shapes[0] is v_r's) from the DeviceArray v_r."""
    out = [DeviceArray(s) for s in shapes]
    npts = _npts(shapes)
    ptrs = (C.c_void_p * len(out))(*[g.ptr for g in out])
    lib().grid_mi355x_potential_pw2rs(C.c_int(len(out)), _ptr(npts, C.c_int),
                                      C.c_void_p(v_r.ptr), ptrs)
    return out


def transfer_rs2pw(rs, border):
    """Real region of a halo grid (DeviceArray [nz + 2bz, ny + 2by, nx + 2bx]);
    border = (bx, by, bz)."""
    b = np.ascontiguousarray(border, dtype=np.int32)
    shape = tuple(n - 2 * w for n, w in zip(rs.shape, b[::-1]))
    pw = DeviceArray(shape)
    lib().grid_mi355x_transfer_rs2pw(_ptr(_npts([shape])[0], C.c_int), _ptr(b, C.c_int),
                                     C.c_void_p(rs.ptr), C.c_void_p(pw.ptr))
    return pw


def transfer_pw2rs(pw, border):
    b = np.ascontiguousarray(border, dtype=np.int32)
    shape = tuple(n + 2 * w for n, w in zip(pw.shape, b[::-1]))
    rs = DeviceArray(shape)
    lib().grid_mi355x_transfer_pw2rs(_ptr(_npts([pw.shape])[0], C.c_int), _ptr(b, C.c_int),
                                     C.c_void_p(pw.ptr), C.c_void_p(rs.ptr))
    return rs


# ---- task-list generation (include/grid_task_gen.h) ----
def generate_task_list(nl):
    """Task list of neighbour list nl (tasklist.NeighbourList), generated on
    the device: dict of the grid_create_task_list arrays (reference order),
    nblocks, block_offsets, buffer_size, ntasks and device_ms (device time of
    the generation)."""
    from .tasklist import call_generate
    bases = [Basis(b) for b in nl.basis_sets]
    arr = (C.POINTER(GridBasisSet) * len(bases))(*[b.h for b in bases])
    out = call_generate(lib().grid_mi355x_generate_task_list, nl, arr)
    out["device_ms"] = float(lib().grid_mi355x_generate_task_list_ms())
    return out


def create_generated_task_list(nl, gen=None):
    """(GridTaskList, TaskListArgs) of neighbour list nl from the device
    generator."""
    gen = generate_task_list(nl) if gen is None else gen
    args = nl.task_list_args(gen)
    return GridTaskList(args), args


# ---- single task (include/grid_pgf_product.h), shaped like the Fortran
#      collocate_pgf_product / integrate_pgf_product, src/grid/grid_api.F:110-490 ----
def _d64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def _i32(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def _pgf_geometry(t):
    return [_d64(t.dh), _d64(t.dh_inv), _d64(t.ra), _d64(t.rab), _i32(t.npts_global),
            _i32(t.npts_local), _i32(t.shift_local), _i32(t.border_width)]


def _arg(x, ctype=C.c_double):
    """A host numpy array or a DeviceArray (device memory) as a pointer."""
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return C.cast(C.c_void_p(x.ptr), C.POINTER(ctype))
    return _ptr(x, ctype)


def collocate_pgf(t, pab=None, func=None, rscale=None, grid=None):
    """grid_mi355x_collocate_pgf_product of task t (attributes as the
    reference's .task files: l ranges, exponents, geometry, o1/o2/n1/n2).
    grid: host array [nz, ny, nx] or DeviceArray, accumulated into (a new
    zero host grid if None); returns it."""
    pab = _d64(t.pab if pab is None else pab) if not isinstance(pab, DeviceArray) else pab
    nx, ny, nz = [int(v) for v in t.npts_local]
    if grid is None:
        grid = np.zeros((nz, ny, nx))
    g = _pgf_geometry(t)
    lib().grid_mi355x_collocate_pgf_product(
        C.c_bool(bool(t.orthorhombic)), C.c_int(int(t.border_mask)),
        C.c_int(int(t.func if func is None else func)), C.c_int(int(t.la_max)),
        C.c_int(int(t.la_min)), C.c_int(int(t.lb_max)), C.c_int(int(t.lb_min)),
        C.c_double(t.zeta), C.c_double(t.zetb),
        C.c_double(t.rscale if rscale is None else rscale), _ptr(g[0]), _ptr(g[1]),
        _ptr(g[2]), _ptr(g[3]), *[_ptr(x, C.c_int) for x in g[4:]], C.c_double(t.radius),
        C.c_int(int(t.o1)), C.c_int(int(t.o2)), C.c_int(int(t.n1)), C.c_int(int(t.n2)),
        _arg(pab), _arg(grid))
    return grid


def integrate_pgf(t, grid=None, compute_tau=False, forces=True, virial=True,
                  derivs=False, hab=None):
    """grid_mi355x_integrate_pgf_product of task t: returns (hab [n2, n1],
    forces [2, 3] or None, virials [2, 3, 3] or None) and with derivs also
    (hdab, hadb, a_hdab).  grid: host array or DeviceArray (t.grid if None);
    hab: host array or DeviceArray accumulated into (zeros if None)."""
    grid = _d64(t.grid if grid is None else grid) if not isinstance(grid, DeviceArray) else grid
    if hab is None:
        hab = np.zeros((int(t.n2), int(t.n1)))
    f = np.zeros((2, 3)) if forces else None
    v = np.zeros((2, 3, 3)) if virial else None
    n2, n1 = int(t.n2), int(t.n1)
    d3 = [np.zeros((n2, n1, 3)), np.zeros((n2, n1, 3)),
          np.zeros((n2, n1, 3, 3))] if derivs else [None, None, None]
    g = _pgf_geometry(t)
    lib().grid_mi355x_integrate_pgf_product(
        C.c_bool(bool(t.orthorhombic)), C.c_bool(bool(compute_tau)),
        C.c_int(int(t.border_mask)), C.c_int(int(t.la_max)), C.c_int(int(t.la_min)),
        C.c_int(int(t.lb_max)), C.c_int(int(t.lb_min)), C.c_double(t.zeta),
        C.c_double(t.zetb), _ptr(g[0]), _ptr(g[1]), _ptr(g[2]), _ptr(g[3]),
        *[_ptr(x, C.c_int) for x in g[4:]], C.c_double(t.radius), C.c_int(int(t.o1)),
        C.c_int(int(t.o2)), C.c_int(n1), C.c_int(n2), _arg(grid), _arg(hab),
        _arg(_d64(t.pab)), _ptr(f), _ptr(v), *[_ptr(x) for x in d3])
    if derivs:
        return hab, f, v, d3
    return hab, f, v


def set_stream(stream_handle):
    lib().grid_mi355x_set_stream(C.c_void_p(stream_handle or 0))


def synchronize():
    lib().grid_mi355x_synchronize()


def set_timing(enable):
    lib().grid_mi355x_set_timing(C.c_bool(enable))


def reset_timing():
    lib().grid_mi355x_reset_timing()


def kernel_time_ms(kernel_id):
    n = C.c_long()
    ms = lib().grid_mi355x_kernel_time_ms(C.c_int(kernel_id), C.byref(n))
    return ms, n.value


def point_launches(kernel_id):
    """(FP64, FP32) eager launches of the point kernels of kernel_id since
    the last reset_timing()."""
    a, b = C.c_long(), C.c_long()
    lib().grid_mi355x_point_launches(C.c_int(kernel_id), C.byref(a), C.byref(b))
    return a.value, b.value


class _DevView:
    """A device range owned by someone else (a batch buffer), shaped."""

    def __init__(self, ptr, shape):
        self.ptr = int(ptr)
        self.shape = tuple(shape)
        self.size = int(np.prod(self.shape))
        self.nbytes = 8 * self.size

    def data_ptr(self):
        return self.ptr

    def numel(self):
        return self.size

    def numpy(self):
        a = np.empty(self.shape)
        lib().grid_mi355x_memcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                 C.c_size_t(self.nbytes), C.c_int(2))
        return a

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != self.size:
            raise ValueError(f"{a.size} values for a buffer of {self.size}")
        lib().grid_mi355x_memcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p),
                                 C.c_size_t(self.nbytes), C.c_int(1))


class GridBatch:
    """Several GridTaskLists merged into one batch (grid_mi355x_create_batch):
    every launch covers all lists, a run replays a captured hipGraph.  The
    batch owns the HBM buffers of each list (pab/hab, grids, forces, virial)."""

    def __init__(self, task_lists):
        self.lists = list(task_lists)
        arr = (C.c_void_p * len(self.lists))(*[t.h.value for t in self.lists])
        self.h = C.c_void_p()
        lib().grid_mi355x_create_batch(C.c_int(len(self.lists)), arr, C.byref(self.h))

    def free(self):
        if getattr(self, "h", None) and self.h.value:
            lib().grid_mi355x_free_batch(self.h)
        self.h = None

    def __del__(self):
        self.free()

    def buffers(self, i):
        """(pab, hab, forces, virial) device views of list i."""
        p, h, f, v = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        n = C.c_long()
        lib().grid_mi355x_batch_buffers(self.h, C.c_int(i), C.byref(p), C.byref(h), C.byref(n),
                                        C.byref(f), C.byref(v))
        a = self.lists[i].args
        return (_DevView(p.value, (n.value,)), _DevView(h.value, (n.value,)),
                _DevView(f.value, (a.natoms, 3)), _DevView(v.value, (3, 3)))

    def grid(self, i, level, potential=False):
        shape = tuple(int(x) for x in self.lists[i].args.grid_shapes()[level])
        ptr = lib().grid_mi355x_batch_grid(self.h, C.c_int(i), C.c_int(level),
                                           C.c_int(1 if potential else 0))
        return _DevView(ptr, shape)

    def set_inputs(self, i, pab=None, V=None):
        """Uploads list i's density blocks (its buffer_size values) and/or
        potential grids."""
        if pab is not None:
            buf = self.buffers(i)[0]
            full = np.zeros(buf.size)
            n = min(len(pab), buf.size)  # the blocks the list addresses
            full[:n] = pab[:n]
            buf.upload(full)
        if V is not None:
            for l, v in enumerate(V):
                self.grid(i, l, potential=True).upload(v)

    def run(self, what=3, func=GRID_FUNC_AB, compute_tau=False, forces=False):
        lib().grid_mi355x_batch_run(self.h, C.c_int(what), C.c_int(func), C.c_bool(compute_tau),
                                    C.c_bool(forces))

    def set_work_queue(self, enable):
        """Persistent launches over device work queues (True) or one
        workgroup per unit (False) for the batch's point kernels."""
        lib().grid_mi355x_batch_set_work_queue(self.h, C.c_bool(enable))

    def work_queue(self):
        return bool(lib().grid_mi355x_batch_work_queue(self.h))

    def autotune(self, what=3, func=GRID_FUNC_AB, compute_tau=False, forces=False):
        """grid_mi355x_autotune_batch: keeps the faster launch shape; True if
        the work queue won."""
        return bool(lib().grid_mi355x_autotune_batch(self.h, C.c_int(what), C.c_int(func),
                                                     C.c_bool(compute_tau), C.c_bool(forces)))

    def outputs(self, i):
        """(rho levels, hab, forces, virial) of list i, on the host."""
        a = self.lists[i].args
        pab, hab, f, v = self.buffers(i)
        rho = [self.grid(i, l).numpy() for l in range(a.nlevels)]
        h = np.zeros(a.buffer_size)
        hv = hab.numpy()
        n = min(len(hv), a.buffer_size)
        h[:n] = hv[:n]
        return rho, h, f.numpy(), v.numpy()

    def task_points(self):
        n = C.c_longlong()
        lib().grid_mi355x_batch_task_points(self.h, C.byref(n))
        return int(n.value)
