"""Host-side description of one grid task list.

`TaskListArgs` holds exactly the arguments of grid_create_task_list
(src/grid/grid_task_list.h:59-68) as numpy arrays, plus the density-matrix
buffer size.  `BasisSpec` mirrors grid_basis_set (src/grid/common/
grid_basis_set.h:14-26).  Both are plain data: the same object feeds this
package's HIP library, the test oracle and the synthetic workload generator.
"""
from dataclasses import dataclass, field

import numpy as np


def ncoset(l):
    """Number of Cartesian orbitals with angular momentum <= l."""
    return 0 if l < 0 else (l + 1) * (l + 2) * (l + 3) // 6


@dataclass
class BasisSpec:
    lmin: np.ndarray        # [nset] int32
    lmax: np.ndarray        # [nset] int32
    npgf: np.ndarray        # [nset] int32
    nsgf_set: np.ndarray    # [nset] int32
    first_sgf: np.ndarray   # [nset] int32, one based
    sphi: np.ndarray        # [nsgf, maxco] float64
    zet: np.ndarray         # [nset, maxpgf] float64

    @property
    def nset(self):
        return len(self.lmax)

    @property
    def nsgf(self):
        return int(self.sphi.shape[0])

    @property
    def maxco(self):
        return int(self.sphi.shape[1])

    @property
    def maxpgf(self):
        return int(self.zet.shape[1])

    def ncoa(self, iset):
        return int(self.npgf[iset]) * ncoset(int(self.lmax[iset]))


def _i32(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


@dataclass
class TaskListArgs:
    orthorhombic: bool
    basis_sets: list                     # [nkinds] BasisSpec
    atom_positions: np.ndarray           # [natoms, 3]
    atom_kinds: np.ndarray               # [natoms] one based
    block_offsets: np.ndarray            # [nblocks] zero based, in doubles
    buffer_size: int                     # length of pab/hab buffers (doubles)
    level_list: np.ndarray               # per task, one based
    iatom_list: np.ndarray
    jatom_list: np.ndarray
    iset_list: np.ndarray
    jset_list: np.ndarray
    ipgf_list: np.ndarray
    jpgf_list: np.ndarray
    border_mask_list: np.ndarray
    block_num_list: np.ndarray
    radius_list: np.ndarray
    rab_list: np.ndarray                 # [ntasks, 3]
    npts_global: np.ndarray              # [nlevels, 3]
    npts_local: np.ndarray
    shift_local: np.ndarray
    border_width: np.ndarray
    dh: np.ndarray                       # [nlevels, 3, 3]
    dh_inv: np.ndarray
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        for name in ("atom_kinds", "block_offsets", "level_list", "iatom_list",
                     "jatom_list", "iset_list", "jset_list", "ipgf_list",
                     "jpgf_list", "border_mask_list", "block_num_list",
                     "npts_global", "npts_local", "shift_local", "border_width"):
            setattr(self, name, _i32(getattr(self, name)))
        for name in ("atom_positions", "radius_list", "rab_list", "dh", "dh_inv"):
            setattr(self, name, _f64(getattr(self, name)))

    @property
    def ntasks(self):
        return int(len(self.level_list))

    @property
    def nlevels(self):
        return int(len(self.npts_local))

    @property
    def natoms(self):
        return int(len(self.atom_kinds))

    @property
    def nkinds(self):
        return len(self.basis_sets)

    @property
    def nblocks(self):
        return int(len(self.block_offsets))

    def grid_shapes(self):
        """[nz, ny, nx] of every level's local grid."""
        return [(int(n[2]), int(n[1]), int(n[0])) for n in self.npts_local]

    def block_shape(self, iatom, jatom):
        """(rows, cols) in Fortran order of the block of atoms i, j (one based):
        rows belong to the lower-numbered atom."""
        a, b = (iatom, jatom) if iatom <= jatom else (jatom, iatom)
        na = self.basis_sets[self.atom_kinds[a - 1] - 1].nsgf
        nb = self.basis_sets[self.atom_kinds[b - 1] - 1].nsgf
        return na, nb


@dataclass
class NeighbourList:
    """Inputs of task-list generation (src/task_list_methods.F:117-501 for
    replicated grids): the orbital neighbour list, the basis sets with their
    screening radii, the grid levels' cutoffs and the grid geometry the
    resulting task list is created on."""
    orthorhombic: bool
    basis_sets: list                     # [nkinds] BasisSpec
    atom_positions: np.ndarray           # [natoms, 3] pbc(r)
    atom_kinds: np.ndarray               # [natoms] one based
    pair_atoms: np.ndarray               # [npairs, 2] iatom, jatom, one based
    pair_rab: np.ndarray                 # [npairs, 3]
    pair_image: object                   # [npairs] one based, or None
    set_radius: np.ndarray               # [nkinds, maxset]
    pgf_radius: np.ndarray               # [nkinds, maxset, maxpgf]
    kind_radius: object                  # [nkinds] or None (largest set radius)
    cutoff: np.ndarray                   # [nlevels]
    rel_cutoff: float
    eps_rho: float
    npts_global: np.ndarray              # [nlevels, 3]
    dh: np.ndarray                       # [nlevels, 3, 3]
    dh_inv: np.ndarray

    def __post_init__(self):
        for name in ("atom_kinds", "pair_atoms", "npts_global"):
            setattr(self, name, _i32(getattr(self, name)))
        for name in ("atom_positions", "pair_rab", "set_radius", "pgf_radius", "cutoff",
                     "dh", "dh_inv"):
            setattr(self, name, _f64(getattr(self, name)))
        if self.pair_image is not None:
            self.pair_image = _i32(self.pair_image)
        if self.kind_radius is not None:
            self.kind_radius = _f64(self.kind_radius)

    @property
    def npairs(self):
        return int(len(self.pair_atoms))

    @property
    def maxset(self):
        return int(self.set_radius.shape[1])

    @property
    def maxpgf(self):
        return int(self.pgf_radius.shape[2])

    def task_list_args(self, gen):
        """TaskListArgs of a generated list (gen: dict of the generator's
        output arrays), on this list's grid geometry (not distributed)."""
        nb = int(gen["nblocks"])
        ia, ja, blk = gen["iatom_list"], gen["jatom_list"], gen["block_num_list"]
        bi = np.zeros(nb, np.int64)
        bj = np.zeros(nb, np.int64)
        bi[blk - 1] = np.minimum(ia, ja)
        bj[blk - 1] = np.maximum(ia, ja)
        nlev = len(self.npts_global)
        return TaskListArgs(
            orthorhombic=self.orthorhombic, basis_sets=self.basis_sets,
            atom_positions=self.atom_positions, atom_kinds=self.atom_kinds,
            block_offsets=gen["block_offsets"], buffer_size=int(gen["buffer_size"]),
            level_list=gen["level_list"], iatom_list=ia, jatom_list=ja,
            iset_list=gen["iset_list"], jset_list=gen["jset_list"],
            ipgf_list=gen["ipgf_list"], jpgf_list=gen["jpgf_list"],
            border_mask_list=gen["border_mask_list"], block_num_list=blk,
            radius_list=gen["radius_list"], rab_list=gen["rab_list"],
            npts_global=self.npts_global, npts_local=self.npts_global,
            shift_local=np.zeros((nlev, 3)), border_width=np.zeros((nlev, 3)),
            dh=self.dh, dh_inv=self.dh_inv,
            meta=dict(nblocks=nb, block_i=bi, block_j=bj, generated=True))


def call_generate(fn, nl, basis_array):
    """Calls a task-list generator with the C signature of
    grid_mi355x_generate_task_list (fn: a ctypes function; basis_array: the
    kinds' basis-set structs as a ctypes array of pointers) and returns its
    outputs as a dict of numpy arrays.  Sized by a first call with capacity
    0, which returns the number of tasks."""
    import ctypes as C

    def p(a, ct):
        return None if a is None else a.ctypes.data_as(C.POINTER(ct))

    fn.restype = C.c_longlong
    nblocks = C.c_int(0)
    buffer_size = C.c_longlong(0)

    def run(cap):
        out = {k: np.zeros(max(cap, 1), np.int32) for k in
               ("level_list", "iatom_list", "jatom_list", "iset_list", "jset_list",
                "ipgf_list", "jpgf_list", "border_mask_list", "block_num_list")}
        out["radius_list"] = np.zeros(max(cap, 1))
        out["rab_list"] = np.zeros((max(cap, 1), 3))
        out["block_offsets"] = np.zeros(max(nl.npairs, 1), np.int32)
        n = fn(C.c_int(nl.npairs), p(nl.pair_atoms, C.c_int), p(nl.pair_rab, C.c_double),
               p(nl.pair_image, C.c_int), C.c_int(len(nl.atom_kinds)),
               p(nl.atom_kinds, C.c_int), p(nl.atom_positions, C.c_double),
               C.c_int(len(nl.basis_sets)), basis_array, C.c_int(nl.maxset),
               C.c_int(nl.maxpgf), p(nl.set_radius, C.c_double), p(nl.pgf_radius, C.c_double),
               p(nl.kind_radius, C.c_double), C.c_int(len(nl.cutoff)),
               p(nl.cutoff, C.c_double), C.c_double(nl.rel_cutoff), C.c_double(nl.eps_rho),
               C.c_longlong(cap),
               *[p(out[k], C.c_int) for k in ("level_list", "iatom_list", "jatom_list",
                                               "iset_list", "jset_list", "ipgf_list",
                                               "jpgf_list", "border_mask_list",
                                               "block_num_list")],
               p(out["radius_list"], C.c_double), p(out["rab_list"], C.c_double),
               C.byref(nblocks), p(out["block_offsets"], C.c_int), C.byref(buffer_size))
        return int(n), out

    n, _ = run(0)
    n2, out = run(n)
    assert n2 == n, (n, n2)
    for k in list(out):
        if k == "block_offsets":
            out[k] = out[k][:nblocks.value]
        else:
            out[k] = out[k][:n]
    out["ntasks"] = n
    out["nblocks"] = nblocks.value
    out["buffer_size"] = buffer_size.value
    return out
