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
