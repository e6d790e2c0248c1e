"""Turns a single-task known-answer fixture into a one-task task list, with
the conventions of the reference's batched replay (src/grid/grid_replay.c):

  - identity decontraction basis per atom, npgf = n/ncoset(lmax)   (:114-148)
  - 2 atoms, 2 kinds, 1 level, 1 block at offset 0, ipgf = o1/ncoset+1 (:154-214)
  - pab_blocks[j*n1+i] = 0.5 * f * pab[j][i], f = rscale for collocate,
    1 for integrate                                                   (:377-382)
  - integrate uses compute_tau = (func == DADB)                       (:252)
"""
import numpy as np

from cp2k_amd.tasklist import BasisSpec, TaskListArgs, ncoset


def dummy_basis(size, lmin, lmax, zet):
    npgf = size // ncoset(lmax)
    assert npgf * ncoset(lmax) == size
    return BasisSpec(lmin=np.array([lmin]), lmax=np.array([lmax]),
                     npgf=np.array([npgf]), nsgf_set=np.array([size]),
                     first_sgf=np.array([1]), sphi=np.eye(size),
                     zet=np.full((1, npgf), zet))


def task_list(t, collocate, cycles=1):
    a = dummy_basis(t.n1, t.la_min, t.la_max, t.zeta)
    b = dummy_basis(t.n2, t.lb_min, t.lb_max, t.zetb)
    ipgf = t.o1 // ncoset(t.la_max) + 1
    jpgf = t.o2 // ncoset(t.lb_max) + 1
    n = cycles
    args = TaskListArgs(
        orthorhombic=t.orthorhombic, basis_sets=[a, b],
        atom_positions=np.array([t.ra, t.ra + t.rab]), atom_kinds=[1, 2],
        block_offsets=[0], buffer_size=t.n1 * t.n2,
        level_list=np.ones(n), iatom_list=np.ones(n), jatom_list=np.full(n, 2),
        iset_list=np.ones(n), jset_list=np.ones(n), ipgf_list=np.full(n, ipgf),
        jpgf_list=np.full(n, jpgf), border_mask_list=np.full(n, t.border_mask),
        block_num_list=np.ones(n), radius_list=np.full(n, t.radius),
        rab_list=np.tile(t.rab, (n, 1)), npts_global=[t.npts_global],
        npts_local=[t.npts_local], shift_local=[t.shift_local],
        border_width=[t.border_width], dh=[t.dh], dh_inv=[t.dh_inv],
        meta=dict(block_i=np.array([1]), block_j=np.array([2])))
    f = t.rscale if collocate else 1.0
    pab_blocks = (0.5 * f * t.pab).ravel().copy()  # [j][i] row-major = [n2][n1]
    return args, pab_blocks


def rel_err(test, ref):
    """max |test-ref| / max(1, |ref|): the reference's own metric
    (src/grid/grid_task_list.c:250-254)."""
    test = np.asarray(test, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if ref.size == 0:
        return 0.0
    return float(np.max(np.abs(test - ref) / np.maximum(1.0, np.abs(ref))))


def block_weights(args):
    """Per element of the pab/hab buffer of a generated list: 1 for a block
    of one atom with itself, 2 for a block of two atoms (the off-site rscale
    of collocate, src/grid/ref/grid_ref_task_list.c:369; hab is unscaled)."""
    import numpy as np
    w = np.zeros(args.buffer_size)
    bi, bj = args.meta["block_i"], args.meta["block_j"]
    ends = list(args.block_offsets[1:]) + [args.buffer_size]
    for b in range(args.nblocks):
        w[int(args.block_offsets[b]):int(ends[b])] = 1.0 if bi[b] == bj[b] else 2.0
    return w
