"""Full-size properties of the 1024-atom triclinic workload (the headline
list: 2,199,464 general-cell tasks on 200/112/64/36 grids), FP64, where the
oracle would take minutes per call:

- the collocate / integrate adjoint identity
    sum_g rho[P](g) V(g) = sum_blocks w_b <P_b, H_b[V]>   (w: 1 on-site, 2 off)
  with the rounding bound 1e-11 of sum |rho V|;
- the forces of the integrate call against central differences of the
  energy E(R) = sum_b w_b <P_b, H_b[V](R)> along a random direction, for two
  atoms, moving the atom (its position and the rab of its tasks) and
  keeping the list's tasks and radii (eps_rho makes the truncation
  invisible at this step size).
"""
import copy

import numpy as np
import pytest

from cp2k_amd import grid, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def lib():
    from cp2k_amd import build
    if not build.LIB.exists():
        build.build()
    grid.set_precision(grid.FP64)
    return grid.lib()


def block_weights(args):
    w = np.zeros(args.buffer_size)
    bi, bj = args.meta["block_i"], args.meta["block_j"]
    ends = list(args.block_offsets[1:]) + [args.buffer_size]
    for b in range(args.nblocks):
        w[args.block_offsets[b]:ends[b]] = 1.0 if bi[b] == bj[b] else 2.0
    return w


def moved(args, a, step):
    """The list with atom a (zero based) moved by the vector `step`."""
    A = copy.copy(args)
    A.atom_positions = args.atom_positions.copy()
    A.atom_positions[a] += step
    rab = args.rab_list.copy()
    rab[args.iatom_list == a + 1] -= step
    rab[args.jatom_list == a + 1] += step
    A.rab_list = np.ascontiguousarray(rab)
    return A


@pytest.mark.slow
def test_tri1024_adjoint_identity_and_forces():
    args, pab, V = synth.config("tri1024", seed=0)
    print("tri1024 full size:", args.ntasks, "tasks", flush=True)
    w = block_weights(args)
    gt = grid.GridTaskList(args)
    rho = gt.collocate(grid.GRID_FUNC_AB, pab)
    hab, f, _ = gt.integrate(V, pab, forces=True)
    del gt
    lhs = sum(float(np.sum(r * v)) for r, v in zip(rho, V))
    scale = sum(float(np.sum(np.abs(r * v))) for r, v in zip(rho, V))
    rhs = float(np.sum(w * pab * hab))
    print(f"adjoint: {lhs:.15e} vs {rhs:.15e} (scale {scale:.3e})", flush=True)
    assert abs(lhs - rhs) <= 1e-11 * scale
    del rho

    rng = np.random.default_rng(5)
    h = 1e-4
    fmax = float(np.max(np.linalg.norm(f, axis=1)))
    for a in (0, 517):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        e = []
        for sgn in (1.0, -1.0):
            gm = grid.GridTaskList(moved(args, a, sgn * h * u))
            hm, _, _ = gm.integrate(V)
            del gm
            e.append(float(np.sum(w * pab * hm)))
        fd = (e[0] - e[1]) / (2 * h)
        fu = float(f[a] @ u)
        print(f"atom {a}: F.u {fu:.10e}  finite difference {fd:.10e}", flush=True)
        assert abs(fu - fd) <= 1e-4 * max(float(np.linalg.norm(f[a])), 1e-3 * fmax)
