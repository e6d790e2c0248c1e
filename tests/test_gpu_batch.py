"""Batched task lists (grid_mi355x_create_batch / grid_mi355x_batch_run).

A batch of lists with different cells, grid sizes and level counts (cubic
one-level lists, a multi-level water box, a triclinic list, an l <= 4 list
on five levels) must give each
list exactly what the list gives on its own (summation order aside, 1e-12
as in the FP64 parity tests).  A run replays a captured hipGraph: every
replay is checked, with eager calls and synchronisations in between.
"""
import pytest

import replay
from cp2k_amd import grid, precision, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12
FTOL = 1e-10


@pytest.fixture(scope="module")
def lists():
    from cp2k_amd import build
    if not build.LIB.exists():
        build.build()
    grid.set_precision(grid.FP64)
    out = synth.batch_config(4, seed=5, atoms=(8, 20))
    out.append(synth.config("h2o2", seed=7, n=32))
    out.append(synth.config("tri10", seed=8))
    # la, lb <= 4: lp up to 8 in collocate and 11 with forces and virial
    out.append(synth.config("hl2048", seed=9, keep=6))
    return out


def tol_of(i, lists):
    # the wrap-stress triclinic list sums thousands of images per point with
    # cancellation: 1e-11, as in the FP64 parity tests
    return 1e-11 if lists[i][0].meta.get("config") == "tri10" else TOL


def eager(args, pab, V):
    gt = grid.GridTaskList(args)
    rho = gt.collocate(100, pab)
    hab, f, v = gt.integrate(V, pab, forces=True, virial=True)
    return rho, hab, f, v


def check(got, want, tag, tol=TOL):
    rho, hab, f, v = got
    rho0, hab0, f0, v0 = want
    for a, b in zip(rho, rho0):
        assert replay.rel_err(a, b) < tol, tag
    assert replay.rel_err(hab, hab0) < tol, tag
    assert replay.rel_err(f, f0) < FTOL, tag
    assert replay.rel_err(v, v0) < FTOL, tag


def test_batch_matches_single_lists(lists):
    want = [eager(a, p, V) for a, p, V in lists]
    gts = [grid.GridTaskList(a) for a, _, _ in lists]
    batch = grid.GridBatch(gts)
    for i, (a, p, V) in enumerate(lists):
        batch.set_inputs(i, pab=p, V=V)
    assert batch.task_points() == sum(g.task_points() for g in gts)
    for rep in range(3):
        print("batch replay", rep, flush=True)
        batch.run(3, forces=True)
        grid.synchronize()
        for i in range(len(lists)):
            check(batch.outputs(i), want[i], (rep, i), tol_of(i, lists))
        # eager work between replays, on the same stream
        a, p, V = lists[rep]
        check(eager(a, p, V), want[rep], ("eager", rep), tol_of(rep, lists))
        grid.synchronize()


def test_batch_first_list_against_oracle(lists):
    a, p, V = lists[0]
    batch = grid.GridBatch([grid.GridTaskList(x) for x, _, _ in lists[:2]])
    batch.set_inputs(0, pab=p, V=V)
    batch.set_inputs(1, pab=lists[1][1], V=lists[1][2])
    batch.run(1)
    batch.run(2, forces=True)
    grid.synchronize()
    rho, hab, f, v = batch.outputs(0)
    ot = O.TaskList(a)
    for g, w in zip(rho, ot.collocate(100, p)):
        assert replay.rel_err(g, w) < TOL
    h2, f2, v2 = ot.integrate(V, p, forces=True, virial=True)
    assert replay.rel_err(hab, h2) < TOL
    assert replay.rel_err(f, f2) < FTOL
    assert replay.rel_err(v, v2) < FTOL


def test_batch_fp32_mixed(lists):
    want = [eager(a, p, V) for a, p, V in lists]
    batch = grid.GridBatch([grid.GridTaskList(a) for a, _, _ in lists])
    for i, (a, p, V) in enumerate(lists):
        batch.set_inputs(i, pab=p, V=V)
    grid.set_precision(grid.FP32_MIXED)
    try:
        for rep in range(2):
            batch.run(3, forces=True)
            grid.synchronize()
            for i in range(len(lists)):
                rho, hab, f, v = batch.outputs(i)
                r = precision.report(rho=(rho, want[i][0]), hab=(hab, want[i][1]),
                                     forces=(f, want[i][2]))
                print("batch fp32", rep, i, {k: x for k, x in r.items() if k != "rho_levels"},
                      flush=True)
                assert r["rho_linf"] <= precision.RHO_LINF_BAR, (rep, i, r)
                assert r["forces_linf"] <= precision.FORCE_BAR, (rep, i, r)
    finally:
        grid.set_precision(grid.FP64)


def _small(lists):
    # two lists are enough: the cubic one-level list and the water box
    return [lists[0], lists[4]]


def test_batch_graphs_survive_buffer_growth(lists):
    """Runs whose l shift grows the per-task buffers (collocate, integrate
    with forces and virial, tau) alternate; every replay of an earlier graph
    must still match the eager calls."""
    sub = _small(lists)
    gts = [grid.GridTaskList(a) for a, _, _ in sub]
    batch = grid.GridBatch(gts)
    for i, (a, p, V) in enumerate(sub):
        batch.set_inputs(i, pab=p, V=V)
    want_rho = [g.collocate(100, p) for g, (_, p, _) in zip(gts, sub)]
    want_f = [g.integrate(V, p, forces=True, virial=True) for g, (_, p, V) in zip(gts, sub)]
    want_tau = [g.integrate(V, p, compute_tau=True) for g, (_, p, V) in zip(gts, sub)]
    seq = [(1, False, False), (2, False, True), (1, False, False), (3, True, False),
           (2, False, True), (1, False, False)]
    for step, (what, tau, forces) in enumerate(seq):
        print("batch sequence", step, what, tau, forces, flush=True)
        batch.run(what, compute_tau=tau, forces=forces)
        grid.synchronize()
        for i in range(len(sub)):
            rho, hab, f, v = batch.outputs(i)
            if what & 1:
                for x, y in zip(rho, want_rho[i]):
                    assert replay.rel_err(x, y) < TOL, (step, i)
            if what & 2:
                w = want_tau[i] if tau else want_f[i]
                assert replay.rel_err(hab, w[0]) < TOL, (step, i)
                if forces:
                    assert replay.rel_err(f, w[1]) < FTOL, (step, i)
                    assert replay.rel_err(v, w[2]) < FTOL, (step, i)


def test_batch_graphs_alternate_precision(lists):
    """FP32-mixed and FP64 integrate graphs of one batch hold their own level
    tables: replays alternate precision and each matches the eager call of
    its precision."""
    sub = _small(lists)
    gts = [grid.GridTaskList(a) for a, _, _ in sub]
    batch = grid.GridBatch(gts)
    for i, (a, p, V) in enumerate(sub):
        batch.set_inputs(i, pab=p, V=V)
    want = {}
    try:
        for prec in (grid.FP64, grid.FP32_MIXED):
            grid.set_precision(prec)
            want[prec] = [g.integrate(V, p, forces=True, virial=True)
                          for g, (_, p, V) in zip(gts, sub)]
        for step, prec in enumerate([grid.FP32_MIXED, grid.FP64, grid.FP32_MIXED, grid.FP64,
                                     grid.FP32_MIXED]):
            print("batch precision", step, prec, flush=True)
            grid.set_precision(prec)
            batch.run(2, forces=True)
            grid.synchronize()
            for i in range(len(sub)):
                _, hab, f, v = batch.outputs(i)
                w = want[prec][i]
                # eager and replay run the same kernels: FP64 summation order only
                assert replay.rel_err(hab, w[0]) < TOL, (step, i)
                assert replay.rel_err(f, w[1]) < FTOL, (step, i)
                assert replay.rel_err(v, w[2]) < FTOL, (step, i)
    finally:
        grid.set_precision(grid.FP64)


def test_batch_work_queue(lists):
    """The same batch with its point kernels as persistent launches over
    device work queues: every replay equals the lists alone; switching the
    launch shape back and forth (and the autotuner's choice) keeps results."""
    want = [eager(a, p, V) for a, p, V in lists]
    batch = grid.GridBatch([grid.GridTaskList(a) for a, _, _ in lists])
    for i, (a, p, V) in enumerate(lists):
        batch.set_inputs(i, pab=p, V=V)
    for step, queue in enumerate([True, True, False, True]):
        print("batch work queue", step, queue, flush=True)
        batch.set_work_queue(queue)
        assert batch.work_queue() == queue
        batch.run(3, forces=True)
        grid.synchronize()
        for i in range(len(lists)):
            check(batch.outputs(i), want[i], (step, queue, i), tol_of(i, lists))
    chosen = batch.autotune(3, forces=True)
    assert batch.work_queue() == chosen
    batch.run(3, forces=True)
    grid.synchronize()
    for i in range(len(lists)):
        check(batch.outputs(i), want[i], ("tuned", chosen, i), tol_of(i, lists))


def test_batch_32_lists_against_oracle():
    """32 lists at the batched workload's shapes (8-64 atoms, cubic 6-14 A
    cells, one level of 32^3-112^3 points, l <= 2) in one batch with the work
    queue, each list against the oracle."""
    lists = synth.batch_config(32, seed=11, atoms=(8, 64))
    batch = grid.GridBatch([grid.GridTaskList(a) for a, _, _ in lists])
    for i, (a, p, V) in enumerate(lists):
        batch.set_inputs(i, pab=p, V=V)
    batch.set_work_queue(True)
    batch.run(3, forces=True)
    grid.synchronize()
    for i, (a, p, V) in enumerate(lists):
        print("batch32 list", i, a.ntasks, flush=True)
        rho, hab, f, v = batch.outputs(i)
        ot = O.TaskList(a)
        for g, w in zip(rho, ot.collocate(100, p)):
            assert replay.rel_err(g, w) < TOL, i
        h2, f2, v2 = ot.integrate(V, p, forces=True, virial=True)
        assert replay.rel_err(hab, h2) < TOL, i
        assert replay.rel_err(f, f2) < FTOL, i
        assert replay.rel_err(v, v2) < FTOL, i


def test_batch_256_adjoint_identity():
    """The full 256-list batched workload: per list, the density collocated
    from P and the hab integrated from the potential V satisfy
    sum_g rho[P] V = sum_blocks s_ij <P, H[V]> (s = 2 off-diagonal), the
    adjoint identity of collocate and integrate."""
    import numpy as np
    lists = synth.batch_config(256, seed=0)
    batch = grid.GridBatch([grid.GridTaskList(a) for a, _, _ in lists])
    for i, (a, p, V) in enumerate(lists):
        batch.set_inputs(i, pab=p, V=V)
    batch.set_work_queue(True)
    batch.run(3)
    grid.synchronize()
    for i, (a, p, V) in enumerate(lists):
        rho, hab, _, _ = batch.outputs(i)
        lhs = sum(float(np.sum(r * v)) for r, v in zip(rho, V))
        rhs = float(np.dot(replay.block_weights(a) * p, hab))
        assert abs(lhs - rhs) <= 1e-10 * (abs(lhs) + 1.0), (i, lhs, rhs)
