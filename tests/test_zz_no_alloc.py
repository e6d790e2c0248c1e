"""grid_mi355x_collocate_task_list_noalloc against
grid_mi355x_collocate_task_list_dev (FP64: summation order only, 1e-12 with
the reference's metric), two calls in a row on lists of different size (the
second smaller), with the oracle as a third check on the second list."""
import pytest

import replay
from cp2k_amd import grid, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module", autouse=True)
def lib():
    from cp2k_amd import build
    if not build.LIB.exists():
        build.build()
    grid.set_precision(grid.FP64)
    return grid.lib()


def dev_and_noalloc(args, pab, func):
    gt = grid.GridTaskList(args)
    pab_d = grid.DeviceArray.from_numpy(pab)
    want_d = [grid.DeviceArray(s) for s in args.grid_shapes()]
    got_d = [grid.DeviceArray(s) for s in args.grid_shapes()]
    gt.collocate_dev(func, pab_d, want_d)
    grid.synchronize()
    want = [g.numpy() for g in want_d]
    gt.collocate_noalloc(func, pab_d, got_d)
    grid.synchronize()
    return [g.numpy() for g in got_d], want


@pytest.mark.parametrize("func", [100, 801])
def test_noalloc_matches_dev_two_calls(func):
    big = synth.config("h2o2", seed=21, n=32)       # orthorhombic, 4 levels
    small = synth.config("tri4", seed=22)           # triclinic, fewer tasks
    assert small[0].ntasks < big[0].ntasks
    for tag, (args, pab, _) in (("first", big), ("second", small)):
        print("noalloc", tag, func, args.ntasks, flush=True)
        got, want = dev_and_noalloc(args, pab, func)
        for g, w in zip(got, want):
            assert replay.rel_err(g, w) < TOL, (tag, func)
    args, pab, _ = small
    for g, w in zip(got, O.TaskList(args).collocate(func, pab)):
        assert replay.rel_err(g, w) < TOL, ("oracle", func)

