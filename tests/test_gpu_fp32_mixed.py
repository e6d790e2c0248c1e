"""FP32-mixed mode of the HIP backend against the FP64 oracle.

Bars: rho L-inf (max|d| / max|ref|) <= 2e-6 on every level, forces
max|d| / max|ref| <= 1e-6.  hab and virial are held to the same 2e-6 / 1e-6.
Each case also checks that the FP32 kernels ran, two ways: the library's
launch counters show FP32 point-kernel launches and no FP64 launch of the
kinds that have an FP32 form, and the FP32-mixed result differs from the
FP64 one by more than 1e-10 relative (FP32 rounding is ~1e-8; two FP64 runs
differ by summation order only, ~1e-15).  FP64-only kernels: the per-task
kernels, and the orthorhombic integrate with lp 7..11 (matrix cores).
"""
import pytest

import replay
from cp2k_amd import grid, precision, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HAB_BAR = 2e-6
VIRIAL_BAR = 1e-6

# the synthetic lists of the FP64 parity tests: C1 in full, C2, C3
# (triclinic), C5 (l <= 4) with a subset of the atoms, and two wrap-stress lists
CONFIGS = [("ref32", 1, {}), ("h2o64", 2, dict(keep=6)), ("tri1024", 3, dict(keep=24)),
           ("hl2048", 4, dict(keep=16)), ("tri10", 5, {}), ("hl24", 6, {})]


@pytest.fixture(scope="module", autouse=True)
def fp32_mode():
    from cp2k_amd import build
    if not build.LIB.exists():
        build.build()
    yield
    grid.set_precision(grid.FP64)


@pytest.fixture(scope="module", params=CONFIGS, ids=lambda c: c[0])
def built(request):
    name, seed, kw = request.param
    print("fp32-mixed config", name, kw, flush=True)
    args, pab, V = synth.config(name, seed=seed, flip=0.4, **kw)
    return name, args, pab, V, grid.GridTaskList(args), O.TaskList(args)


# point-kernel ids (grid.KERNEL_*) whose every launch has an FP32 form
FP32_ONLY_COLLOCATE = (grid.KERNEL_TILES, grid.KERNEL_DIRECT, grid.KERNEL_GENERAL_POINTS)
FP32_ONLY_INTEGRATE = (grid.KERNEL_GENERAL_POINTS,)
POINT_IDS = (grid.KERNEL_TILES, grid.KERNEL_DIRECT, grid.KERNEL_INTEGRATE_POINTS,
             grid.KERNEL_PER_TASK, grid.KERNEL_GENERAL_POINTS)
NOISE = 1e-10


def run(gt, mode, fn):
    """fn(gt) in precision mode; returns (result, {id: (fp64, fp32) launches})."""
    grid.set_precision(mode)
    grid.reset_timing()
    try:
        out = fn(gt)
        return out, {k: grid.point_launches(k) for k in POINT_IDS}
    finally:
        grid.set_precision(grid.FP64)


def check_launches(name, launches, fp32_only):
    assert sum(v[1] for v in launches.values()) > 0, (name, launches)
    for k in fp32_only:
        assert launches[k][0] == 0, (name, grid.KERNEL_NAMES[k], launches)


def test_collocate_fp32_mixed(built):
    name, args, pab, V, gt, ot = built
    got, launches = run(gt, grid.FP32_MIXED, lambda g: g.collocate(100, pab))
    want = ot.collocate(100, pab)
    errs = precision.rho_errors(got, want)
    print(name, "rho", errs, launches, flush=True)
    for e in errs:
        assert e["linf"] <= precision.RHO_LINF_BAR, (name, errs)
    check_launches(name, launches, FP32_ONLY_COLLOCATE)
    fp64 = gt.collocate(100, pab)
    for a, b in zip(fp64, want):
        assert replay.rel_err(a, b) < 1e-11, name
    assert max(precision.linf_rel(a, b) for a, b in zip(got, fp64)) > NOISE, name


@pytest.mark.parametrize("tau", [False, True])
def test_integrate_fp32_mixed(built, tau):
    name, args, pab, V, gt, ot = built
    (h1, f1, v1), launches = run(
        gt, grid.FP32_MIXED,
        lambda g: g.integrate(V, pab, compute_tau=tau, forces=True, virial=True))
    h2, f2, v2 = ot.integrate(V, pab, compute_tau=tau, forces=True, virial=True)
    r = precision.report(hab=(h1, h2), forces=(f1, f2), virial=(v1, v2))
    print(name, tau, r, launches, flush=True)
    assert r["hab_linf"] <= HAB_BAR, (name, r)
    assert r["forces_linf"] <= precision.FORCE_BAR, (name, r)
    assert r["virial_linf"] <= VIRIAL_BAR, (name, r)
    general = gt.task_counts()[1] > 0
    if general or not tau:
        # orthorhombic lists with tau, forces and virial (l shift 5) hold no
        # task with la_max + lb_max + 5 <= 6: all of it takes the FP64
        # matrix-core kernel there
        check_launches(name, launches, FP32_ONLY_INTEGRATE)
        h3, _, _ = gt.integrate(V, pab, compute_tau=tau, forces=True, virial=True)
        assert precision.linf_rel(h1, h3) > NOISE, name


def test_precision_switch_round_trip():
    args, pab, V = synth.config("h2o2", seed=11, n=16)
    gt = grid.GridTaskList(args)
    a = gt.collocate(100, pab)
    grid.set_precision(grid.FP32_MIXED)
    assert grid.get_precision() == grid.FP32_MIXED
    grid.set_precision(grid.FP64)
    b = gt.collocate(100, pab)
    for x, y in zip(a, b):
        assert replay.rel_err(y, x) < 1e-12
