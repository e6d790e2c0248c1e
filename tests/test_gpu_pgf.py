"""Single-task entry points (grid_mi355x_collocate_pgf_product /
grid_mi355x_integrate_pgf_product, the drop-ins for the functions that
src/grid/grid_api.F:110-490 binds) against the reference's known-answer
files and the oracle.  FP64 bar: |d|/max(1,|ref|) <= 1e-12 on grids and hab
(src/grid/grid_task_list.c:241), forces / virials / hab derivatives 1e-10
against the oracle (the known-answer files: the replay's 1e-4 derivative
weight, src/grid/grid_replay.c:445)."""
import numpy as np
import pytest

import brute
import replay
import taskfile
from cp2k_amd import grid
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12
FTOL = 1e-10
DERIV_WEIGHT = 1e-4
FIXTURES = taskfile.all_fixtures()


@pytest.fixture(scope="module", autouse=True)
def lib():
    from cp2k_amd import build
    if not build.LIB.exists():
        build.build()
    grid.set_precision(grid.FP64)
    return grid.lib()


@pytest.fixture(scope="module", params=FIXTURES, ids=lambda p: p.name.split(".")[0])
def fixture(request):
    return taskfile.load(request.param)


def test_pgf_fixture(fixture):
    t = fixture
    print("pgf fixture", t.name, flush=True)
    g = grid.collocate_pgf(t)
    assert replay.rel_err(g, t.grid) < TOL
    hab, f, v = grid.integrate_pgf(t, compute_tau=t.compute_tau)
    assert replay.rel_err(hab, t.hab) < TOL
    assert replay.rel_err(f, t.forces) * DERIV_WEIGHT < TOL
    assert replay.rel_err(v[0] + v[1], t.virial) * DERIV_WEIGHT < TOL
    ho, fo, vo = O.integrate_pgf(t, compute_tau=t.compute_tau)
    assert replay.rel_err(hab, ho) < TOL
    assert replay.rel_err(f, fo) < 1e-8   # the l=3/3 files: cancellation noise
    assert replay.rel_err(v, vo) < 1e-8


def _task(rng, func, ortho, border, la, lb):
    t = brute.SyntheticTask(rng, la, lb, ortho=ortho, func=func, big_radius=False,
                            zrange=(0.2, 20.0))
    if border:
        t.border_mask = int(rng.integers(1, 64))
        t.border_width = rng.integers(1, 4, 3).astype(np.int32)
    return t


@pytest.mark.parametrize("ortho,border", [(True, False), (False, False), (True, True)])
def test_pgf_collocate_all_funcs(ortho, border):
    """Every grid_func on random l ranges, geometry and border masks; the
    grid is accumulated into (a non-zero starting grid)."""
    rng = np.random.default_rng(7 + 2 * ortho + border)
    print("pgf collocate funcs", ortho, border, flush=True)
    for func in brute.FUNCS:
        t = _task(rng, func, ortho, border, (int(rng.integers(0, 2)), int(rng.integers(1, 4))),
                  (int(rng.integers(0, 2)), int(rng.integers(1, 3))))
        g0 = rng.normal(size=tuple(int(x) for x in t.npts_local[::-1]))
        a = grid.collocate_pgf(t, grid=g0.copy())
        b = O.collocate_pgf(t, grid=g0.copy())
        assert replay.rel_err(a, b) < TOL, func


@pytest.mark.parametrize("ortho,border", [(True, False), (False, False), (True, True)])
@pytest.mark.parametrize("tau", [False, True])
def test_pgf_integrate(ortho, border, tau):
    rng = np.random.default_rng(31 + 4 * ortho + 2 * border + tau)
    print("pgf integrate", ortho, border, tau, flush=True)
    for rep in range(3):
        t = _task(rng, 100, ortho, border, (rep % 2, 1 + rep), (0, 2))
        V = rng.normal(size=tuple(int(x) for x in t.npts_local[::-1]))
        h1, f1, v1 = grid.integrate_pgf(t, grid=V, compute_tau=tau)
        h2, f2, v2 = O.integrate_pgf(t, grid=V, compute_tau=tau)
        assert replay.rel_err(h1, h2) < TOL
        assert replay.rel_err(f1, f2) < FTOL
        assert replay.rel_err(v1, v2) < FTOL


@pytest.mark.parametrize("ortho", [True, False])
def test_pgf_hab_derivatives(ortho):
    """hdab, hadb, a_hdab with and without forces / virials requested."""
    rng = np.random.default_rng(51 + ortho)
    print("pgf derivatives", ortho, flush=True)
    for rep in range(4):
        t = _task(rng, 100, ortho, False, (rep % 2, 1 + rep % 3), (0, 2))
        V = rng.normal(size=tuple(int(x) for x in t.npts_local[::-1]))
        with_f = rep % 2 == 1
        a = grid.integrate_pgf(t, grid=V, forces=with_f, virial=with_f, derivs=True)
        b = O.integrate_pgf(t, grid=V, forces=with_f, virial=with_f, derivs=True)
        assert replay.rel_err(a[0], b[0]) < TOL
        if with_f:
            assert replay.rel_err(a[1], b[1]) < FTOL
            assert replay.rel_err(a[2], b[2]) < FTOL
        for x, y in zip(a[3], b[3]):
            assert replay.rel_err(x, y) < FTOL


def test_pgf_device_buffers_accumulate():
    """Grid, potential and hab in HBM: two collocations add up to twice one,
    and an integrate into a non-zero device hab adds to it."""
    rng = np.random.default_rng(77)
    t = _task(rng, 100, True, False, (0, 2), (1, 2))
    shape = tuple(int(x) for x in t.npts_local[::-1])
    gd = grid.DeviceArray(shape)
    gd.copy_from(grid.DeviceArray.from_numpy(np.zeros(shape)))
    grid.collocate_pgf(t, grid=gd)
    grid.collocate_pgf(t, grid=gd)
    one = O.collocate_pgf(t)
    assert replay.rel_err(gd.numpy(), 2.0 * one) < TOL
    V = rng.normal(size=shape)
    h0 = rng.normal(size=(t.n2, t.n1))
    hd = grid.DeviceArray.from_numpy(h0)
    grid.integrate_pgf(t, grid=grid.DeviceArray.from_numpy(V), forces=False, virial=False,
                       hab=hd)
    ho = O.integrate_pgf(t, grid=V, forces=False, virial=False)[0]
    assert replay.rel_err(hd.numpy(), h0 + ho) < TOL


def test_pgf_skip_and_errors():
    """A task whose radius maps to no point leaves everything unchanged."""
    rng = np.random.default_rng(3)
    t = _task(rng, 100, True, False, (0, 1), (0, 1))
    t.radius = 0.4 * float(np.min(np.diag(t.dh)))
    g0 = rng.normal(size=tuple(int(x) for x in t.npts_local[::-1]))
    assert np.array_equal(grid.collocate_pgf(t, grid=g0.copy()), g0)
    h, f, v = grid.integrate_pgf(t, grid=g0)
    assert not h.any() and not f.any() and not v.any()
