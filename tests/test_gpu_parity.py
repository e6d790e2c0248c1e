"""Parity of the HIP backend (through its C-ABI) with the CPU oracle and the
reference's known-answer files.  FP64 bar: the reference's own validation
metric |d|/max(1,|ref|) <= 1e-12 on grids and hab (src/grid/grid_task_list.c
:241,:367), forces/virial <= 1e-8 (:386,:410).  The synthetic lists hold
forces/virial to 1e-10; the known-answer fixtures keep the reference's 1e-8
because some (l=3/3) have forces that are pure cancellation noise (~1e-9).
"""
import numpy as np
import pytest

import replay
import taskfile
from cp2k_amd import grid, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12
FTOL = 1e-10
FTOL_REF = 1e-8  # src/grid/grid_task_list.c:386,410
DERIV_WEIGHT = 1e-4  # src/grid/grid_replay.c:445
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


def test_fixture_collocate(fixture):
    t = fixture
    print("fixture collocate", t.name, flush=True)
    args, pab = replay.task_list(t, collocate=True)
    g = grid.GridTaskList(args).collocate(t.func, pab)[0]
    assert replay.rel_err(g, t.grid) < TOL
    assert replay.rel_err(g, O.TaskList(args).collocate(t.func, pab)[0]) < TOL


def test_fixture_integrate(fixture):
    t = fixture
    print("fixture integrate", t.name, flush=True)
    args, pab = replay.task_list(t, collocate=False)
    hab, f, v = grid.GridTaskList(args).integrate([t.grid], pab, compute_tau=t.compute_tau,
                                                  forces=True, virial=True)
    assert replay.rel_err(hab.reshape(t.n2, t.n1), t.hab) < TOL
    assert replay.rel_err(f, t.forces) * DERIV_WEIGHT < TOL
    assert replay.rel_err(v, t.virial) * DERIV_WEIGHT < TOL
    ho, fo, vo = O.TaskList(args).integrate([t.grid], pab, compute_tau=t.compute_tau,
                                            forces=True, virial=True)
    assert replay.rel_err(hab, ho) < TOL
    assert replay.rel_err(f, fo) < FTOL_REF
    assert replay.rel_err(v, vo) < FTOL_REF


# Benchmark cells and grids with a subset of the atoms: C1 in full, and C2,
# C3 (triclinic, general kernel), C5 (l <= 4, five levels) with keep=N atoms.
# The wrap-stress lists put 10-24 atoms into cells of 9-16 bohr with 8^3-12^3
# coarse grids, where every point sums thousands of periodic images with
# cancellation; there the summation-order noise of FP64 itself reaches ~1e-12,
# so they are held to 1e-11.
CONFIGS = [("ref32", 1, {}, TOL), ("h2o64", 2, dict(keep=6), TOL),
           ("tri1024", 3, dict(keep=24), TOL), ("hl2048", 4, dict(keep=16), TOL),
           ("tri10", 5, {}, 1e-11), ("hl24", 6, {}, 1e-11)]


@pytest.fixture(scope="module", params=CONFIGS, ids=lambda c: c[0])
def built(request):
    name, seed, kw, tol = request.param
    print("config", name, kw, flush=True)
    args, pab, V = synth.config(name, seed=seed, flip=0.4, **kw)
    return name, tol, args, pab, V, grid.GridTaskList(args), O.TaskList(args)


def test_config_collocate(built):
    name, tol, args, pab, V, gt, ot = built
    for got, want in zip(gt.collocate(100, pab), ot.collocate(100, pab)):
        assert replay.rel_err(got, want) < tol, name


@pytest.mark.parametrize("tau", [False, True])
def test_config_integrate_forces_virial(built, tau):
    name, tol, args, pab, V, gt, ot = built
    h1, f1, v1 = gt.integrate(V, pab, compute_tau=tau, forces=True, virial=True)
    h2, f2, v2 = ot.integrate(V, pab, compute_tau=tau, forces=True, virial=True)
    assert replay.rel_err(h1, h2) < tol, name
    assert replay.rel_err(f1, f2) < max(FTOL, 100 * tol), name
    assert replay.rel_err(v1, v2) < max(FTOL, 100 * tol), name


def test_config_integrate_hab_only(built):
    name, tol, args, pab, V, gt, ot = built
    h1, f1, v1 = gt.integrate(V)
    h2, _, _ = ot.integrate(V)
    assert f1 is None and v1 is None
    assert replay.rel_err(h1, h2) < tol, name


# every grid_func of src/grid/common/grid_constants.h:14-48, on an
# orthorhombic list (wave-per-task and tile kernels, per-task kernels for
# lp > 8) and on a triclinic list (general kernels, per-task for lp > 6)
ALL_FUNCS = [100, 200, 301, 302, 303, 411, 412, 413, 421, 422, 423, 431, 432, 433,
             501, 502, 503, 601, 602, 603, 701, 702, 703, 801, 802, 803,
             901, 902, 903, 904, 905, 906, 1001, 1002, 1003]


@pytest.fixture(scope="module", params=["h2o1", "tri6"])
def func_list(request):
    kw = dict(n=24) if request.param == "h2o1" else {}
    args, pab, V = synth.config(request.param, seed=17, flip=0.5, **kw)
    return request.param, args, pab, grid.GridTaskList(args), O.TaskList(args)


@pytest.mark.parametrize("func", ALL_FUNCS)
def test_funcs(func_list, func):
    name, args, pab, gt, ot = func_list
    print("func", name, func, flush=True)
    got = gt.collocate(func, pab)
    want = ot.collocate(func, pab)
    for g, w in zip(got, want):
        assert replay.rel_err(g, w) < TOL, (name, func)


def test_device_api_matches_host_api():
    args, pab, V = synth.config("h2o2", seed=9, n=32)
    gt = grid.GridTaskList(args)
    host = gt.collocate(100, pab)
    pab_d = grid.DeviceArray.from_numpy(pab)
    grids_d = [grid.DeviceArray(s) for s in args.grid_shapes()]
    gt.collocate_dev(100, pab_d, grids_d)
    grid.synchronize()
    for h, d in zip(host, grids_d):
        assert replay.rel_err(d.numpy(), h) < 1e-14
    V_d = [grid.DeviceArray.from_numpy(v) for v in V]
    hab_d = grid.DeviceArray(args.buffer_size)
    f_d = grid.DeviceArray((args.natoms, 3))
    v_d = grid.DeviceArray((3, 3))
    gt.integrate_dev(V_d, hab_d, pab=pab_d, forces=f_d, virial=v_d)
    grid.synchronize()
    h, f, v = gt.integrate(V, pab, forces=True, virial=True)
    assert replay.rel_err(hab_d.numpy(), h) < 1e-14
    assert replay.rel_err(f_d.numpy(), f) < 1e-13
    assert replay.rel_err(v_d.numpy(), v) < 1e-13


def test_device_calls_in_flight():
    """Device-API calls return before their kernels run: back-to-back
    collocate/integrate pairs (more than the library's eight per-call level
    tables) without a synchronisation must each use their own buffers."""
    args, pab, V = synth.config("h2o2", seed=9, n=32)
    gt = grid.GridTaskList(args)
    want_rho = gt.collocate(100, pab)
    want_hab, _, _ = gt.integrate(V)
    pab_d = grid.DeviceArray.from_numpy(pab)
    rho_d = [grid.DeviceArray(s) for s in args.grid_shapes()]
    V_d = [grid.DeviceArray.from_numpy(v) for v in V]
    hab_d = grid.DeviceArray(args.buffer_size)
    for _ in range(12):
        gt.collocate_dev(100, pab_d, rho_d)
        gt.integrate_dev(V_d, hab_d)
    grid.synchronize()
    for v_d, v in zip(V_d, V):
        assert np.array_equal(v_d.numpy(), v)
    for r_d, r in zip(rho_d, want_rho):
        assert replay.rel_err(r_d.numpy(), r) < 1e-14
    assert replay.rel_err(hab_d.numpy(), want_hab) < 1e-14


def test_autotune_keeps_results():
    args, pab, V = synth.config("h2o64", seed=3, keep=4)
    gt = grid.GridTaskList(args)
    pab_d = grid.DeviceArray.from_numpy(pab)
    rho_d = [grid.DeviceArray(s) for s in args.grid_shapes()]
    gt.autotune(100, pab_d, rho_d)
    for level in range(args.nlevels):
        rounds, waves = gt.tuned_shape(level)
        assert rounds in (2, 4, 8) and waves in (6, 8, 12)
    for got, want in zip(gt.collocate(100, pab), O.TaskList(args).collocate(100, pab)):
        assert replay.rel_err(got, want) < TOL


def test_task_points_match_oracle():
    args, pab, V = synth.config("h2o2", seed=4, n=32)
    assert grid.GridTaskList(args).task_points() == O.TaskList(args).task_points()


def test_handle_reuse_rebuilds():
    import ctypes as C
    a1, p1, _ = synth.config("ref6", seed=1)
    a2, p2, _ = synth.config("ref8", seed=2)
    gt = grid.GridTaskList(a1)
    first = gt.collocate(100, p1)
    assert replay.rel_err(first[0], O.TaskList(a1).collocate(100, p1)[0]) < TOL
    # rebuild the same handle with a different list (grid_task_list.c:42-63)
    h = gt.h
    gt2 = grid.GridTaskList.__new__(grid.GridTaskList)
    gt2.args = a2
    gt2._bases = [grid.Basis(b) for b in a2.basis_sets]
    gt2._basis_arr = (C.POINTER(grid.GridBasisSet) * len(gt2._bases))(*[b.h for b in gt2._bases])
    gt2.h = h
    gt2._npts = np.ascontiguousarray(a2.npts_local, dtype=np.int32)
    a = a2
    P = grid._ptr
    grid.lib().grid_create_task_list(
        C.c_bool(a.orthorhombic), C.c_int(a.ntasks), C.c_int(a.nlevels), C.c_int(a.natoms),
        C.c_int(a.nkinds), C.c_int(a.nblocks), P(a.block_offsets, C.c_int),
        P(a.atom_positions), P(a.atom_kinds, C.c_int), gt2._basis_arr,
        P(a.level_list, C.c_int), P(a.iatom_list, C.c_int), P(a.jatom_list, C.c_int),
        P(a.iset_list, C.c_int), P(a.jset_list, C.c_int), P(a.ipgf_list, C.c_int),
        P(a.jpgf_list, C.c_int), P(a.border_mask_list, C.c_int),
        P(a.block_num_list, C.c_int), P(a.radius_list), P(a.rab_list),
        P(a.npts_global, C.c_int), P(a.npts_local, C.c_int), P(a.shift_local, C.c_int),
        P(a.border_width, C.c_int), P(a.dh), P(a.dh_inv), C.byref(gt2.h))
    assert gt2.h.value == h.value
    gt.h = None  # owned by gt2 now
    second = gt2.collocate(100, p2)
    assert replay.rel_err(second[0], O.TaskList(a2).collocate(100, p2)[0]) < TOL


@pytest.mark.slow
def test_h2o64_adjoint_identity():
    """Full-size C2 list: sum_g rho V == sum s_ij <P, H[V]> on the GPU."""
    args, pab, V = synth.config("h2o64", seed=0)
    gt = grid.GridTaskList(args)
    rho = gt.collocate(100, pab)
    hab, _, _ = gt.integrate(V)
    lhs = sum(float(np.sum(r * v)) for r, v in zip(rho, V))
    w = np.zeros(args.buffer_size)
    bi, bj = args.meta["block_i"], args.meta["block_j"]
    ends = list(args.block_offsets[1:]) + [args.buffer_size]
    for b in range(args.nblocks):
        w[args.block_offsets[b]:ends[b]] = 1.0 if bi[b] == bj[b] else 2.0
    rhs = float(np.sum(w * pab * hab))
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + 1.0)
