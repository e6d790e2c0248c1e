"""Matrix blocks <-> buffers on the device and buffer residency
(include/grid_buffers.h) against the oracle: packing HBM-resident blocks
makes the buffer's device copy current, the grid calls then read it with no
host copy, resident result buffers stay in HBM until synchronised, and
unpacking adds the hab slices into the blocks.  FP64 bar as everywhere:
|d|/max(1,|ref|) <= 1e-12 (src/grid/grid_task_list.c:241); the copies
themselves are exact."""
import numpy as np
import pytest

import replay
from cp2k_amd import grid, synth
from oracle import buffers as B
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


def _blocks(args, pab):
    offs = np.asarray(args.block_offsets)
    sizes = np.diff(np.append(offs, args.buffer_size))
    return offs, [pab[o:o + n].copy() for o, n in zip(offs, sizes)]


@pytest.mark.parametrize("name", ["h2o2", "tri4"])
def test_pack_collocate_resident(name):
    args, pab, V = synth.config(name, seed=9, n=24) if name == "h2o2" else synth.config(name, seed=9)
    print("pack/collocate", name, args.ntasks, flush=True)
    offs, blocks = _blocks(args, pab)
    perm = np.random.default_rng(1).permutation(len(blocks))
    dev = [grid.DeviceArray.from_numpy(blocks[i]) for i in perm]
    gt = grid.GridTaskList(args)
    abi = grid.AbiBuffers(gt)
    abi.pab.host()[:] = np.nan  # pinned host memory starts undefined: mark it
    grid.pack_buffer(dev, offs[perm], abi.pab)
    assert abi.pab.device_current()
    assert np.isnan(abi.pab.host()).all()  # packed on the device only
    for r in abi.rho:
        r.host()[:] = np.nan
        r.set_resident(True)
    abi.collocate(100)
    want = O.TaskList(args).collocate(100, B.pack_buffer(blocks, offs, args.buffer_size))
    # tri4's coarsest level is 8^3 points: thousands of periodic images add
    # into each point, so FP64 summation-order noise reaches ~1.5e-12 there
    # (the wrap-stress bar of the parity tests)
    tol = 1e-11 if name == "tri4" else TOL
    for r, w in zip(abi.rho, want):
        assert r.device_current()
        assert np.isnan(r.host()).all()  # results stayed in HBM
        r.sync_host()
        assert replay.rel_err(r.host().reshape(w.shape), w) < tol
    # the same blocks packed from host memory give the same buffer
    hostbuf = grid.OffloadBuf(args.buffer_size)
    grid.pack_buffer(blocks, offs, hostbuf)
    assert not hostbuf.device_current()
    assert np.array_equal(hostbuf.host(), B.pack_buffer(blocks, offs, args.buffer_size))
    abi.free()


@pytest.mark.parametrize("forces", [False, True])
def test_integrate_resident_unpack(forces):
    args, pab, V = synth.config("h2o2", seed=12, n=24)
    print("integrate/unpack", forces, flush=True)
    offs, blocks = _blocks(args, pab)
    gt = grid.GridTaskList(args)
    abi = grid.AbiBuffers(gt)
    abi.set_inputs(pab, V)
    abi.pab.host_written()
    for b in abi.pot:
        b.host_written()
    abi.hab.host()[:] = np.nan
    abi.hab.set_resident(True)
    abi.integrate(forces=forces)
    assert abi.hab.device_current()
    assert np.isnan(abi.hab.host()).all()
    ho, fo, _ = O.TaskList(args).integrate(V, pab if forces else None, forces=forces)
    rng = np.random.default_rng(5)
    init = [rng.normal(size=b.size) for b in blocks]
    dev = [grid.DeviceArray.from_numpy(x) for x in init]
    grid.unpack_buffer(dev, offs, abi.hab)
    want = B.unpack_buffer([x.copy() for x in init], offs, ho)
    for d, w in zip(dev, want):
        assert replay.rel_err(d.numpy(), w) < TOL
    if forces:
        assert replay.rel_err(abi.forces, fo) < 1e-10
    # host blocks: unpack brings the host copy up to date first
    hb = [x.copy() for x in init]
    grid.unpack_buffer(hb, offs, abi.hab)
    for a, w in zip(hb, want):
        assert replay.rel_err(a, w) < TOL
    abi.free()


def test_resident_roundtrip_collocate_integrate():
    """Density packed on the device -> collocate into resident grids ->
    integrate those grids (no host copy in between) -> hab resident ->
    unpack: each call equals the oracle on the same inputs."""
    args, pab, V = synth.config("h2o2", seed=3, n=16)
    print("resident round trip", args.ntasks, flush=True)
    offs, blocks = _blocks(args, pab)
    gt = grid.GridTaskList(args)
    abi = grid.AbiBuffers(gt)
    grid.pack_buffer([grid.DeviceArray.from_numpy(b) for b in blocks], offs, abi.pab)
    for r in abi.rho:
        r.set_resident(True)
    abi.collocate(100)
    abi.pot, saved = abi.rho, abi.pot  # integrate the density just collocated
    import ctypes as C
    abi._pot = (C.POINTER(grid.OffloadBuffer) * args.nlevels)(*[g.h for g in abi.pot])
    abi.hab.set_resident(True)
    abi.integrate()
    # each call against the oracle on the same inputs: the density, then the
    # integration of the GPU's own density
    ot = O.TaskList(args)
    rho_gpu = []
    for r, w in zip(abi.rho, ot.collocate(100, pab)):
        r.sync_host()
        rho_gpu.append(r.host().reshape(w.shape).copy())
        assert replay.rel_err(rho_gpu[-1], w) < TOL
    ho, _, _ = ot.integrate(rho_gpu)
    dev = [grid.DeviceArray.from_numpy(np.zeros(b.size)) for b in blocks]
    grid.unpack_buffer(dev, offs, abi.hab)
    got = B.pack_buffer([d.numpy() for d in dev], offs, args.buffer_size)
    # the potential is the density itself (values up to max|rho|, not O(1)
    # as in the other tests): summation-order noise scales with it
    scale = max(1.0, max(float(np.max(np.abs(r))) for r in rho_gpu))
    assert replay.rel_err(got, ho) < TOL * scale
    abi.pot = saved
    abi.free()
