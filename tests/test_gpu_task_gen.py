"""Device task-list generation (grid_mi355x_generate_task_list) against the
oracle restatement of generate_qs_task_list (oracle/task_gen.c, itself
pinned in test_task_gen.py).

Bar: every integer list, the rab list, the blocks and their offsets are
identical.  Radii come out of a bisection whose decisions compare single-
precision exponentials with eps_rho; the device's exp and the host's expf
may round a value on the other side of such a comparison, so a radius may
end in a neighbouring bracket: at most 1 task in 10^4 may differ, each by
less than the bracket width 1e-2, and every device radius must satisfy the
defining property of exp_radius_very_extended on its own (tests/gen_brute.py).
A list built from the generated arrays collocates and integrates like the
oracle on the same arguments.
"""
import copy

import numpy as np
import pytest

import gen_brute as GB
import replay
from cp2k_amd import grid, synth
from oracle import oracle as O
from test_task_gen import INT_KEYS, neighbour_list

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def lib():
    from cp2k_amd import build
    if not build.LIB.exists():
        build.build()
    grid.set_precision(grid.FP64)
    return grid.lib()


def compare(nl, sample=1500):
    gpu = grid.generate_task_list(nl)
    ora = O.generate_task_list(nl)
    assert gpu["ntasks"] == ora["ntasks"]
    for k in INT_KEYS + ("border_mask_list",):
        np.testing.assert_array_equal(gpu[k], ora[k], err_msg=k)
    np.testing.assert_array_equal(gpu["rab_list"], ora["rab_list"])
    assert gpu["nblocks"] == ora["nblocks"]
    np.testing.assert_array_equal(gpu["block_offsets"], ora["block_offsets"])
    assert gpu["buffer_size"] == ora["buffer_size"]
    d = np.abs(gpu["radius_list"] - ora["radius_list"])
    ndiff = int((d != 0).sum())
    assert ndiff <= max(1, gpu["ntasks"] // 10000), f"{ndiff} radii differ"
    assert d.max(initial=0.0) < 1.0e-2
    if sample:
        bad = GB.radius_bracket_violations(nl, gpu, sample=sample)
        assert not bad, f"device radii {bad[:5]} fail the exp_radius bracket property"
    return gpu, ora, ndiff


@pytest.mark.parametrize("name,kw", [
    ("ref16", {}), ("tri12", {}), ("h2o8", {}), ("hl12", {}),
    ("tri8", dict(images=True)), ("h2o4", dict(kind_radius=True)),
    ("ref24", dict(shuffle=True, images=True)),
])
def test_generate_vs_oracle(name, kw):
    print("generate", name, kw, flush=True)
    compare(neighbour_list(name, **kw))


def test_generate_empty_and_repeat():
    nl = neighbour_list("ref16")
    from cp2k_amd.tasklist import NeighbourList
    empty = NeighbourList(**{**nl.__dict__, "pair_atoms": np.zeros((0, 2)),
                             "pair_rab": np.zeros((0, 3))})
    out = grid.generate_task_list(empty)
    assert out["ntasks"] == 0 and out["nblocks"] == 0
    # a larger list, then a smaller one: the library's buffers are reused
    a = grid.generate_task_list(neighbour_list("tri12"))
    b = grid.generate_task_list(nl)
    c = O.generate_task_list(nl)
    assert a["ntasks"] > b["ntasks"]
    for k in INT_KEYS:
        np.testing.assert_array_equal(b[k], c[k], err_msg=k)


def order_noise(args, pab):
    """Per level, how far the oracle's own FP64 density moves when the tasks
    are summed in reverse order."""
    rev = copy.copy(args)
    perm = np.arange(args.ntasks)[::-1].copy()
    for k in ("level_list", "iatom_list", "jatom_list", "iset_list", "jset_list", "ipgf_list",
              "jpgf_list", "border_mask_list", "block_num_list", "radius_list", "rab_list"):
        setattr(rev, k, np.ascontiguousarray(getattr(args, k)[perm]))
    a = O.TaskList(args).collocate(grid.GRID_FUNC_AB, pab)
    b = O.TaskList(rev).collocate(grid.GRID_FUNC_AB, pab)
    return a, [replay.rel_err(x, y) for x, y in zip(a, b)]


@pytest.mark.parametrize("name,ortho,tol", [("tri10", False, 1e-11), ("h2o6", True, 1e-12)])
def test_generated_list_collocates(name, ortho, tol):
    """The generated arrays feed grid_create_task_list: collocation and
    integration of the list match the oracle's on the same arguments.  The
    10-atom triclinic cell sums thousands of periodic images per point with
    cancellation; on its coarsest level the oracle itself moves by about
    2e-11 when the tasks are summed in another order, so each level is held
    to the larger of the nominal bar and 4x that noise."""
    print("generated list", name, flush=True)
    args0, _, _ = synth.config(name, seed=2)
    nl = args0.meta["nl"]
    gl, args = grid.create_generated_task_list(nl)
    assert args.orthorhombic == ortho
    rng = np.random.default_rng(7)
    pab = synth.random_density(args, rng)
    ol = O.TaskList(args)
    g = gl.collocate(grid.GRID_FUNC_AB, pab)
    o, noise = order_noise(args, pab)
    for a, b, nz in zip(g, o, noise):
        assert replay.rel_err(a, b) < max(tol, 4 * nz)
    V = synth.random_potential(args, rng)
    hab, f, v = gl.integrate(V, pab, forces=True, virial=True)
    ho, fo, vo = ol.integrate(V, pab, forces=True, virial=True)
    assert replay.rel_err(hab, ho) < tol
    assert replay.rel_err(f, fo) < 1e-10


def test_generate_full_size_tri1024():
    """The 1024-atom triclinic workload's neighbour list (about 2.2 million
    tasks) at full size."""
    args, _, _ = synth.config("tri1024", seed=0)
    nl = args.meta["nl"]
    print("generate tri1024 full size", nl.npairs, "entries", flush=True)
    gpu, ora, ndiff = compare(nl, sample=500)
    assert gpu["ntasks"] == args.ntasks
    print(f"tri1024: {gpu['ntasks']} tasks, {gpu['nblocks']} blocks, {ndiff} radii differ, "
          f"device {gpu['device_ms']:.2f} ms", flush=True)
