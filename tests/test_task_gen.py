"""Task-list generation: the oracle (oracle/task_gen.c) pinned against
closed forms and an independent numpy restatement (tests/gen_brute.py), and
the device generator's library export.  The device generator itself is
compared with the oracle in test_gpu_task_gen.py."""
import math

import numpy as np
import pytest

import gen_brute as GB
from cp2k_amd import synth
from cp2k_amd.tasklist import NeighbourList
from oracle import oracle as O

INT_KEYS = ("level_list", "iatom_list", "jatom_list", "iset_list", "jset_list", "ipgf_list",
            "jpgf_list", "block_num_list")


def variant(nl, rng, images=False, kind_radius=False, shuffle=False):
    """The neighbour list with periodic-image indices (every entry repeated
    under images 1..3, so equal (iatom, jatom) pairs meet in the sort and
    the block keys), smaller kind radii (the set loop's kind screening then
    removes tasks on its own) or a shuffled entry order."""
    kw = dict(orthorhombic=nl.orthorhombic, basis_sets=nl.basis_sets,
              atom_positions=nl.atom_positions, atom_kinds=nl.atom_kinds,
              pair_atoms=nl.pair_atoms, pair_rab=nl.pair_rab, pair_image=nl.pair_image,
              set_radius=nl.set_radius, pgf_radius=nl.pgf_radius, kind_radius=nl.kind_radius,
              cutoff=nl.cutoff, rel_cutoff=nl.rel_cutoff, eps_rho=nl.eps_rho,
              npts_global=nl.npts_global, dh=nl.dh, dh_inv=nl.dh_inv)
    if images:
        reps = 3
        kw["pair_atoms"] = np.tile(nl.pair_atoms, (reps, 1))
        kw["pair_rab"] = np.concatenate([nl.pair_rab + 0.01 * r for r in range(reps)])
        kw["pair_image"] = np.repeat(np.arange(1, reps + 1), nl.npairs)
    if kind_radius:
        kw["kind_radius"] = 0.8 * nl.set_radius.max(axis=1)
    if shuffle:
        perm = rng.permutation(len(kw["pair_atoms"]))
        kw["pair_atoms"] = kw["pair_atoms"][perm]
        kw["pair_rab"] = kw["pair_rab"][perm]
        if kw["pair_image"] is not None:
            kw["pair_image"] = kw["pair_image"][perm]
    return NeighbourList(**kw)


def neighbour_list(name, **kw):
    args, _, _ = synth.config(name, seed=kw.pop("seed", 0))
    nl = args.meta["nl"]
    return variant(nl, np.random.default_rng(1), **kw) if kw else nl


# ---------------------------------------------------------------------------
# exp_radius (src/aobasis/ao_util.F:95-190)
# ---------------------------------------------------------------------------
def test_exp_radius_closed_form_l0():
    """l = 0: d exp(-a r^2) = t at r* = sqrt(ln(d / t) / a); the result is
    the upper end of a bracket around r* narrower than epsabs."""
    rng = np.random.default_rng(3)
    for _ in range(300):
        a = 10 ** rng.uniform(-2, 3)
        d = 10 ** rng.uniform(-3, 3)
        t = 10 ** rng.uniform(-14, -4)
        rstar = math.sqrt(math.log(d / t) / a)
        lo, hi = O.exp_radius(0, a, t, d, 1.0e-2, 0.0, bracket=True)
        slack = 1e-5 * rstar  # single-precision exponentials
        assert lo - slack <= rstar <= hi + slack, (a, d, t, lo, rstar, hi)
        assert hi - lo < 1.0e-2


def test_exp_radius_bracket_general_l():
    """l > 0: d r^l exp(-a r^2) is below t at the result and not below t at
    the lower end of the final bracket, which lies beyond the maximum."""
    rng = np.random.default_rng(4)
    for _ in range(300):
        l = int(rng.integers(1, 9))
        a = 10 ** rng.uniform(-2, 2)
        d = 10 ** rng.uniform(-2, 4)
        t = 10 ** rng.uniform(-12, -5)
        lo, hi = O.exp_radius(l, a, t, d, 1.0e-2, 0.0, bracket=True)
        g = lambda r: d * r ** l * math.exp(-a * r * r)  # noqa: E731
        rmax = math.sqrt(0.5 * l / a)
        if hi == 0.0:  # below t everywhere: the maximum is below t
            assert g(rmax) < t * (1 + 1e-5)
            continue
        assert g(hi) < t * (1 + 1e-5)
        assert lo >= rmax * (1 - 1e-12)
        assert g(lo) >= t * (1 - 1e-5)
        assert hi - lo < 1.0e-2


def test_exp_radius_special_cases():
    # zero prefactor, and a Gaussian already below the threshold at its
    # maximum, return rlow
    assert O.exp_radius(2, 1.0, 1e-10, 0.0, 1e-2, 1.5) == 1.5
    assert O.exp_radius(3, 1.0, 1.0, 1e-3, 1e-2, 0.25) == 0.25
    # l = 0 skips the early test at the start point: with rlow beyond the
    # crossing the bisection still runs and ends within a bracket of rlow
    r = O.exp_radius(0, 1.0, 1e-10, 1.0, 1e-2, 50.0)
    assert 50.0 < r < 50.0 + 1e-2
    assert O.exp_radius(1, 1.0, 1e-10, 1.0, 1e-2, 50.0) == 50.0


def test_polycoef_binomial_expansion():
    """poly[k] = max over lxa, lxb of the r^k coefficient of
    (r + rad_a)^lxa (r + rad_b)^lxb (numpy polynomial products)."""
    rng = np.random.default_rng(5)
    for _ in range(60):
        la, lb = int(rng.integers(0, 6)), int(rng.integers(0, 6))
        ra_, rb_ = rng.uniform(0, 3, 2)
        np.testing.assert_allclose(O.product_polycoef(la, lb, ra_, rb_),
                                   GB.polycoef_binomial(la, lb, ra_, rb_), rtol=1e-13, atol=0)


def test_gridlevel_cases():
    """gaussian_gridlevel: the last (coarsest) level whose cutoff + 1e-6
    covers |zetp| * rel_cutoff, level 1 when none does."""
    cutoff = [400.0, 133.0, 44.0, 15.0]
    cases = {0.1: 4, 0.375: 4, 0.376: 3, 1.0: 3, 1.1: 3, 1.2: 2, 3.325: 2, 3.4: 1, 10.0: 1,
             50.0: 1, -0.2: 4}
    for zetp, level in cases.items():
        assert O.gridlevel(cutoff, 40.0, zetp) == level, zetp
    # the 1e-6 margin: exactly at the cutoff and just inside the margin
    assert O.gridlevel([10.0, 5.0], 1.0, 5.0) == 2
    assert O.gridlevel([10.0, 5.0], 1.0, 5.0 + 5e-7) == 2
    assert O.gridlevel([10.0, 5.0], 1.0, 5.0 + 2e-6) == 1


# ---------------------------------------------------------------------------
# the whole list
# ---------------------------------------------------------------------------
def check_against_brute(nl, out, sample=2000):
    ref = GB.brute_task_order(nl)
    assert out["ntasks"] == len(ref["level_list"])
    for k in INT_KEYS:
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)
    np.testing.assert_array_equal(out["rab_list"], ref["rab_list"])
    np.testing.assert_array_equal(out["border_mask_list"], 0)
    assert out["nblocks"] == ref["nblocks"]
    np.testing.assert_array_equal(out["block_offsets"], ref["block_offsets"])
    assert out["buffer_size"] == ref["buffer_size"]
    bad = GB.radius_bracket_violations(nl, out, ref, sample=sample)
    assert not bad, f"{len(bad)} radii fail the exp_radius bracket property, first {bad[:5]}"
    return ref


@pytest.mark.parametrize("name,kw", [
    ("ref16", {}), ("tri12", {}), ("h2o4", {}), ("hl12", {}),
    ("tri8", dict(images=True)), ("h2o4", dict(kind_radius=True)),
    ("ref24", dict(shuffle=True, images=True)),
])
def test_oracle_generation_vs_brute(name, kw):
    nl = neighbour_list(name, **kw)
    out = O.generate_task_list(nl)
    check_against_brute(nl, out)


def test_kind_radius_screening_removes_tasks():
    nl = neighbour_list("h2o4")
    small = variant(nl, None, kind_radius=True)
    assert O.generate_task_list(small)["ntasks"] < O.generate_task_list(nl)["ntasks"]


def test_oracle_generation_matches_synthetic_lists():
    """The synthetic workloads keep a task when the two pgf radii reach |rab|
    (the set and kind screens are implied by it for the radii they use): the
    generator finds the same tasks, blocks and buffer size."""
    args, _, _ = synth.config("tri12")
    out = O.generate_task_list(args.meta["nl"])
    assert out["ntasks"] == args.ntasks
    assert out["nblocks"] == args.nblocks
    assert out["buffer_size"] == args.buffer_size

    def keys(lv, ia, ja):
        return sorted(zip(lv.tolist(), np.minimum(ia, ja).tolist(), np.maximum(ia, ja).tolist()))

    assert keys(out["level_list"], out["iatom_list"], out["jatom_list"]) == \
        keys(args.level_list, args.iatom_list, args.jatom_list)


def test_empty_neighbour_list():
    nl = neighbour_list("ref16")
    empty = NeighbourList(**{**nl.__dict__, "pair_atoms": np.zeros((0, 2)),
                             "pair_rab": np.zeros((0, 3))})
    out = O.generate_task_list(empty)
    assert out["ntasks"] == 0 and out["nblocks"] == 0 and out["buffer_size"] == 0


def test_generator_exported():
    """The product library exports the generator (no device needed)."""
    from cp2k_amd import grid
    L = grid.lib()
    assert hasattr(L, "grid_mi355x_generate_task_list")
    assert hasattr(L, "grid_mi355x_generate_task_list_ms")
