"""Pins the CPU oracle before anything is compared against it.

- golden: the 13 known-answer files of the reference (single task, and
  through a task list with the reference's batched-replay conventions);
- brute force: analytic evaluation of all 35 grid_func operators and of the
  hab/force/virial integrals, with radii so large that truncation is
  invisible (pins the polynomial algebra independently of the reference);
- _ref: the reference's own single-task kernels compiled from its sources
  (random l ranges, funcs, geometry, border masks); skipped when not built;
- task-list layer: decontraction/contraction against a numpy einsum with a
  random contracted basis, and the collocate/integrate adjoint identity.
"""
import numpy as np
import pytest

import brute
import replay
import taskfile
from cp2k_amd import synth
from oracle import oracle as O

FIXTURES = taskfile.all_fixtures()
TOL = 1e-12          # src/grid/grid_task_list.c:241 (grid, hab)
DERIV_WEIGHT = 1e-4  # src/grid/grid_replay.c:445 (forces, virial)


@pytest.fixture(scope="module", params=FIXTURES, ids=lambda p: p.name.split(".")[0])
def fixture(request):
    return taskfile.load(request.param)


def test_single_task_collocate_golden(fixture):
    t = fixture
    assert replay.rel_err(O.collocate_pgf(t), t.grid) < TOL


def test_single_task_integrate_golden(fixture):
    t = fixture
    hab, f, v = O.integrate_pgf(t)
    assert replay.rel_err(hab, t.hab) < TOL
    assert replay.rel_err(f, t.forces) * DERIV_WEIGHT < TOL
    assert replay.rel_err(v[0] + v[1], t.virial) * DERIV_WEIGHT < TOL


def test_batched_golden(fixture):
    """Task-list path with the replay's dummy basis / task list."""
    t = fixture
    args, pab = replay.task_list(t, collocate=True)
    tl = O.TaskList(args)
    g = tl.collocate(t.func, pab)[0]
    assert replay.rel_err(g, t.grid) < TOL
    args, pab = replay.task_list(t, collocate=False)
    tl = O.TaskList(args)
    hab, f, v = tl.integrate([t.grid], pab, compute_tau=t.compute_tau, forces=True,
                             virial=True)
    assert replay.rel_err(hab.reshape(t.n2, t.n1), t.hab) < TOL
    assert replay.rel_err(f, t.forces) * DERIV_WEIGHT < TOL
    assert replay.rel_err(v, t.virial) * DERIV_WEIGHT < TOL


@pytest.mark.parametrize("ortho", [True, False])
@pytest.mark.parametrize("func", brute.FUNCS)
def test_collocate_all_funcs_brute(func, ortho):
    rng = np.random.default_rng(func * 2 + ortho)
    la = [(0, 1), (1, 2), (2, 2), (0, 3)][func % 4]
    lb = [(1, 1), (0, 2), (0, 0), (2, 3)][(func // 7) % 4]
    t = brute.SyntheticTask(rng, la, lb, ortho=ortho, func=func)
    g = O.collocate_pgf(t)
    b = brute.collocate(t, func)
    assert np.max(np.abs(g - b)) < 1e-12 * np.max(np.abs(b))


@pytest.mark.parametrize("ortho", [True, False])
@pytest.mark.parametrize("tau", [False, True])
def test_integrate_brute(ortho, tau):
    rng = np.random.default_rng(11 + 2 * ortho + tau)
    t = brute.SyntheticTask(rng, (1, 2), (0, 2), ortho=ortho)
    V = rng.normal(size=tuple(int(x) for x in t.npts_local[::-1]))
    hab, f, v = O.integrate_pgf(t, grid=V, compute_tau=tau)
    hb, fb, vb = brute.integrate(t, V, tau)
    for a, b in ((hab, hb), (f, fb), (v, vb)):
        assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(b))


def test_s_gaussian_integral_closed_form():
    """sum_g rho dV of an s-s product equals its analytic integral
    rscale * exp(-za zb/zp |rab|^2) (pi/zp)^(3/2) (fine grid, no wrap)."""
    rng = np.random.default_rng(5)
    t = brute.SyntheticTask(rng, (0, 0), (0, 0), ortho=True, h=0.12, n=(96, 96, 96))
    t.n1 = t.n2 = 1
    t.o1 = t.o2 = 0
    t.pab = np.ones((1, 1))
    g = O.collocate_pgf(t)
    zp = t.zeta + t.zetb
    exact = t.rscale * np.exp(-t.zeta * t.zetb / zp * t.rab @ t.rab) * (np.pi / zp) ** 1.5
    assert abs(g.sum() * abs(np.linalg.det(t.dh)) - exact) < 1e-9 * exact


# ---- the reference's own kernels (built from its sources into oracle/_ref) ----
ref = O.ref_lib()
needs_ref = pytest.mark.skipif(ref is None, reason="oracle/_ref not built")


def _random_task(rng, func, ortho, border):
    t = brute.SyntheticTask(rng, (int(rng.integers(0, 2)), int(rng.integers(1, 4))),
                            (int(rng.integers(0, 2)), int(rng.integers(1, 3))),
                            ortho=ortho, func=func, big_radius=False, zrange=(0.2, 20.0))
    t.la_min = min(t.la_min, t.la_max)
    t.lb_min = min(t.lb_min, t.lb_max)
    if border:
        t.border_mask = int(rng.integers(1, 64))
        t.border_width = rng.integers(1, 4, 3).astype(np.int32)
    return t


@needs_ref
@pytest.mark.parametrize("seed", range(6))
def test_oracle_vs_reference_kernels(seed):
    rng = np.random.default_rng(100 + seed)
    for func in brute.FUNCS:
        for ortho, border in ((True, False), (False, False), (True, True)):
            t = _random_task(rng, func, ortho, border)
            a = O.collocate_pgf(t)
            b = O.collocate_pgf(t, libh=ref, prefix="grid_ref")
            assert replay.rel_err(a, b) < TOL, (func, ortho, border)
    for tau in (False, True):
        for ortho, border in ((True, False), (False, False), (True, True)):
            t = _random_task(rng, 100, ortho, border)
            V = rng.normal(size=tuple(int(x) for x in t.npts_local[::-1]))
            h1, f1, v1 = O.integrate_pgf(t, grid=V, compute_tau=tau)
            h2, f2, v2 = O.integrate_pgf(t, grid=V, compute_tau=tau, libh=ref,
                                         prefix="grid_ref")
            assert replay.rel_err(h1, h2) < TOL
            assert replay.rel_err(f1, f2) < 1e-10
            assert replay.rel_err(v1, v2) < 1e-10


@needs_ref
@pytest.mark.parametrize("ortho", [True, False])
def test_oracle_hab_derivatives_vs_reference(ortho):
    """hdab / hadb / a_hdab (grid_ref_integrate.c:126-149) against the
    reference's own single-task integrate, with and without forces."""
    rng = np.random.default_rng(300 + ortho)
    for rep in range(4):
        t = _random_task(rng, 100, ortho, border=(rep == 3 and ortho))
        V = rng.normal(size=tuple(int(x) for x in t.npts_local[::-1]))
        with_f = rep % 2 == 0
        a = O.integrate_pgf(t, grid=V, compute_tau=False, forces=with_f, virial=with_f,
                            derivs=True)
        b = O.integrate_pgf(t, grid=V, compute_tau=False, forces=with_f, virial=with_f,
                            libh=ref, prefix="grid_ref", derivs=True)
        assert replay.rel_err(a[0], b[0]) < TOL
        if with_f:
            assert replay.rel_err(a[1], b[1]) < 1e-10
            assert replay.rel_err(a[2], b[2]) < 1e-10
        for x, y in zip(a[3], b[3]):
            assert replay.rel_err(x, y) < 1e-10


@pytest.mark.parametrize("ortho", [True, False])
def test_hab_derivatives_finite_difference(ortho):
    """hdab = d hab / d ra and hadb = d hab / d rb (the other centre fixed),
    by central differences of the oracle's hab with truncation invisible;
    the density-weighted sums of hdab / hadb are the forces on a / b."""
    rng = np.random.default_rng(41 + ortho)
    t = brute.SyntheticTask(rng, (0, 2), (1, 2), ortho=ortho)
    V = rng.normal(size=tuple(int(x) for x in t.npts_local[::-1]))
    hab, f, v, (hdab, hadb, a_hdab) = O.integrate_pgf(t, grid=V, derivs=True)
    d = 1e-5
    ra0, rab0 = t.ra.copy(), t.rab.copy()
    for i in range(3):
        e = np.zeros(3)
        e[i] = d
        hp = []
        for sgn in (1, -1):
            t.ra, t.rab = ra0 + sgn * e, rab0 - sgn * e  # move a, b stays
            hp.append(O.integrate_pgf(t, grid=V, forces=False, virial=False)[0])
        fd_a = (hp[0] - hp[1]) / (2 * d)
        hp = []
        for sgn in (1, -1):
            t.ra, t.rab = ra0, rab0 + sgn * e  # move b, a stays
            hp.append(O.integrate_pgf(t, grid=V, forces=False, virial=False)[0])
        fd_b = (hp[0] - hp[1]) / (2 * d)
        t.ra, t.rab = ra0, rab0
        scale = np.max(np.abs(fd_a)) + np.max(np.abs(fd_b))
        assert np.max(np.abs(hdab[..., i] - fd_a)) < 1e-6 * scale
        assert np.max(np.abs(hadb[..., i] - fd_b)) < 1e-6 * scale
        mask = np.zeros_like(t.pab)
        a0 = brute._ncoset(t.la_min - 1)
        b0 = brute._ncoset(t.lb_min - 1)
        mask[t.o2 + b0:t.o2 + brute._ncoset(t.lb_max), t.o1 + a0:t.o1 + brute._ncoset(t.la_max)] = 1
        pm = t.pab * mask
        assert abs(np.sum(pm * hdab[..., i]) - f[0, i]) < 1e-10 * np.max(np.abs(f))
        assert abs(np.sum(pm * hadb[..., i]) - f[1, i]) < 1e-10 * np.max(np.abs(f))
    assert np.allclose(np.einsum("ba,baij->ij", pm, a_hdab), v[0], rtol=0, atol=1e-10 * np.max(np.abs(v)))


def test_buffer_pack_unpack():
    """rs_pack_buffer / rs_unpack_buffer restatement: the blocks of a
    generated density matrix pack back into the generator's buffer, and
    pack -> unpack into zero blocks returns the blocks (any order)."""
    from oracle import buffers as B
    args, pab, _ = synth.config("h2o2", seed=4, n=16)
    offs = np.asarray(args.block_offsets)
    sizes = np.diff(np.append(offs, args.buffer_size))
    blocks = [pab[o:o + n].copy() for o, n in zip(offs, sizes)]
    perm = np.random.default_rng(2).permutation(len(blocks))
    buf = B.pack_buffer([blocks[i] for i in perm], offs[perm], args.buffer_size)
    assert np.array_equal(buf, pab)
    back = [np.zeros(n) for n in sizes]
    B.unpack_buffer(back, offs, buf)
    assert all(np.array_equal(a, b) for a, b in zip(back, blocks))
    B.unpack_buffer(back, offs, buf)
    assert all(np.array_equal(a, 2 * b) for a, b in zip(back, blocks))


def _band_limited(rng, shape_box, nterms=6):
    """A real trigonometric sum with Miller vectors strictly inside the box
    of shape_box (no Nyquist plane): [(h, a, b)]."""
    lim = [(n - 1) // 2 for n in shape_box[::-1]]  # (x, y, z) of an [nz, ny, nx] shape
    return [(np.array([rng.integers(-m, m + 1) for m in lim]), rng.normal(), rng.normal())
            for _ in range(nterms)]


def _sample(terms, shape):
    """The sum at the points of a grid of shape [nz, ny, nx] of one cell:
    sum a cos(2 pi h.x) + b sin(2 pi h.x), x_d = j_d / n_d (h = (hx, hy, hz))."""
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx,
                          indexing="ij")
    out = np.zeros(shape)
    for h, a, b in terms:
        ph = 2 * np.pi * (h[0] * x + h[1] * y + h[2] * z)
        out += a * np.cos(ph) + b * np.sin(ph)
    return out


@pytest.mark.parametrize("shapes", [[(27, 25, 21), (15, 15, 9), (9, 7, 5)],
                                    [(24, 20, 18), (12, 10, 8)]])
def test_rs_pw_closed_forms(shapes):
    """density_rs2pw of per-level band-limited functions is their sum on the
    finest grid; potential_pw2rs of a band-limited potential samples it on
    every level times N_0 / N_l; charge is conserved; and the two are
    adjoint: <rho(rs), v> = sum_l <rs_l, v_l>."""
    from oracle import rs_pw
    rng = np.random.default_rng(len(shapes))
    coarsest = shapes[-1]
    per_level = [_band_limited(rng, coarsest) for _ in shapes]
    rs = [_sample(t, s) for t, s in zip(per_level, shapes)]
    rho_r, rho_g = rs_pw.density_rs2pw(rs)
    want = sum(_sample(t, shapes[0]) for t in per_level)
    assert np.max(np.abs(rho_r - want)) < 1e-12 * np.max(np.abs(want))
    vt = _band_limited(rng, coarsest)
    out = rs_pw.potential_pw2rs(_sample(vt, shapes[0]), shapes)
    for s, o in zip(shapes, out):
        w = _sample(vt, s) * (np.prod(shapes[0]) / np.prod(s))
        assert np.max(np.abs(o - w)) < 1e-12 * np.max(np.abs(w))
    # arbitrary (not band-limited) data
    rs = [rng.normal(size=s) for s in shapes]
    rho_r, rho_g = rs_pw.density_rs2pw(rs)
    N0 = np.prod(shapes[0])
    assert abs(rho_r.sum() / N0 - sum(g.sum() / g.size for g in rs)) < 1e-12 * N0
    assert abs(rho_g[0, 0, 0].real - sum(g.mean() for g in rs)) < 1e-13
    v = rng.normal(size=shapes[0])
    vl = rs_pw.potential_pw2rs(v, shapes)
    lhs = float(np.sum(rho_r * v))
    rhs = sum(float(np.sum(g * w)) for g, w in zip(rs, vl))
    assert abs(lhs - rhs) < 1e-10 * (abs(lhs) + 1.0)


def test_rs_pw_transfer_halo():
    from oracle import rs_pw
    rng = np.random.default_rng(3)
    pw = rng.normal(size=(6, 5, 7))
    rs = rs_pw.transfer_pw2rs(pw, (2, 1, 3))
    assert rs.shape == (12, 7, 11)
    assert np.array_equal(rs_pw.transfer_rs2pw(rs, (2, 1, 3)), pw)
    assert rs[0, 0, 0] == pw[-3, -1, -2] and rs[-1, -1, -1] == pw[2, 0, 1]


# ---- task-list layer ----
def _tasks_by_hand(args, pab_blocks, func):
    """Collocation of a task list via numpy decontraction + single-task oracle."""
    grids = [np.zeros(s) for s in args.grid_shapes()]
    for t in range(args.ntasks):
        ia, ja = int(args.iatom_list[t]), int(args.jatom_list[t])
        A = args.basis_sets[args.atom_kinds[ia - 1] - 1]
        B = args.basis_sets[args.atom_kinds[ja - 1] - 1]
        s1, s2 = int(args.iset_list[t]) - 1, int(args.jset_list[t]) - 1
        blk = int(args.block_num_list[t]) - 1
        off = int(args.block_offsets[blk])
        if ia <= ja:
            S = pab_blocks[off:off + A.nsgf * B.nsgf].reshape(B.nsgf, A.nsgf).T
        else:
            S = pab_blocks[off:off + A.nsgf * B.nsgf].reshape(A.nsgf, B.nsgf)
        a0, b0 = int(A.first_sgf[s1]) - 1, int(B.first_sgf[s2]) - 1
        Sa = S[a0:a0 + A.nsgf_set[s1], b0:b0 + B.nsgf_set[s2]]
        ncoa, ncob = A.ncoa(s1), B.ncoa(s2)
        pa = A.sphi[a0:a0 + A.nsgf_set[s1], :ncoa]
        pb = B.sphi[b0:b0 + B.nsgf_set[s2], :ncob]
        cart = np.einsum("sc,sd,de->ce", pb, Sa.T, pa)  # [ncob][ncoa]
        lev = int(args.level_list[t]) - 1

        class T:
            pass
        k = T()
        k.orthorhombic = args.orthorhombic
        k.border_mask = int(args.border_mask_list[t])
        k.la_max, k.la_min = int(A.lmax[s1]), int(A.lmin[s1])
        k.lb_max, k.lb_min = int(B.lmax[s2]), int(B.lmin[s2])
        k.zeta = A.zet[s1, int(args.ipgf_list[t]) - 1]
        k.zetb = B.zet[s2, int(args.jpgf_list[t]) - 1]
        k.rscale = 1.0 if ia == ja else 2.0
        k.dh, k.dh_inv = args.dh[lev], args.dh_inv[lev]
        k.ra = args.atom_positions[ia - 1]
        k.rab = args.rab_list[t]
        k.npts_global, k.npts_local = args.npts_global[lev], args.npts_local[lev]
        k.shift_local, k.border_width = args.shift_local[lev], args.border_width[lev]
        k.radius = args.radius_list[t]
        k.o1 = (int(args.ipgf_list[t]) - 1) * (ncoa // int(A.npgf[s1]))
        k.o2 = (int(args.jpgf_list[t]) - 1) * (ncob // int(B.npgf[s2]))
        k.n1, k.n2 = ncoa, ncob
        k.pab = cart
        O.collocate_pgf(k, func=func, grid=grids[lev])
    return grids


def test_task_list_contraction_by_hand():
    args, pab, _ = synth.config("h2o2", seed=3, n=32, flip=0.5)
    sel = np.arange(0, args.ntasks, max(1, args.ntasks // 400))
    sub = _subset(args, sel)
    tl = O.TaskList(sub)
    got = tl.collocate(100, pab)
    want = _tasks_by_hand(sub, pab, 100)
    for g, w in zip(got, want):
        assert np.max(np.abs(g - w)) <= 1e-12 * max(1.0, np.max(np.abs(w)))


def _subset(args, sel):
    import dataclasses
    kw = {f.name: getattr(args, f.name) for f in dataclasses.fields(args)}
    for name in ("level_list", "iatom_list", "jatom_list", "iset_list", "jset_list",
                 "ipgf_list", "jpgf_list", "border_mask_list", "block_num_list",
                 "radius_list", "rab_list"):
        kw[name] = kw[name][sel]
    return type(args)(**kw)


@pytest.mark.parametrize("name", ["ref6", "h2o2", "tri12"])
def test_adjoint_identity(name):
    """sum_g rho[P](g) V(g) = sum_blocks s_ij <P_blk, H_blk[V]>, s = 2 off-site
    (rscale in collocate, grid_ref_task_list.c:369; hab unscaled)."""
    args, pab, V = synth.config(name, seed=7, flip=0.5)
    tl = O.TaskList(args)
    rho = tl.collocate(100, pab)
    hab, _, _ = tl.integrate(V)
    lhs = sum(float(np.sum(r * v)) for r, v in zip(rho, V))
    w = np.zeros(args.buffer_size)
    bi, bj = args.meta["block_i"], args.meta["block_j"]
    for b in range(args.nblocks):
        off = int(args.block_offsets[b])
        n = (args.block_offsets[b + 1] if b + 1 < args.nblocks else args.buffer_size) - off
        w[off:off + n] = 1.0 if bi[b] == bj[b] else 2.0
    rhs = float(np.sum(w * pab * hab))
    assert abs(lhs - rhs) <= 1e-10 * (abs(lhs) + 1.0)
