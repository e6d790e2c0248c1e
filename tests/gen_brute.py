"""Independent numpy checks of task-list generation (the oracle's
oracle_generate_task_list and the device generator are compared with these).

- brute_task_order: the screening of task_list_inner_loop
  (src/task_list_methods.F:652-719) as boolean masks over all candidates
  (pair, iset, jset, ipgf, jpgf), the grid-level rule of gaussian_gridlevel
  (src/pw_env/gaussian_gridlevels.F:146-166), the stable order of
  tasks_less_than (:2931-2960) as one lexicographic sort with the generation
  position as the last key, and the blocks of get_atom_pair (:2070-2136) as
  np.unique over (min atom, max atom, image).
- radius_bracket_violations: the product radius of
  exp_radius_very_extended (src/aobasis/ao_util.F:207-290) through what it
  must satisfy: every term prefactor * poly[i] * R^i exp(-zetp R^2) of the
  polynomial envelope is below eps at R, and the term that set R is not
  below eps a bracket width (1e-2) below R, with poly from the binomial
  expansion of (r + |ra - rp|)^lxa (r + |rb - rp|)^lxb.
"""
import math

import numpy as np
from numpy.polynomial import polynomial as npoly


def brute_candidates(nl):
    """All tasks in generation order: dict of arrays pair, iset, jset, ipgf,
    jpgf (one based sets / pgfs)."""
    pa = nl.pair_atoms
    ka = nl.atom_kinds[pa[:, 0] - 1] - 1
    kb = nl.atom_kinds[pa[:, 1] - 1] - 1
    dab = np.sqrt((nl.pair_rab ** 2).sum(axis=1))
    if nl.kind_radius is None:
        kind_r = nl.set_radius.max(axis=1)
    else:
        kind_r = nl.kind_radius
    rows = []
    nk = len(nl.basis_sets)
    for a in range(nk):
        for b in range(nk):
            sel = np.nonzero((ka == a) & (kb == b))[0]
            if sel.size == 0:
                continue
            A, B = nl.basis_sets[a], nl.basis_sets[b]
            d = dab[sel]
            for s1 in range(A.nset):
                for s2 in range(B.nset):
                    for p1 in range(int(A.npgf[s1])):
                        for p2 in range(int(B.npgf[s2])):
                            keep = ((nl.set_radius[a, s1] + kind_r[b] >= d)
                                    & (nl.set_radius[a, s1] + nl.set_radius[b, s2] >= d)
                                    & (nl.pgf_radius[a, s1, p1] + nl.set_radius[b, s2] >= d)
                                    & (nl.pgf_radius[a, s1, p1] + nl.pgf_radius[b, s2, p2] >= d))
                            m = sel[keep]
                            if m.size:
                                rows.append(np.stack([m, np.full(m.size, s1 + 1),
                                                      np.full(m.size, s2 + 1),
                                                      np.full(m.size, p1 + 1),
                                                      np.full(m.size, p2 + 1)], axis=1))
    if not rows:
        return np.zeros((0, 5), np.int64)
    t = np.concatenate(rows)
    order = np.lexsort((t[:, 4], t[:, 3], t[:, 2], t[:, 1], t[:, 0]))
    return t[order]


def gridlevel_np(zetp, cutoff, rel_cutoff):
    need = np.abs(zetp) * rel_cutoff
    ok = (np.asarray(cutoff)[None, :] + 1e-6) >= need[:, None]
    idx = np.arange(1, len(cutoff) + 1)[None, :]
    return np.where(ok, idx, 0).max(axis=1).clip(min=1)


def brute_task_order(nl):
    """The generator's output lists, restated with numpy."""
    t = brute_candidates(nl)
    p, s1, s2, p1, p2 = t.T
    ia = nl.pair_atoms[p, 0]
    ja = nl.pair_atoms[p, 1]
    img = np.ones(len(p), np.int64) if nl.pair_image is None else nl.pair_image[p]
    ka = nl.atom_kinds[ia - 1] - 1
    kb = nl.atom_kinds[ja - 1] - 1
    za = np.array([nl.basis_sets[k].zet[s - 1, q - 1] for k, s, q in zip(ka, s1, p1)])
    zb = np.array([nl.basis_sets[k].zet[s - 1, q - 1] for k, s, q in zip(kb, s2, p2)])
    level = gridlevel_np(za + zb, nl.cutoff, nl.rel_cutoff)
    gen = np.arange(len(p))
    order = np.lexsort((gen, p2, p1, s2, s1, ja, ia, img, level))
    row, col = np.minimum(ia, ja), np.maximum(ia, ja)
    keys = np.stack([row, col, img], axis=1)
    uniq, block = np.unique(keys, axis=0, return_inverse=True)
    block = block.reshape(-1) + 1
    nsgf = np.array([b.nsgf for b in nl.basis_sets])
    sizes = nsgf[nl.atom_kinds[uniq[:, 0] - 1] - 1] * nsgf[nl.atom_kinds[uniq[:, 1] - 1] - 1]
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    return dict(level_list=level[order], iatom_list=ia[order], jatom_list=ja[order],
                iset_list=s1[order], jset_list=s2[order], ipgf_list=p1[order],
                jpgf_list=p2[order], block_num_list=block[order],
                rab_list=nl.pair_rab[p[order]], nblocks=len(uniq), block_offsets=offsets,
                buffer_size=int(sizes.sum()), zeta=za[order], zetb=zb[order])


def polycoef_binomial(la_max, lb_max, rad_a, rad_b):
    """max over lxa <= la_max, lxb <= lb_max of the coefficients of r^k in
    (r + rad_a)^lxa (r + rad_b)^lxb."""
    poly = np.zeros(la_max + lb_max + 1)
    for lxa in range(la_max + 1):
        for lxb in range(lb_max + 1):
            c = npoly.polymul(npoly.polypow([rad_a, 1.0], lxa), npoly.polypow([rad_b, 1.0], lxb))
            poly[:len(c)] = np.maximum(poly[:len(c)], c)
    return poly


def radius_bracket_violations(nl, out, ref=None, rel=1e-5, width=0.0101, sample=None,
                              seed=0):
    """Indices of tasks whose radius fails the defining property of
    exp_radius_very_extended (see module docstring).  ref: the brute lists
    (for zeta / zetb); computed when None.  sample: check that many tasks,
    drawn at random (all when None)."""
    ref = brute_task_order(nl) if ref is None else ref
    R = out["radius_list"]
    bad = []
    lmax = {k: b.lmax for k, b in enumerate(nl.basis_sets)}
    idx = np.arange(len(R))
    if sample is not None and sample < len(R):
        idx = np.sort(np.random.default_rng(seed).choice(len(R), sample, replace=False))
    for i in idx:
        ia, ja = out["iatom_list"][i], out["jatom_list"][i]
        ka = nl.atom_kinds[ia - 1] - 1
        kb = nl.atom_kinds[ja - 1] - 1
        la = int(lmax[ka][out["iset_list"][i] - 1])
        lb = int(lmax[kb][out["jset_list"][i] - 1])
        za, zb = ref["zeta"][i], ref["zetb"][i]
        zp = za + zb
        rab = out["rab_list"][i]
        d = math.sqrt(float((rab ** 2).sum()))
        pref = math.exp(-za * zb / zp * d * d)
        poly = polycoef_binomial(la, lb, zb / zp * d, za / zp * d) * pref
        r = R[i]

        def terms(x):
            return np.array([poly[k] * x ** k * math.exp(-zp * x * x) for k in range(len(poly))])

        if np.any(terms(r) >= nl.eps_rho * (1 + rel)):
            bad.append(i)
            continue
        if r - width <= 0.0:
            continue
        # the term that set R was not below eps at the lower end of the final
        # bracket, which lies above R - width and at or beyond that term's
        # maximum sqrt(k / (2 zetp)), where the term decreases
        hit = False
        for k in range(len(poly)):
            y = max(r - width, math.sqrt(0.5 * k / zp))
            if y < r and poly[k] * y ** k * math.exp(-zp * y * y) >= nl.eps_rho * (1 - rel):
                hit = True
                break
        if not hit:
            bad.append(i)
    return bad
