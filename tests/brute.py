"""Brute-force evaluation of single-task collocation/integration.

Independent of the reference's algorithm: every grid point's value is the
analytic Gaussian-monomial product (and its derivatives) evaluated directly
at the point's minimum image, with no truncation.  Used with a radius large
enough that truncation is invisible, so it pins the polynomial algebra
(prepare_pab stencils of all 35 functions, re-centring, coefficient
contractions, hab/force/virial getters) without retyping any of it.

A derivative of phi_a(r) = (r-A)^a exp(-z|r-A|^2) is kept as a sum of
terms c * (r-A)^e * exp(...), using d/dx_k (x^e g) = e_k x^(e-1_k) g - 2 z x^(e+1_k) g.
"""
import itertools

import numpy as np


def orbitals(lmin, lmax):
    """(coset index, (lx,ly,lz)) for shells lmin..lmax, CP2K ordering."""
    out = []
    idx = 0
    for l in range(0, lmax + 1):
        for lx in range(l, -1, -1):
            for lz in range(0, l - lx + 1):
                if l >= lmin:
                    out.append((idx, (lx, l - lx - lz, lz)))
                idx += 1
    return out


def deriv(terms, k, z):
    """Apply d/dx_k to a list of (coef, exponent tuple)."""
    out = []
    for c, e in terms:
        if e[k] > 0:
            e1 = list(e)
            e1[k] -= 1
            out.append((c * e[k], tuple(e1)))
        e2 = list(e)
        e2[k] += 1
        out.append((-2.0 * z * c, tuple(e2)))
    return out


def times_coord(terms, k):
    out = []
    for c, e in terms:
        e1 = list(e)
        e1[k] += 1
        out.append((c, tuple(e1)))
    return out


def evaluate(terms, x, gauss):
    v = np.zeros(x.shape[:-1])
    for c, e in terms:
        v += c * x[..., 0] ** e[0] * x[..., 1] ** e[1] * x[..., 2] ** e[2]
    return v * gauss


def displacements(t):
    """Cartesian displacement of each local grid point ([nz,ny,nx,3]) from the
    product centre rp, minimum image; requires shift_local == 0."""
    nx, ny, nz = [int(v) for v in t.npts_local]
    N = np.array([int(v) for v in t.npts_global], dtype=float)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    g = np.stack([i, j, k], -1).astype(float)
    f = t.zetb / (t.zeta + t.zetb)
    rp = t.ra + f * t.rab
    gp = rp @ t.dh_inv
    dg = g - gp
    dg -= N * np.round(dg / N)
    return dg @ t.dh, f


FUNCS = ([100, 200] + [301, 302, 303] + [411, 412, 413, 421, 422, 423, 431, 432, 433]
         + [501, 502, 503] + [601, 602, 603] + [701, 702, 703] + [801, 802, 803]
         + [901, 902, 903, 904, 905, 906] + [1001, 1002, 1003])


def product_terms(func, ea, eb, za, zb):
    """func applied to (phi_a, phi_b): list of (termsA, termsB, which) pairs
    whose products sum to the collocated function; `which` selects the
    coordinate origin of an extra (r-B) factor (handled in termsB)."""
    A = [(1.0, ea)]
    B = [(1.0, eb)]
    if func == 100:
        return [(A, B)]
    if func == 200:
        return [([(0.5 * c, e) for c, e in deriv(A, k, za)], deriv(B, k, zb)) for k in range(3)]
    if 301 <= func <= 303:
        d = func - 301
        return [(A, deriv(B, d, zb)), ([(-c, e) for c, e in deriv(A, d, za)], B)]
    if 411 <= func <= 433:
        d, r = (func // 10) % 10 - 1, func % 10 - 1
        return [(A, times_coord(deriv(B, d, zb), r)),
                ([(-c, e) for c, e in deriv(A, d, za)], times_coord(B, r))]
    if 501 <= func <= 503:
        d = func - 501
        return [(A, deriv(B, d, zb)), (deriv(A, d, za), B)]
    if 601 <= func <= 603:
        d = func - 601
        return [(deriv(A, d, za), deriv(B, d, zb))]
    if 701 <= func <= 703:
        d1 = func - 701
        d2 = (d1 + 1) % 3
        return [(deriv(deriv(A, d2, za), d1, za), deriv(deriv(B, d2, zb), d1, zb))]
    if 801 <= func <= 803:
        d = func - 801
        return [(deriv(deriv(A, d, za), d, za), deriv(deriv(B, d, zb), d, zb))]
    if 901 <= func <= 903:
        d = func - 901
        return [([(-c, e) for c, e in deriv(A, d, za)], B)]
    if 904 <= func <= 906:
        d = func - 904
        return [(A, [(-c, e) for c, e in deriv(B, d, zb)])]
    if 1001 <= func <= 1003:
        d = func - 1001
        return [([(2.0 * za * c, e) for c, e in times_coord(A, d)], B)]
    raise ValueError(func)


def collocate(t, func):
    """rho on the local grid, [nz,ny,nx]."""
    d, f = displacements(t)
    xa = d + f * t.rab          # r - A
    xb = d - (1.0 - f) * t.rab  # r - B
    ga = np.exp(-t.zeta * np.sum(xa * xa, -1))
    gb = np.exp(-t.zetb * np.sum(xb * xb, -1))
    rho = np.zeros(d.shape[:-1])
    for ia, ea in orbitals(t.la_min, t.la_max):
        for ib, eb in orbitals(t.lb_min, t.lb_max):
            p = t.pab[t.o2 + ib, t.o1 + ia]
            if p == 0.0:
                continue
            for TA, TB in product_terms(func, ea, eb, t.zeta, t.zetb):
                rho += p * evaluate(TA, xa, ga) * evaluate(TB, xb, gb)
    return t.rscale * rho


def integrate(t, V, compute_tau):
    """hab [n2][n1] (only the o1/o2 sub-block), forces [2,3], virials [2,3,3]."""
    d, f = displacements(t)
    xa = d + f * t.rab
    xb = d - (1.0 - f) * t.rab
    ga = np.exp(-t.zeta * np.sum(xa * xa, -1))
    gb = np.exp(-t.zetb * np.sum(xb * xb, -1))
    za, zb = t.zeta, t.zetb
    hab = np.zeros((t.n2, t.n1))
    forces = np.zeros((2, 3))
    vir = np.zeros((2, 3, 3))

    def pair_sum(TA, TB):
        return float(np.sum(V * evaluate(TA, xa, ga) * evaluate(TB, xb, gb)))

    def neg(T):
        return [(-c, e) for c, e in T]

    for ia, ea in orbitals(t.la_min, t.la_max):
        for ib, eb in orbitals(t.lb_min, t.lb_max):
            A = [(1.0, ea)]
            B = [(1.0, eb)]
            # the "vab" pairs: (termsA, termsB, weight)
            if compute_tau:
                pairs = [(deriv(A, k, za), deriv(B, k, zb), 0.5) for k in range(3)]
            else:
                pairs = [(A, B, 1.0)]
            h = sum(w * pair_sum(TA, TB) for TA, TB, w in pairs)
            hab[t.o2 + ib, t.o1 + ia] += h
            p = t.pab[t.o2 + ib, t.o1 + ia]
            for i in range(3):
                forces[0, i] += p * sum(w * pair_sum(neg(deriv(TA, i, za)), TB)
                                        for TA, TB, w in pairs)
                forces[1, i] += p * sum(w * pair_sum(TA, neg(deriv(TB, i, zb)))
                                        for TA, TB, w in pairs)
                for j in range(3):
                    vir[0, i, j] += p * sum(
                        w * pair_sum(times_coord(neg(deriv(TA, j, za)), i), TB)
                        for TA, TB, w in pairs)
                    vir[1, i, j] += p * sum(
                        w * pair_sum(TA, times_coord(neg(deriv(TB, j, zb)), i))
                        for TA, TB, w in pairs)
    return hab, forces, vir


class SyntheticTask:
    """A single task with the fixture attributes used by oracle/brute."""

    def __init__(self, rng, la, lb, ortho=True, n=None, func=100, big_radius=True,
                 zrange=(2.0, 6.0), skew=0.2, h=0.3):
        self.orthorhombic = ortho
        self.border_mask = 0
        self.func = func
        self.la_min, self.la_max = la
        self.lb_min, self.lb_max = lb
        self.zeta = float(np.exp(rng.uniform(*np.log(zrange))))
        self.zetb = float(np.exp(rng.uniform(*np.log(zrange))))
        self.rscale = float(rng.choice([1.0, 2.0]))
        if n is None:
            n = (40, 36, 38) if ortho else (46, 44, 42)
        if ortho:
            self.dh = np.diag(h * rng.uniform(0.9, 1.1, 3))
        else:
            self.dh = h * (np.eye(3) + skew * rng.uniform(-1, 1, (3, 3)))
        self.dh_inv = np.linalg.inv(self.dh)
        self.npts_global = np.array(n, dtype=np.int32)
        self.npts_local = self.npts_global.copy()
        self.shift_local = np.zeros(3, dtype=np.int32)
        self.border_width = np.zeros(3, dtype=np.int32)
        cell = self.dh * self.npts_global[:, None]
        self.ra = rng.uniform(0, 1, 3) @ cell
        self.rab = rng.normal(size=3) * 0.6
        zp = self.zeta + self.zetb
        # radius where exp(-zetp r^2) ~ 1e-30 (truncation invisible); the
        # cube must still fit into the grid without wrapping
        self.radius = float(np.sqrt(69.0 / zp)) if big_radius else float(rng.uniform(1.0, 2.5))
        nca, ncb = _ncoset(self.la_max), _ncoset(self.lb_max)
        npa, npb = int(rng.integers(1, 3)), int(rng.integers(1, 3))
        self.n1, self.n2 = npa * nca, npb * ncb
        self.o1 = int(rng.integers(0, npa)) * nca
        self.o2 = int(rng.integers(0, npb)) * ncb
        self.pab = rng.normal(size=(self.n2, self.n1))
        self.compute_tau = False


def _ncoset(l):
    return 0 if l < 0 else (l + 1) * (l + 2) * (l + 3) // 6


def lpairs(max_l=3):
    """(la_min, la_max) choices."""
    return [(lo, hi) for hi in range(max_l + 1) for lo in range(hi + 1)]


__all__ = ["collocate", "integrate", "SyntheticTask", "FUNCS", "orbitals", "itertools"]
