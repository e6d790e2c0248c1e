"""CPU restatement (numpy) of the real-space multigrid <-> plane-wave steps
around the grid calls, one process, Fourier interpolation.

TEST INFRASTRUCTURE ONLY: imported by tests/ -- never by the product package.

Grids are [nz, ny, nx] arrays (x fastest), level 0 the finest.
  density_rs2pw    src/pw_env/rs_pw_interface.F:64-128 (pw_interp branch):
                   rho_g = sum over levels of FWFFT(level) placed by Miller
                   index into the finest G box (pw_axpy through the grids'
                   reference mapping, src/pw/pw_methods.F:748-925), rho_r =
                   BWFFT(rho_g); one level: rho_r is the grid, rho_g its
                   transform.
  potential_pw2rs  :139-226: level 0 a copy of v_r, level l the G
                   coefficients of FWFFT(v_r) inside level l's box, BWFFT,
                   times dvol_l / dvol_0 = N_0 / N_l.
  transfer_rs2pw / src/pw/realspace_grid_types.F:651-737 (serial case):
  transfer_pw2rs   the real region of a halo grid, and its periodic fill.
FWFFT is exp(-i ...) with the factor 1/N, BWFFT exp(+i ...) unscaled
(src/pw/fft_tools.F:146, 394-397).  Miller indices of an axis of n points:
-(n // 2) .. -(n // 2) + n - 1 (src/pw/pw_grid_info.F:247-254).  Complex
results are taken to real by their real part (the copy from complex to real
of pw_transfer).

Pinned in tests/test_oracle.py by closed forms (band-limited trigonometric
sums sampled on every level), charge conservation, and the adjoint identity
<density_rs2pw(rs), v> = sum_l <rs_l, potential_pw2rs(v)_l>.
"""
import numpy as np


def miller(n):
    """Miller index held by each FFT index of an axis of n points."""
    k = np.arange(n)
    return np.where(k <= n - 1 - n // 2, k, k - n)


def fwfft(a):
    return np.fft.fftn(a) / a.size


def bwfft(g):
    return np.fft.ifftn(g) * g.size


def _box(fine_shape, coarse_shape):
    """Index arrays of the coarse G box inside the fine one (per axis)."""
    return np.ix_(*[miller(c) % f for c, f in zip(coarse_shape, fine_shape)])


def density_rs2pw(rs_grids):
    """Returns (rho_r, rho_g) on the finest grid."""
    rs = [np.asarray(g, dtype=np.float64) for g in rs_grids]
    if len(rs) == 1:
        return rs[0].copy(), fwfft(rs[0])
    fine = rs[0].shape
    G = np.zeros(fine, dtype=np.complex128)
    for g in rs:
        G[_box(fine, g.shape)] += fwfft(g)
    return bwfft(G).real, G


def potential_pw2rs(v_r, shapes):
    """Level grids of the given shapes (shapes[0] == v_r.shape)."""
    v = np.asarray(v_r, dtype=np.float64)
    out = [v.copy()]
    if len(shapes) > 1:
        G = fwfft(v)
        for s in shapes[1:]:
            out.append(bwfft(G[_box(v.shape, s)]).real * (v.size / np.prod(s)))
    return out


def transfer_rs2pw(rs, border):
    """border = (bx, by, bz); rs is [nz + 2 bz, ny + 2 by, nx + 2 bx]."""
    bx, by, bz = border
    nz, ny, nx = (rs.shape[0] - 2 * bz, rs.shape[1] - 2 * by, rs.shape[2] - 2 * bx)
    return rs[bz:bz + nz, by:by + ny, bx:bx + nx].copy()


def transfer_pw2rs(pw, border):
    bx, by, bz = border
    nz, ny, nx = pw.shape
    iz = (np.arange(nz + 2 * bz) - bz) % nz
    iy = (np.arange(ny + 2 * by) - by) % ny
    ix = (np.arange(nx + 2 * bx) - bx) % nx
    return pw[np.ix_(iz, iy, ix)].copy()
