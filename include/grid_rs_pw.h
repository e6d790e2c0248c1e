/*
 * grid_rs_pw.h -- real-space multigrids <-> plane-wave grids in HBM.
 *
 * The caller's steps right after grid_collocate_task_list and right before
 * grid_integrate_task_list, so that the grids stay on the device between
 * them: density_rs2pw and potential_pw2rs (src/pw_env/rs_pw_interface.F
 * :64-226, Fourier interpolation, one process) and transfer_rs2pw /
 * transfer_pw2rs (src/pw/realspace_grid_types.F:651-737, the serial case).
 * The reference implements these in Fortran over its pw_grid / rs_grid
 * types; a maintainer binds these entry points through ISO_C_BINDING with
 * the arrays of those types (rs%r, pw%array) in HBM.
 *
 * All arrays are device (HBM) pointers; every grid is [nz][ny][nx], x
 * fastest, with npts given as {nx, ny, nz}; level 0 is the finest grid (the
 * plane-wave grid of the density and potential).  G-space arrays are the
 * full FFT box [nz][ny][nx] of complex values (real, imaginary interleaved)
 * in FFT index order: index k along an axis of n points holds Miller index
 * k for k <= n - 1 - n/2, else k - n; the reference's rho_gspace is the
 * gather of this box at its pw_grid's G vectors.  Forward transforms carry
 * the factor 1/N, backward ones none (src/pw/fft_tools.F:394-397).  Calls
 * return when the results are in place.
 */
#ifndef GRID_RS_PW_H
#define GRID_RS_PW_H

#ifdef __cplusplus
extern "C" {
#endif

/* density_rs2pw, src/pw_env/rs_pw_interface.F:64-128: rho_r (finest grid)
 * and, if rho_g is not NULL, rho_g (finest G box) from the collocated level
 * grids rs_grids[0 .. nlevels).  Coarser levels must not be larger than the
 * finest along any axis. */
void grid_mi355x_density_rs2pw(const int nlevels, const int (*npts)[3],
                               const double *const *rs_grids, double *rho_r,
                               double *rho_g);

/* potential_pw2rs, src/pw_env/rs_pw_interface.F:139-226: the level grids
 * rs_grids[0 .. nlevels) from the potential v_r on the finest grid (level 0
 * a copy; coarser levels its Fourier interpolation times the volume element
 * ratio N_0 / N_l). */
void grid_mi355x_potential_pw2rs(const int nlevels, const int (*npts)[3],
                                 const double *v_r, double *const *rs_grids);

/* transfer_rs2pw, src/pw/realspace_grid_types.F:651-687 (serial): pw
 * [npts] = the real region of rs, a grid with border[d] halo points on each
 * side of axis d ([npts + 2 border]). */
void grid_mi355x_transfer_rs2pw(const int npts[3], const int border[3],
                                const double *rs, double *pw);

/* transfer_pw2rs, src/pw/realspace_grid_types.F:693-743 (serial): rs
 * [npts + 2 border] from pw [npts], the halo filled periodically. */
void grid_mi355x_transfer_pw2rs(const int npts[3], const int border[3],
                                const double *pw, double *rs);

#ifdef __cplusplus
}
#endif

#endif
