/*
 * grid_pgf_product.h -- single-task collocate / integrate on the MI355X.
 *
 * Drop-ins for the two C functions that CP2K's Fortran module binds for one
 * product of primitive Gaussians (src/grid/grid_api.F:110-490,
 * collocate_pgf_product / integrate_pgf_product, used by GAPW, XAS and other
 * callers outside the task lists).  Same arguments, layout and meaning as
 * the reference declarations cited below; a maintainer points the
 * BIND(C, name=...) of those interfaces at these names.
 *
 * Every array argument may be host memory or device (HBM) memory; the
 * library decides per pointer.  Host arrays are copied to the device and
 * back on every call (for a host grid that is the whole local grid), so
 * callers that keep their grids in HBM avoid the PCIe traffic.  Both calls
 * return when the results are in place.  Results are accumulated (+=) into
 * grid, hab, forces, virials, hdab, hadb and a_hdab, as in the reference.
 */
#ifndef GRID_PGF_PRODUCT_H
#define GRID_PGF_PRODUCT_H

#include <stdbool.h>

#include "grid_constants.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces grid_cpu_collocate_pgf_product, src/grid/cpu/grid_cpu_collocate.h:47-55
 * (reference semantics: src/grid/ref/grid_ref_collocate.c).  pab is
 * [n2][n1] (the pgf pair's Cartesian block at offsets o1, o2); grid is the
 * local grid [npts_local[2]][npts_local[1]][npts_local[0]]. */
void grid_mi355x_collocate_pgf_product(
    const bool orthorhombic, const int border_mask, const enum grid_func func,
    const int la_max, const int la_min, const int lb_max, const int lb_min,
    const double zeta, const double zetb, const double rscale,
    const double dh[3][3], const double dh_inv[3][3], const double ra[3],
    const double rab[3], const int npts_global[3], const int npts_local[3],
    const int shift_local[3], const int border_width[3], const double radius,
    const int o1, const int o2, const int n1, const int n2, const double *pab,
    double *grid);

/* Replaces grid_cpu_integrate_pgf_product, src/grid/cpu/grid_cpu_integrate.h:30-40
 * (reference semantics: src/grid/ref/grid_ref_integrate.c:44-154).  hab and
 * pab are [n2][n1]; forces [2][3] (atom a, atom b) and virials [2][3][3]
 * are optional (NULL); hdab, hadb ([n2][n1][3]) and a_hdab ([n2][n1][3][3])
 * are the unweighted derivatives of each hab element, optional and not
 * defined with compute_tau.  Forces (or hdab) need pab; virials (or a_hdab)
 * need forces (or hdab). */
void grid_mi355x_integrate_pgf_product(
    const bool orthorhombic, const bool compute_tau, const int border_mask,
    const int la_max, const int la_min, const int lb_max, const int lb_min,
    const double zeta, const double zetb, const double dh[3][3],
    const double dh_inv[3][3], const double ra[3], const double rab[3],
    const int npts_global[3], const int npts_local[3], const int shift_local[3],
    const int border_width[3], const double radius, const int o1, const int o2,
    const int n1, const int n2, const double *grid, double *hab,
    const double *pab, double *forces, double *virials, double *hdab,
    double *hadb, double *a_hdab);

#ifdef __cplusplus
}
#endif

#endif
