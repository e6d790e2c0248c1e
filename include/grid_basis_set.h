/*
 * grid_basis_set.h -- contracted Gaussian basis of one atomic kind.
 *
 * Same struct layout and entry points as src/grid/common/grid_basis_set.h:
 * 14-64.  The task list keeps the pointers it is given (the basis must
 * outlive it) and snapshots sphi/zet into HBM at creation.
 */
#ifndef GRID_BASIS_SET_H
#define GRID_BASIS_SET_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nset;       /* number of sets */
  int nsgf;       /* contracted spherical functions = block dimension */
  int maxco;      /* max Cartesian functions of a set (row stride of sphi) */
  int maxpgf;     /* max primitives of a set (row stride of zet) */
  int *lmin;      /* [nset] */
  int *lmax;      /* [nset] */
  int *npgf;      /* [nset] */
  int *nsgf_set;  /* [nset] */
  int *first_sgf; /* [nset], one based */
  double *sphi;   /* [nsgf][maxco] decontraction matrix */
  double *zet;    /* [nset][maxpgf] exponents */
} grid_basis_set;

#ifndef __cplusplus
/* src/grid/common/grid_basis_set.h:52-58 */
void grid_create_basis_set(const int nset, const int nsgf, const int maxco,
                           const int maxpgf, const int lmin[nset],
                           const int lmax[nset], const int npgf[nset],
                           const int nsgf_set[nset], const int first_sgf[nset],
                           const double sphi[nsgf][maxco],
                           const double zet[nset][maxpgf],
                           grid_basis_set **basis_set);
#else
void grid_create_basis_set(const int nset, const int nsgf, const int maxco,
                           const int maxpgf, const int *lmin, const int *lmax,
                           const int *npgf, const int *nsgf_set,
                           const int *first_sgf, const double *sphi,
                           const double *zet, grid_basis_set **basis_set);
#endif

/* src/grid/common/grid_basis_set.h:64 */
void grid_free_basis_set(grid_basis_set *basis_set);

#ifdef __cplusplus
}
#endif

#endif
