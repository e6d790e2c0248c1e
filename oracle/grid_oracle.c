/*
 * grid_oracle.c -- CPU restatement of CP2K's FP64 "ref" grid backend.
 *
 * TEST INFRASTRUCTURE ONLY (see grid_oracle.h for who may load it and how
 * each part is pinned).  Build with -ffp-contract=off: the truncation-region
 * arithmetic must round exactly like the reference's.
 */
#include "grid_oracle.h"

#include <assert.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Cartesian orbital bookkeeping, src/grid/common/grid_common.h:81-163.      */
/* Within shell l the orbitals are ordered by lx descending, then lz         */
/* ascending; all shells up to l are stacked (ncoset).                       */
/* ------------------------------------------------------------------------- */

static inline int ncart(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6; }

static inline int cart_index(int lx, int ly, int lz) {
  const int l = lx + ly + lz;
  if (l == 0) return 0;
  return ncart(l - 1) + ((l - lx) * (l - lx + 1)) / 2 + lz;
}

/* Fortran MODULO: result in [0, m). */
static inline int pmod(int a, int m) { return ((a % m) + m) % m; }

typedef struct {
  int l[3];
} ang;

static inline ang raise(ang a, int i) {
  a.l[i] += 1;
  return a;
}
/* Lowering clamps at zero, as down() in grid_common.h:143-147. */
static inline ang lower(ang a, int i) {
  a.l[i] = a.l[i] > 0 ? a.l[i] - 1 : 0;
  return a;
}
static inline int aidx(ang a) { return cart_index(a.l[0], a.l[1], a.l[2]); }

/* cab is stored [idx(b)][idx(a)] with row length n1. */
typedef struct {
  double *m;
  int n1;
} cmat;

static inline void cadd(cmat *c, ang a, ang b, double v) {
  c->m[aidx(b) * c->n1 + aidx(a)] += v;
}
static inline double cget(const cmat *c, ang a, ang b) {
  return c->m[aidx(b) * c->n1 + aidx(a)];
}

/* ------------------------------------------------------------------------- */
/* grid_func stencils, src/grid/common/grid_prepare_pab.h:33-430.            */
/* Each turns one density element p = pab(a,b) into Cartesian coefficients   */
/* of products of (possibly raised/lowered) primitives.                      */
/* ------------------------------------------------------------------------- */

/* oneterm_dijdij, grid_prepare_pab.h:218-240: mixed second derivative on b. */
static void mixed_b(int d, double fa, ang a, ang b, double zb, cmat *c) {
  const int i1 = d, i2 = (d + 1) % 3;
  cadd(c, a, lower(lower(b, i2), i1), b.l[i1] * b.l[i2] * fa);
  cadd(c, a, raise(lower(b, i2), i1), -2.0 * zb * b.l[i2] * fa);
  cadd(c, a, lower(raise(b, i2), i1), -2.0 * zb * b.l[i1] * fa);
  cadd(c, a, raise(raise(b, i2), i1), +4.0 * zb * zb * fa);
}

/* oneterm_diidii, grid_prepare_pab.h:272-280: pure second derivative on b. */
static void pure_b(int d, double fa, ang a, ang b, double zb, cmat *c) {
  cadd(c, a, lower(lower(b, d), d), b.l[d] * (b.l[d] - 1) * fa);
  cadd(c, a, b, -2.0 * zb * (2 * b.l[d] + 1) * fa);
  cadd(c, a, raise(raise(b, d), d), +4.0 * zb * zb * fa);
}

static void apply_func(int func, ang a, ang b, double za, double zb, double p,
                       cmat *c) {
  if (func == 100) { /* AB, :33-38 */
    cadd(c, a, b, p);
  } else if (func == 200) { /* DADB, :44-59 */
    for (int i = 0; i < 3; i++) {
      cadd(c, lower(a, i), lower(b, i), 0.5 * a.l[i] * b.l[i] * p);
      cadd(c, lower(a, i), raise(b, i), -1.0 * a.l[i] * zb * p);
      cadd(c, raise(a, i), lower(b, i), -1.0 * za * b.l[i] * p);
      cadd(c, raise(a, i), raise(b, i), 2.0 * za * zb * p);
    }
  } else if (func >= 301 && func <= 303) { /* ADBmDAB, :65-82 */
    const int d = func - 301;
    cadd(c, a, lower(b, d), +b.l[d] * p);
    cadd(c, a, raise(b, d), -2.0 * zb * p);
    cadd(c, lower(a, d), b, -a.l[d] * p);
    cadd(c, raise(a, d), b, +2.0 * za * p);
  } else if (func >= 411 && func <= 433) { /* ARDBmDARB, :88-106 */
    const int d = (func / 10) % 10 - 1, r = func % 10 - 1;
    const ang br = raise(b, r);
    cadd(c, a, lower(br, d), b.l[d] * p);
    cadd(c, a, raise(br, d), -2.0 * zb * p);
    cadd(c, lower(a, d), br, -a.l[d] * p);
    cadd(c, raise(a, d), br, +2.0 * za * p);
  } else if (func >= 501 && func <= 503) { /* DABpADB, :112-129 */
    const int d = func - 501;
    cadd(c, a, lower(b, d), b.l[d] * p);
    cadd(c, a, raise(b, d), -2.0 * zb * p);
    cadd(c, lower(a, d), b, a.l[d] * p);
    cadd(c, raise(a, d), b, -2.0 * za * p);
  } else if (func >= 601 && func <= 603) { /* DX,DY,DZ, :195-212 */
    const int d = func - 601;
    cadd(c, lower(a, d), lower(b, d), a.l[d] * b.l[d] * p);
    cadd(c, lower(a, d), raise(b, d), -2.0 * a.l[d] * zb * p);
    cadd(c, raise(a, d), lower(b, d), -2.0 * za * b.l[d] * p);
    cadd(c, raise(a, d), raise(b, d), +4.0 * za * zb * p);
  } else if (func >= 701 && func <= 703) { /* DXDY,DYDZ,DZDX, :246-266 */
    const int d1 = func - 701, d2 = (d1 + 1) % 3;
    mixed_b(d1, a.l[d1] * a.l[d2] * p, lower(lower(a, d2), d1), b, zb, c);
    mixed_b(d1, -2.0 * za * a.l[d2] * p, raise(lower(a, d2), d1), b, zb, c);
    mixed_b(d1, -2.0 * za * a.l[d1] * p, lower(raise(a, d2), d1), b, zb, c);
    mixed_b(d1, 4.0 * za * za * p, raise(raise(a, d2), d1), b, zb, c);
  } else if (func >= 801 && func <= 803) { /* DXDX,DYDY,DZDZ, :286-303 */
    const int d = func - 801;
    pure_b(d, a.l[d] * (a.l[d] - 1) * p, lower(lower(a, d), d), b, zb, c);
    pure_b(d, -2.0 * za * (2 * a.l[d] + 1) * p, a, b, zb, c);
    pure_b(d, 4.0 * za * za * p, raise(raise(a, d), d), b, zb, c);
  } else if (func >= 901 && func <= 903) { /* DAB, :138-149 */
    const int d = func - 901;
    cadd(c, lower(a, d), b, -a.l[d] * p);
    cadd(c, raise(a, d), b, +2.0 * za * p);
  } else if (func >= 904 && func <= 906) { /* ADB, :158-170 */
    const int d = func - 904;
    cadd(c, a, lower(b, d), -b.l[d] * p);
    cadd(c, a, raise(b, d), +2.0 * zb * p);
  } else if (func >= 1001 && func <= 1003) { /* CORE, :177-189 (zero terms omitted) */
    const int d = func - 1001;
    cadd(c, raise(a, d), b, +2.0 * za * p);
  } else {
    abort(); /* unknown function; callers validate func first */
  }
}

/* l-range widening per func, src/grid/common/grid_prepare_pab.h:447-517.
 * Returns 0 for an unknown func. d = {la_max, la_min, lb_max, lb_min}. */
static int func_ldiffs(int func, int d[4]) {
  int a = 0, b = 0;
  if (func == 100) {
    a = 0; b = 0;
  } else if (func == 200 || (func >= 301 && func <= 303) ||
             (func >= 501 && func <= 503) || (func >= 601 && func <= 603) ||
             (func >= 901 && func <= 906) || (func >= 1001 && func <= 1003)) {
    a = 1; b = 1;
  } else if (func >= 411 && func <= 433 && func % 10 >= 1 && func % 10 <= 3 &&
             (func / 10) % 10 >= 1 && (func / 10) % 10 <= 3) {
    d[0] = 1; d[1] = -1; d[2] = 2; d[3] = -1;
    return 1;
  } else if ((func >= 701 && func <= 703) || (func >= 801 && func <= 803)) {
    a = 2; b = 2;
  } else {
    return 0;
  }
  d[0] = a; d[1] = -a; d[2] = b; d[3] = -b;
  return 1;
}

/* ------------------------------------------------------------------------- */
/* (x-a)^i (x-b)^j  ->  sum_l alpha_l (x-p)^l, per direction;                 */
/* src/grid/ref/grid_ref_collint.h:836-861.                                  */
/* alpha is [3][lb_max+1][la_max+1][lp+1].                                   */
/* ------------------------------------------------------------------------- */
static void expansion_table(int la_max, int lb_max, const double ra[3],
                            const double rb[3], const double rp[3],
                            double *alpha) {
  const int lp = la_max + lb_max;
  const int na = la_max + 1, nb = lb_max + 1, np = lp + 1;
  memset(alpha, 0, sizeof(double) * 3 * na * nb * np);
  for (int d = 0; d < 3; d++) {
    const double pa = rp[d] - ra[d];
    const double pb = rp[d] - rb[d];
    for (int ia = 0; ia <= la_max; ia++) {
      for (int ib = 0; ib <= lb_max; ib++) {
        double *row = &alpha[((d * nb + ib) * na + ia) * np];
        double bin_a = 1.0, pow_a = 1.0;
        for (int k = 0; k <= ia; k++) {
          double bin_b = 1.0, pow_b = 1.0;
          for (int l = 0; l <= ib; l++) {
            row[ia - k + ib - l] += bin_a * bin_b * pow_a * pow_b;
            bin_b *= ((double)(ib - l)) / ((double)(l + 1));
            pow_b *= pb;
          }
          bin_a *= ((double)(ia - k)) / ((double)(k + 1));
          pow_a *= pa;
        }
      }
    }
  }
}

/* cab <-> cxyz, src/grid/ref/grid_ref_collint.h:827-911.
 * cxyz is [lp+1]^3 indexed [lz][ly][lx]; cab is [ncart(lb_max)][ncart(la_max)].
 * forward: cxyz += T cab; adjoint: cab += T^T cxyz. */
static void cab_cxyz(bool forward, int la_max, int la_min, int lb_max,
                     int lb_min, double prefactor, const double ra[3],
                     const double rb[3], const double rp[3], double *cab,
                     double *cxyz) {
  const int lp = la_max + lb_max;
  const int na = la_max + 1, nb = lb_max + 1, np = lp + 1;
  double *alpha = malloc(sizeof(double) * 3 * na * nb * np);
  expansion_table(la_max, lb_max, ra, rb, rp, alpha);
  const int n1 = ncart(la_max);
#define ALPHA(d, ib, ia) (&alpha[(((d) * nb + (ib)) * na + (ia)) * np])
  for (int lza = 0; lza <= la_max; lza++)
    for (int lya = 0; lya <= la_max - lza; lya++)
      for (int lxa = (la_min - lza - lya > 0 ? la_min - lza - lya : 0);
           lxa <= la_max - lza - lya; lxa++) {
        const int ico = cart_index(lxa, lya, lza);
        for (int lzb = 0; lzb <= lb_max; lzb++)
          for (int lyb = 0; lyb <= lb_max - lzb; lyb++)
            for (int lxb = (lb_min - lzb - lyb > 0 ? lb_min - lzb - lyb : 0);
                 lxb <= lb_max - lzb - lyb; lxb++) {
              const int jco = cart_index(lxb, lyb, lzb);
              double *cv = &cab[jco * n1 + ico];
              const double *ax = ALPHA(0, lxb, lxa);
              const double *ay = ALPHA(1, lyb, lya);
              const double *az = ALPHA(2, lzb, lza);
              for (int lzp = 0; lzp <= lza + lzb; lzp++)
                for (int lyp = 0; lyp <= lya + lyb; lyp++)
                  for (int lxp = 0; lxp <= lxa + lxb; lxp++) {
                    const double t = ax[lxp] * ay[lyp] * az[lzp] * prefactor;
                    double *xv = &cxyz[(lzp * np + lyp) * np + lxp];
                    if (forward)
                      *xv += *cv * t;
                    else
                      *cv += *xv * t;
                  }
            }
      }
#undef ALPHA
  free(alpha);
}

/* ------------------------------------------------------------------------- */
/* Orthorhombic kernel, src/grid/ref/grid_ref_collint.h:30-327.             */
/* The cube is walked in mirrored pairs of planes (k, 1-k) and rows (j, 1-j) */
/* for k, j <= 0; the reference's pair distance (2k-1)/2 equals k under C    */
/* integer division for k <= 0.  Row bounds use the discretised radius.      */
/* ------------------------------------------------------------------------- */
typedef struct {
  int center[3];
  double roff[3];
  double disr;
  int lb[3]; /* ub = 1 - lb */
  int cmax;
} ortho_geom;

static void ortho_geometry(const double dh[3][3], const double dh_inv[3][3],
                           const double rp[3], const int ng[3], const int nl[3],
                           const int shift[3], double radius, ortho_geom *g) {
  for (int d = 0; d < 3; d++) {
    double s = 0.0;
    for (int e = 0; e < 3; e++) s += dh_inv[e][d] * rp[e];
    g->center[d] = (int)floor(s);
  }
  for (int d = 0; d < 3; d++)
    g->roff[d] = rp[d] - ((double)g->center[d]) * dh[d][d];
  const double drmin = fmin(dh[0][0], fmin(dh[1][1], dh[2][2]));
  g->disr = drmin * fmax(1.0, ceil(radius / drmin));
  g->cmax = 0;
  for (int d = 0; d < 3; d++) {
    g->lb[d] = (int)ceil(-1e-8 - g->disr * dh_inv[d][d]);
    const int ub = 1 - g->lb[d];
    if (ng[d] != nl[d]) { /* non-periodic direction: cube must not wrap */
      const int off = pmod(g->center[d] + g->lb[d] - shift[d], ng[d]) - g->lb[d];
      assert(off + ub < nl[d]);
      assert(off + g->lb[d] >= 0);
      (void)off; (void)ub;
    }
    if (ub > g->cmax) g->cmax = ub;
  }
}

static inline int ortho_row_start(double remain, double hinv) {
  return (int)ceil(-1e-8 - sqrt(fmax(0.0, remain)) * hinv);
}

/* Walks the ortho truncation region; mode 0 collocate, 1 integrate,
 * 2 count only. */
static long ortho_pass(int mode, int lp, double zetp, const double dh[3][3],
                       const double dh_inv[3][3], const double rp[3],
                       const int ng[3], const int nl[3], const int shift[3],
                       double radius, double *cxyz, double *grid) {
  ortho_geom g;
  ortho_geometry(dh, dh_inv, rp, ng, nl, shift, radius, &g);
  const int cmax = g.cmax, ext = 2 * cmax + 1, np = lp + 1;
  double *pol = NULL;
  int *map = malloc(sizeof(int) * 3 * ext);
  for (int d = 0; d < 3; d++)
    for (int c = -cmax; c <= cmax; c++)
      map[d * ext + c + cmax] = pmod(g.center[d] + c - shift[d], ng[d]);
  if (mode != 2) {
    pol = malloc(sizeof(double) * 3 * np * ext);
    for (int d = 0; d < 3; d++) {
      const double h = dh[d][d];
      for (int c = g.lb[d]; c <= 1 - g.lb[d]; c++) {
        const double t = c * h - g.roff[d];
        double v = exp(-zetp * t * t);
        for (int l = 0; l <= lp; l++) {
          pol[(d * np + l) * ext + c + cmax] = v;
          v *= t;
        }
      }
    }
  }
#define POL(d, l, c) pol[((d) * np + (l)) * ext + (c) + cmax]
  double *cxy = malloc(sizeof(double) * np * np);
  double *cx = malloc(sizeof(double) * np);
  const long sxy = (long)nl[0] * nl[1];
  long count = 0;
  const double r2 = g.disr * g.disr;
  for (int k1 = g.lb[2]; k1 <= 0; k1++) {
    const double kr = k1 * dh[2][2];
    const double kremain = r2 - kr * kr;
    const int jstart = ortho_row_start(kremain, dh_inv[1][1]);
    for (int kk = 0; kk < 2; kk++) {
      const int k = kk == 0 ? k1 : 1 - k1;
      const long kg = map[2 * ext + k + cmax];
      if (mode == 0) {
        memset(cxy, 0, sizeof(double) * np * np);
        for (int lz = 0; lz <= lp; lz++)
          for (int ly = 0; ly <= lp - lz; ly++)
            for (int lx = 0; lx <= lp - lz - ly; lx++)
              cxy[ly * np + lx] += cxyz[(lz * np + ly) * np + lx] * POL(2, lz, k);
      } else if (mode == 1) {
        memset(cxy, 0, sizeof(double) * np * np);
      }
      for (int j1 = jstart; j1 <= 0; j1++) {
        const double jr = j1 * dh[1][1];
        const double jremain = kremain - jr * jr;
        const int istart = ortho_row_start(jremain, dh_inv[0][0]);
        const int iend = 1 - istart;
        for (int jj = 0; jj < 2; jj++) {
          const int j = jj == 0 ? j1 : 1 - j1;
          const long base = kg * sxy + (long)map[1 * ext + j + cmax] * nl[0];
          if (mode == 2) {
            count += iend - istart + 1;
            continue;
          }
          if (mode == 0) {
            for (int lx = 0; lx <= lp; lx++) cx[lx] = 0.0;
            for (int ly = 0; ly <= lp; ly++)
              for (int lx = 0; lx <= lp - ly; lx++)
                cx[lx] += cxy[ly * np + lx] * POL(1, ly, j);
            for (int i = istart; i <= iend; i++) {
              double v = 0.0;
              for (int lx = 0; lx <= lp; lx++) v += cx[lx] * POL(0, lx, i);
              grid[base + map[i + cmax]] += v;
            }
          } else {
            for (int lx = 0; lx <= lp; lx++) cx[lx] = 0.0;
            for (int i = istart; i <= iend; i++) {
              const double gv = grid[base + map[i + cmax]];
              for (int lx = 0; lx <= lp; lx++) cx[lx] += gv * POL(0, lx, i);
            }
            for (int ly = 0; ly <= lp; ly++)
              for (int lx = 0; lx <= lp - ly; lx++)
                cxy[ly * np + lx] += cx[lx] * POL(1, ly, j);
          }
        }
      }
      if (mode == 1) {
        for (int lz = 0; lz <= lp; lz++)
          for (int ly = 0; ly <= lp - lz; ly++)
            for (int lx = 0; lx <= lp - lz - ly; lx++)
              cxyz[(lz * np + ly) * np + lx] += cxy[ly * np + lx] * POL(2, lz, k);
      }
    }
  }
#undef POL
  free(cx);
  free(cxy);
  free(map);
  free(pol);
  return count;
}

/* ------------------------------------------------------------------------- */
/* General (non-orthorhombic or border-masked) kernel,                       */
/* src/grid/ref/grid_ref_collint.h:333-794.  The polynomial is re-expanded   */
/* in lattice offsets (i-gp0, j-gp1, k-gp2); the enclosing index box comes   */
/* from the 27 corner points of the cube rp +- radius; each row's i-range    */
/* solves |r|^2 <= radius^2.  The Gaussian is exp(-zetp |r|^2) evaluated     */
/* directly (the reference factors it into three 2-D tables).               */
/* ------------------------------------------------------------------------- */

static double factorial(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; i++) f *= i;
  return f;
}

/* cxyz <-> cijk multinomial change of variables, grid_ref_collint.h:697-762.
 * forward: cijk += M cxyz; adjoint: cxyz += M^T cijk. */
static void cxyz_cijk(bool forward, int lp, const double dh[3][3], double *cxyz,
                      double *cijk) {
  const int np = lp + 1;
  /* hp[k][a][b] = dh[a][b]^k */
  double *hp = malloc(sizeof(double) * np * 9);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      hp[0 * 9 + a * 3 + b] = 1.0;
      for (int k = 1; k <= lp; k++) hp[k * 9 + a * 3 + b] = hp[(k - 1) * 9 + a * 3 + b] * dh[a][b];
    }
#define HP(k, a, b) hp[(k) * 9 + (a) * 3 + (b)]
  for (int klx = 0; klx <= lp; klx++)
    for (int jlx = 0; jlx <= lp - klx; jlx++)
      for (int ilx = 0; ilx <= lp - klx - jlx; ilx++) {
        const int lx = ilx + jlx + klx;
        for (int kly = 0; kly <= lp - lx; kly++)
          for (int jly = 0; jly <= lp - lx - kly; jly++)
            for (int ily = 0; ily <= lp - lx - kly - jly; ily++) {
              const int ly = ily + jly + kly;
              for (int klz = 0; klz <= lp - lx - ly; klz++)
                for (int jlz = 0; jlz <= lp - lx - ly - klz; jlz++)
                  for (int ilz = 0; ilz <= lp - lx - ly - klz - jlz; ilz++) {
                    const int lz = ilz + jlz + klz;
                    const int il = ilx + ily + ilz, jl = jlx + jly + jlz,
                              kl = klx + kly + klz;
                    const double m =
                        HP(ilx, 0, 0) * HP(jlx, 1, 0) * HP(klx, 2, 0) *
                        HP(ily, 0, 1) * HP(jly, 1, 1) * HP(kly, 2, 1) *
                        HP(ilz, 0, 2) * HP(jlz, 1, 2) * HP(klz, 2, 2) *
                        factorial(lx) * factorial(ly) * factorial(lz) /
                        (factorial(ilx) * factorial(ily) * factorial(ilz) *
                         factorial(jlx) * factorial(jly) * factorial(jlz) *
                         factorial(klx) * factorial(kly) * factorial(klz));
                    double *ijk = &cijk[(kl * np + jl) * np + il];
                    double *xyz = &cxyz[(lz * np + ly) * np + lx];
                    if (forward)
                      *ijk += *xyz * m;
                    else
                      *xyz += *ijk * m;
                  }
            }
      }
#undef HP
  free(hp);
}

typedef struct {
  double gp[3];
  int imin[3], imax[3];
  int bounds[3][2];
} general_geom;

static void general_geometry(int border_mask, const double dh_inv[3][3],
                             const double rp[3], const int nl[3],
                             const int border_width[3], double radius,
                             general_geom *g) {
  for (int d = 0; d < 3; d++) {
    g->bounds[d][0] = 0;
    g->bounds[d][1] = nl[d] - 1;
    if (border_mask & (1 << (2 * d))) g->bounds[d][0] += border_width[d];
    if (border_mask & (1 << (2 * d + 1))) g->bounds[d][1] -= border_width[d];
  }
  for (int d = 0; d < 3; d++) {
    g->gp[d] = 0.0;
    for (int e = 0; e < 3; e++) g->gp[d] += dh_inv[e][d] * rp[e];
  }
  for (int d = 0; d < 3; d++) {
    g->imin[d] = INT_MAX;
    g->imax[d] = INT_MIN;
  }
  for (int si = -1; si <= 1; si++)
    for (int sj = -1; sj <= 1; sj++)
      for (int sk = -1; sk <= 1; sk++) {
        const double x = rp[0] + si * radius;
        const double y = rp[1] + sj * radius;
        const double z = rp[2] + sk * radius;
        for (int d = 0; d < 3; d++) {
          const double t = dh_inv[0][d] * x + dh_inv[1][d] * y + dh_inv[2][d] * z;
          const int lo = (int)floor(t), hi = (int)ceil(t);
          if (lo < g->imin[d]) g->imin[d] = lo;
          if (hi > g->imax[d]) g->imax[d] = hi;
        }
      }
}

/* Row i-range from the quadratic in i, grid_ref_collint.h:450-464.
 * Returns 0 when the row misses the sphere. */
static int general_row(const double dh[3][3], const double gp[3], int j, int k,
                       double radius, int *ismin, int *ismax) {
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = 0; i < 3; i++) {
    const double v = (0 - gp[0]) * dh[0][i] + (j - gp[1]) * dh[1][i] + (k - gp[2]) * dh[2][i];
    a += dh[0][i] * dh[0][i];
    b += 2.0 * v * dh[0][i];
    c += v * v;
  }
  const double d = b * b - 4.0 * a * (c - radius * radius);
  if (!(0.0 < d)) return 0;
  const double sd = sqrt(d);
  const double inv_2a = 1.0 / (2.0 * a);
  *ismin = (int)ceil((-b - sd) * inv_2a);
  *ismax = (int)floor((-b + sd) * inv_2a);
  return 1;
}

static long general_pass(int mode, int border_mask, int lp, double zetp,
                         const double dh[3][3], const double dh_inv[3][3],
                         const double rp[3], const int ng[3], const int nl[3],
                         const int shift[3], const int border_width[3],
                         double radius, double *cxyz, double *grid) {
  general_geom g;
  general_geometry(border_mask, dh_inv, rp, nl, border_width, radius, &g);
  const int np = lp + 1;
  int range[3];
  int *map[3];
  for (int d = 0; d < 3; d++) {
    range[d] = g.imax[d] - g.imin[d] + 1;
    map[d] = malloc(sizeof(int) * range[d]);
    for (int t = g.imin[d]; t <= g.imax[d]; t++) {
      const int tg = pmod(t - shift[d], ng[d]);
      map[d][t - g.imin[d]] =
          (g.bounds[d][0] <= tg && tg <= g.bounds[d][1]) ? tg : INT_MIN;
    }
  }
  double *cijk = NULL, *cij = NULL, *ci = NULL;
  if (mode != 2) {
    cijk = calloc((size_t)np * np * np, sizeof(double));
    cij = malloc(sizeof(double) * np * np);
    ci = malloc(sizeof(double) * np);
    if (mode == 0) cxyz_cijk(true, lp, dh, cxyz, cijk);
  }
  const long sxy = (long)nl[0] * nl[1];
  long count = 0;
  for (int k = g.imin[2]; k <= g.imax[2]; k++) {
    const int kg = map[2][k - g.imin[2]];
    if (kg < g.bounds[2][0] || g.bounds[2][1] < kg) continue;
    const double dk = k - g.gp[2];
    if (mode == 0) {
      memset(cij, 0, sizeof(double) * np * np);
      double p = 1.0;
      for (int kl = 0; kl <= lp; kl++) {
        for (int jl = 0; jl <= lp - kl; jl++)
          for (int il = 0; il <= lp - kl - jl; il++)
            cij[jl * np + il] += cijk[(kl * np + jl) * np + il] * p;
        p *= dk;
      }
    } else if (mode == 1) {
      memset(cij, 0, sizeof(double) * np * np);
    }
    for (int j = g.imin[1]; j <= g.imax[1]; j++) {
      const int jg = map[1][j - g.imin[1]];
      if (jg < 0) continue;
      const double dj = j - g.gp[1];
      int ismin, ismax;
      if (!general_row(dh, g.gp, j, k, radius, &ismin, &ismax)) continue;
      /* The sphere lies inside the index box; clamp guards rounding. */
      if (ismin < g.imin[0]) ismin = g.imin[0];
      if (ismax > g.imax[0]) ismax = g.imax[0];
      const long base = (long)kg * sxy + (long)jg * nl[0];
      if (mode == 2) {
        for (int i = ismin; i <= ismax; i++)
          if (map[0][i - g.imin[0]] >= 0) count++;
        continue;
      }
      if (mode == 0) {
        for (int il = 0; il <= lp; il++) ci[il] = 0.0;
        double p = 1.0;
        for (int jl = 0; jl <= lp; jl++) {
          for (int il = 0; il <= lp - jl; il++) ci[il] += cij[jl * np + il] * p;
          p *= dj;
        }
      } else {
        for (int il = 0; il <= lp; il++) ci[il] = 0.0;
      }
      for (int i = ismin; i <= ismax; i++) {
        const int ig = map[0][i - g.imin[0]];
        if (ig < 0) continue;
        const double di = i - g.gp[0];
        const double rx = di * dh[0][0] + dj * dh[1][0] + dk * dh[2][0];
        const double ry = di * dh[0][1] + dj * dh[1][1] + dk * dh[2][1];
        const double rz = di * dh[0][2] + dj * dh[1][2] + dk * dh[2][2];
        const double gauss = exp(-zetp * (rx * rx + ry * ry + rz * rz));
        double dip = gauss;
        if (mode == 0) {
          double v = 0.0;
          for (int il = 0; il <= lp; il++) {
            v += ci[il] * dip;
            dip *= di;
          }
          grid[base + ig] += v;
        } else {
          const double gv = grid[base + ig];
          for (int il = 0; il <= lp; il++) {
            ci[il] += gv * dip;
            dip *= di;
          }
        }
      }
      if (mode == 1) {
        double p = 1.0;
        for (int jl = 0; jl <= lp; jl++) {
          for (int il = 0; il <= lp - jl; il++) cij[jl * np + il] += ci[il] * p;
          p *= dj;
        }
      }
    }
    if (mode == 1) {
      double p = 1.0;
      for (int kl = 0; kl <= lp; kl++) {
        for (int jl = 0; jl <= lp - kl; jl++)
          for (int il = 0; il <= lp - kl - jl; il++)
            cijk[(kl * np + jl) * np + il] += cij[jl * np + il] * p;
        p *= dk;
      }
    }
  }
  if (mode == 1) cxyz_cijk(false, lp, dh, cxyz, cijk);
  for (int d = 0; d < 3; d++) free(map[d]);
  free(cijk);
  free(cij);
  free(ci);
  return count;
}

/* ------------------------------------------------------------------------- */
/* cab_to_grid, src/grid/ref/grid_ref_collint.h:917-965.                     */
/* ------------------------------------------------------------------------- */
static bool too_small(double radius, const double dh[3][3]) {
  double hmax = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) hmax = fmax(hmax, fabs(dh[i][j]));
  return 2.0 * radius < hmax;
}

static long cab_grid(int mode, bool orthorhombic, int border_mask, int la_max,
                     int la_min, int lb_max, int lb_min, double zeta,
                     double zetb, double rscale, const double dh[3][3],
                     const double dh_inv[3][3], const double ra[3],
                     const double rab[3], const int ng[3], const int nl[3],
                     const int shift[3], const int border_width[3],
                     double radius, double *cab, double *grid) {
  if (too_small(radius, dh)) return 0;
  const double zetp = zeta + zetb;
  const double f = zetb / zetp;
  const double rab2 = rab[0] * rab[0] + rab[1] * rab[1] + rab[2] * rab[2];
  const double prefactor = rscale * exp(-zeta * f * rab2);
  double rp[3], rb[3];
  for (int i = 0; i < 3; i++) {
    rp[i] = ra[i] + f * rab[i];
    rb[i] = ra[i] + rab[i];
  }
  const int lp = la_max + lb_max;
  const int np = lp + 1;
  double *cxyz = NULL;
  if (mode != 2) cxyz = calloc((size_t)np * np * np, sizeof(double));
  if (mode == 0)
    cab_cxyz(true, la_max, la_min, lb_max, lb_min, prefactor, ra, rb, rp, cab, cxyz);
  long n;
  if (orthorhombic && border_mask == 0)
    n = ortho_pass(mode, lp, zetp, dh, dh_inv, rp, ng, nl, shift, radius, cxyz, grid);
  else
    n = general_pass(mode, border_mask, lp, zetp, dh, dh_inv, rp, ng, nl, shift,
                     border_width, radius, cxyz, grid);
  if (mode == 1)
    cab_cxyz(false, la_max, la_min, lb_max, lb_min, prefactor, ra, rb, rp, cab, cxyz);
  free(cxyz);
  return n;
}

long oracle_task_points(bool orthorhombic, int border_mask, int lp, double zeta,
                        double zetb, const double dh[3][3],
                        const double dh_inv[3][3], const double ra[3],
                        const double rab[3], const int npts_global[3],
                        const int npts_local[3], const int shift_local[3],
                        const int border_width[3], double radius) {
  (void)lp;
  return cab_grid(2, orthorhombic, border_mask, 0, 0, 0, 0, zeta, zetb, 1.0, dh,
                  dh_inv, ra, rab, npts_global, npts_local, shift_local,
                  border_width, radius, NULL, NULL);
}

/* src/grid/ref/grid_ref_collocate.c:19-50 */
void oracle_collocate_pgf_product(
    bool orthorhombic, int border_mask, int func, int la_max, int la_min,
    int lb_max, int lb_min, double zeta, double zetb, double rscale,
    const double dh[3][3], const double dh_inv[3][3], const double ra[3],
    const double rab[3], const int npts_global[3], const int npts_local[3],
    const int shift_local[3], const int border_width[3], double radius, int o1,
    int o2, int n1, int n2, const double *pab, double *grid) {
  int ld[4];
  if (!func_ldiffs(func, ld)) abort();
  const int la_max_c = la_max + ld[0], lb_max_c = lb_max + ld[2];
  const int la_min_c = la_min + ld[1] > 0 ? la_min + ld[1] : 0;
  const int lb_min_c = lb_min + ld[3] > 0 ? lb_min + ld[3] : 0;
  const int m1 = ncart(la_max_c), m2 = ncart(lb_max_c);
  double *cab = calloc((size_t)m1 * m2, sizeof(double));
  cmat c = {cab, m1};
  /* grid_ref_prepare_pab.c:54-82: loop over the requested shells of a, b */
  for (int lxa = 0; lxa <= la_max; lxa++)
    for (int lxb = 0; lxb <= lb_max; lxb++)
      for (int lya = 0; lya <= la_max - lxa; lya++)
        for (int lyb = 0; lyb <= lb_max - lxb; lyb++)
          for (int lza = (la_min - lxa - lya > 0 ? la_min - lxa - lya : 0);
               lza <= la_max - lxa - lya; lza++)
            for (int lzb = (lb_min - lxb - lyb > 0 ? lb_min - lxb - lyb : 0);
                 lzb <= lb_max - lxb - lyb; lzb++) {
              const ang a = {{lxa, lya, lza}}, b = {{lxb, lyb, lzb}};
              const double p = pab[(o2 + aidx(b)) * n1 + o1 + aidx(a)];
              apply_func(func, a, b, zeta, zetb, p, &c);
            }
  (void)n2;
  cab_grid(0, orthorhombic, border_mask, la_max_c, la_min_c, lb_max_c, lb_min_c,
           zeta, zetb, rscale, dh, dh_inv, ra, rab, npts_global, npts_local,
           shift_local, border_width, radius, cab, grid);
  free(cab);
}

/* ------------------------------------------------------------------------- */
/* hab / force / virial extraction, src/grid/common/grid_process_vab.h.      */
/* ------------------------------------------------------------------------- */
static double fa_plain(ang a, ang b, int i, double za, const cmat *c) {
  return 2.0 * za * cget(c, raise(a, i), b) - a.l[i] * cget(c, lower(a, i), b);
}
static double fb_plain(ang a, ang b, int i, double zb, const double rab[3],
                       const cmat *c) {
  return 2.0 * zb * (cget(c, raise(a, i), b) - rab[i] * cget(c, a, b)) -
         b.l[i] * cget(c, a, lower(b, i));
}
static double va_plain(ang a, ang b, int i, int j, double za, const cmat *c) {
  return 2.0 * za * cget(c, raise(raise(a, j), i), b) -
         a.l[j] * cget(c, raise(lower(a, j), i), b);
}
static double vb_plain(ang a, ang b, int i, int j, double zb,
                       const double rab[3], const cmat *c) {
  return 2.0 * zb *
             (cget(c, raise(raise(a, j), i), b) - cget(c, raise(a, i), b) * rab[j] -
              cget(c, raise(a, j), b) * rab[i] + cget(c, a, b) * rab[j] * rab[i]) -
         b.l[j] * cget(c, a, raise(lower(b, j), i));
}

/* The tau (<nabla a|V|nabla b>) forms expand each quantity over the four
 * derivative terms of grid_process_vab.h:41-205. */
#define TAU_SUM(EXPR_DD, EXPR_UD, EXPR_DU, EXPR_UU)                            \
  do {                                                                         \
    for (int k = 0; k < 3; k++) {                                              \
      const ang ad = lower(a, k), au = raise(a, k);                            \
      const ang bd = lower(b, k), bu = raise(b, k);                            \
      s += 0.5 * a.l[k] * b.l[k] * (EXPR_DD);                                  \
      s -= za * b.l[k] * (EXPR_UD);                                            \
      s -= a.l[k] * zb * (EXPR_DU);                                            \
      s += 2.0 * za * zb * (EXPR_UU);                                          \
    }                                                                          \
  } while (0)

static double get_hab(ang a, ang b, double za, double zb, const cmat *c, bool tau) {
  if (!tau) return cget(c, a, b);
  double s = 0.0;
  TAU_SUM(cget(c, ad, bd), cget(c, au, bd), cget(c, ad, bu), cget(c, au, bu));
  return s;
}
static double get_fa(ang a, ang b, int i, double za, double zb, const cmat *c, bool tau) {
  if (!tau) return fa_plain(a, b, i, za, c);
  double s = 0.0;
  TAU_SUM(fa_plain(ad, bd, i, za, c), fa_plain(au, bd, i, za, c),
          fa_plain(ad, bu, i, za, c), fa_plain(au, bu, i, za, c));
  return s;
}
static double get_fb(ang a, ang b, int i, double za, double zb, const double rab[3],
                     const cmat *c, bool tau) {
  if (!tau) return fb_plain(a, b, i, zb, rab, c);
  double s = 0.0;
  TAU_SUM(fb_plain(ad, bd, i, zb, rab, c), fb_plain(au, bd, i, zb, rab, c),
          fb_plain(ad, bu, i, zb, rab, c), fb_plain(au, bu, i, zb, rab, c));
  return s;
}
static double get_va(ang a, ang b, int i, int j, double za, double zb, const cmat *c,
                     bool tau) {
  if (!tau) return va_plain(a, b, i, j, za, c);
  double s = 0.0;
  TAU_SUM(va_plain(ad, bd, i, j, za, c), va_plain(au, bd, i, j, za, c),
          va_plain(ad, bu, i, j, za, c), va_plain(au, bu, i, j, za, c));
  return s;
}
static double get_vb(ang a, ang b, int i, int j, double za, double zb,
                     const double rab[3], const cmat *c, bool tau) {
  if (!tau) return vb_plain(a, b, i, j, zb, rab, c);
  double s = 0.0;
  TAU_SUM(vb_plain(ad, bd, i, j, zb, rab, c), vb_plain(au, bd, i, j, zb, rab, c),
          vb_plain(ad, bu, i, j, zb, rab, c), vb_plain(au, bu, i, j, zb, rab, c));
  return s;
}
#undef TAU_SUM

/* src/grid/ref/grid_ref_integrate.c:95-203 (+ process_get_ldiffs,
 * grid_process_vab.h:222-251) */
void oracle_integrate_pgf_product(
    bool orthorhombic, bool compute_tau, int border_mask, int la_max,
    int la_min, int lb_max, int lb_min, double zeta, double zetb,
    const double dh[3][3], const double dh_inv[3][3], const double ra[3],
    const double rab[3], const int npts_global[3], const int npts_local[3],
    const int shift_local[3], const int border_width[3], double radius, int o1,
    int o2, int n1, int n2, const double *grid, double *hab, const double *pab,
    double *forces, double *virials) {
  oracle_integrate_pgf_product_full(
      orthorhombic, compute_tau, border_mask, la_max, la_min, lb_max, lb_min, zeta, zetb, dh,
      dh_inv, ra, rab, npts_global, npts_local, shift_local, border_width, radius, o1, o2, n1,
      n2, grid, hab, pab, forces, virials, NULL, NULL, NULL);
}

void oracle_integrate_pgf_product_full(
    bool orthorhombic, bool compute_tau, int border_mask, int la_max,
    int la_min, int lb_max, int lb_min, double zeta, double zetb,
    const double dh[3][3], const double dh_inv[3][3], const double ra[3],
    const double rab[3], const int npts_global[3], const int npts_local[3],
    const int shift_local[3], const int border_width[3], double radius, int o1,
    int o2, int n1, int n2, const double *grid, double *hab, const double *pab,
    double *forces, double *virials, double *hdab, double *hadb, double *a_hdab) {
  (void)n2;
  /* grid_ref_integrate.c:56-60: the derivative outputs widen l like forces
   * (hdab, hadb) and the virial (a_hdab) */
  const bool do_f = forces != NULL || hdab != NULL || hadb != NULL;
  const bool do_v = virials != NULL || a_hdab != NULL;
  assert(!do_v || do_f);
  assert(!(hdab || hadb || a_hdab) || !compute_tau);
  int dla_max = 0, dla_min = 0, dlb_max = 0, dlb_min = 0;
  if (do_f || do_v) { dla_max += 1; dla_min -= 1; dlb_min -= 1; }
  if (do_v) { dla_max += 1; dlb_max += 1; }
  if (compute_tau) { dla_max += 1; dlb_max += 1; dla_min -= 1; dlb_min -= 1; }
  const int la_max_l = la_max + dla_max, lb_max_l = lb_max + dlb_max;
  const int la_min_l = la_min + dla_min > 0 ? la_min + dla_min : 0;
  const int lb_min_l = lb_min + dlb_min > 0 ? lb_min + dlb_min : 0;
  const int m1 = ncart(la_max_l), m2 = ncart(lb_max_l);
  double *cab = calloc((size_t)m1 * m2, sizeof(double));
  cab_grid(1, orthorhombic, border_mask, la_max_l, la_min_l, lb_max_l, lb_min_l,
           zeta, zetb, 1.0, dh, dh_inv, ra, rab, npts_global, npts_local,
           shift_local, border_width, radius, cab, (double *)grid);
  const cmat c = {cab, m1};
  const double za = zeta, zb = zetb;
  for (int la = la_min; la <= la_max; la++)
    for (int ax = 0; ax <= la; ax++)
      for (int ay = 0; ay <= la - ax; ay++) {
        const ang a = {{ax, ay, la - ax - ay}};
        for (int lb = lb_min; lb <= lb_max; lb++)
          for (int bx = 0; bx <= lb; bx++)
            for (int by = 0; by <= lb - bx; by++) {
              const ang b = {{bx, by, lb - bx - by}};
              const long e = (long)(o2 + aidx(b)) * n1 + o1 + aidx(a);
              hab[e] += get_hab(a, b, za, zb, &c, compute_tau);
              if (forces) {
                const double pv = pab[e];
                for (int i = 0; i < 3; i++) {
                  forces[i] += pv * get_fa(a, b, i, za, zb, &c, compute_tau);
                  forces[3 + i] += pv * get_fb(a, b, i, za, zb, rab, &c, compute_tau);
                }
              }
              if (virials) {
                const double pv = pab[e];
                for (int i = 0; i < 3; i++)
                  for (int j = 0; j < 3; j++) {
                    virials[i * 3 + j] += pv * get_va(a, b, i, j, za, zb, &c, compute_tau);
                    virials[9 + i * 3 + j] +=
                        pv * get_vb(a, b, i, j, za, zb, rab, &c, compute_tau);
                  }
              }
              /* grid_ref_integrate.c:126-149: the same getters without the
               * density weight, per hab element */
              for (int i = 0; i < 3; i++) {
                if (hdab) hdab[3 * e + i] += get_fa(a, b, i, za, zb, &c, false);
                if (hadb) hadb[3 * e + i] += get_fb(a, b, i, za, zb, rab, &c, false);
                if (a_hdab)
                  for (int j = 0; j < 3; j++)
                    a_hdab[9 * e + 3 * i + j] += get_va(a, b, i, j, za, zb, &c, false);
              }
            }
      }
  free(cab);
}

/* ------------------------------------------------------------------------- */
/* Task list, src/grid/ref/grid_ref_task_list.c.                             */
/* ------------------------------------------------------------------------- */
typedef struct {
  int level, iatom, jatom, iset, jset, ipgf, jpgf, border_mask, block_num;
  int order; /* position in the caller's list: makes the sort total */
  double radius, rab[3];
} otask;

typedef struct {
  int npts_global[3], npts_local[3], shift_local[3], border_width[3];
  double dh[3][3], dh_inv[3][3];
} olayout;

struct oracle_task_list {
  bool orthorhombic;
  int ntasks, nlevels, natoms, nkinds, nblocks, maxco;
  int *block_offsets;
  double *atom_positions;
  int *atom_kinds;
  const oracle_basis **basis;
  otask *tasks;
  olayout *layouts;
  int *first, *last; /* [nlevels][nblocks] task ranges, last < first if none */
};

/* Task order: level, block, iset, jset (grid_ref_task_list.c:26-37). */
static int task_cmp(const void *pa, const void *pb) {
  const otask *a = pa, *b = pb;
  if (a->level != b->level) return a->level - b->level;
  if (a->block_num != b->block_num) return a->block_num - b->block_num;
  if (a->iset != b->iset) return a->iset - b->iset;
  if (a->jset != b->jset) return a->jset - b->jset;
  return a->order - b->order;
}

oracle_task_list *oracle_create_task_list(
    bool orthorhombic, int ntasks, int nlevels, int natoms, int nkinds,
    int nblocks, const int *block_offsets, const double *atom_positions,
    const int *atom_kinds, const oracle_basis *const *basis_sets,
    const int *level_list, const int *iatom_list, const int *jatom_list,
    const int *iset_list, const int *jset_list, const int *ipgf_list,
    const int *jpgf_list, const int *border_mask_list,
    const int *block_num_list, const double *radius_list,
    const double *rab_list, const int *npts_global, const int *npts_local,
    const int *shift_local, const int *border_width, const double *dh,
    const double *dh_inv) {
  oracle_task_list *tl = calloc(1, sizeof(*tl));
  tl->orthorhombic = orthorhombic;
  tl->ntasks = ntasks;
  tl->nlevels = nlevels;
  tl->natoms = natoms;
  tl->nkinds = nkinds;
  tl->nblocks = nblocks;
  tl->block_offsets = malloc(sizeof(int) * (nblocks + 1));
  memcpy(tl->block_offsets, block_offsets, sizeof(int) * nblocks);
  tl->atom_positions = malloc(sizeof(double) * 3 * (natoms + 1));
  memcpy(tl->atom_positions, atom_positions, sizeof(double) * 3 * natoms);
  tl->atom_kinds = malloc(sizeof(int) * (natoms + 1));
  memcpy(tl->atom_kinds, atom_kinds, sizeof(int) * natoms);
  tl->basis = malloc(sizeof(oracle_basis *) * (nkinds + 1));
  for (int i = 0; i < nkinds; i++) tl->basis[i] = basis_sets[i];
  tl->maxco = 0;
  for (int i = 0; i < nkinds; i++)
    if (basis_sets[i]->maxco > tl->maxco) tl->maxco = basis_sets[i]->maxco;
  tl->tasks = malloc(sizeof(otask) * (ntasks + 1));
  for (int t = 0; t < ntasks; t++) {
    otask *k = &tl->tasks[t];
    k->level = level_list[t];
    k->iatom = iatom_list[t];
    k->jatom = jatom_list[t];
    k->iset = iset_list[t];
    k->jset = jset_list[t];
    k->ipgf = ipgf_list[t];
    k->jpgf = jpgf_list[t];
    k->border_mask = border_mask_list[t];
    k->block_num = block_num_list[t];
    k->radius = radius_list[t];
    for (int d = 0; d < 3; d++) k->rab[d] = rab_list[3 * t + d];
    k->order = t;
  }
  tl->layouts = malloc(sizeof(olayout) * (nlevels + 1));
  for (int l = 0; l < nlevels; l++) {
    olayout *y = &tl->layouts[l];
    for (int i = 0; i < 3; i++) {
      y->npts_global[i] = npts_global[3 * l + i];
      y->npts_local[i] = npts_local[3 * l + i];
      y->shift_local[i] = shift_local[3 * l + i];
      y->border_width[i] = border_width[3 * l + i];
      for (int j = 0; j < 3; j++) {
        y->dh[i][j] = dh[9 * l + 3 * i + j];
        y->dh_inv[i][j] = dh_inv[9 * l + 3 * i + j];
      }
    }
  }
  qsort(tl->tasks, ntasks, sizeof(otask), task_cmp);
  const long nlb = (long)nlevels * nblocks;
  tl->first = malloc(sizeof(int) * (nlb + 1));
  tl->last = malloc(sizeof(int) * (nlb + 1));
  for (long i = 0; i < nlb; i++) {
    tl->first[i] = 0;
    tl->last[i] = -1;
  }
  for (int t = 0; t < ntasks; t++) {
    const long key = (long)(tl->tasks[t].level - 1) * nblocks + tl->tasks[t].block_num - 1;
    if (t == 0 || tl->tasks[t - 1].level != tl->tasks[t].level ||
        tl->tasks[t - 1].block_num != tl->tasks[t].block_num)
      tl->first[key] = t;
    tl->last[key] = t;
  }
  return tl;
}

void oracle_free_task_list(oracle_task_list *tl) {
  if (!tl) return;
  free(tl->block_offsets);
  free(tl->atom_positions);
  free(tl->atom_kinds);
  free(tl->basis);
  free(tl->tasks);
  free(tl->layouts);
  free(tl->first);
  free(tl->last);
  free(tl);
}

/* Decontraction of one (block, iset, jset) sub-block to primitive Cartesian
 * pab[ncob][ncoa] = sphi_b^T S^T sphi_a, grid_ref_task_list.c:237-271.
 * The block is stored column-major with the lower-numbered atom as the row
 * atom; `transpose` (iatom <= jatom) selects the layout. */
static void decontract(const oracle_basis *ib, const oracle_basis *jb, int iset,
                       int jset, bool transpose, const double *block, double *pab) {
  const int ncoa = ib->npgf[iset] * ncart(ib->lmax[iset]);
  const int ncob = jb->npgf[jset] * ncart(jb->lmax[jset]);
  const int nsa = ib->nsgf_set[iset], nsb = jb->nsgf_set[jset];
  const int sa0 = ib->first_sgf[iset] - 1, sb0 = jb->first_sgf[jset] - 1;
  double *work = calloc((size_t)nsb * ncoa + 1, sizeof(double));
  for (int sb = 0; sb < nsb; sb++)
    for (int sa = 0; sa < nsa; sa++) {
      const double s = transpose ? block[(long)(sb0 + sb) * ib->nsgf + sa0 + sa]
                                 : block[(long)(sa0 + sa) * jb->nsgf + sb0 + sb];
      const double *ph = &ib->sphi[(long)(sa0 + sa) * ib->maxco];
      for (int ca = 0; ca < ncoa; ca++) work[sb * ncoa + ca] += s * ph[ca];
    }
  for (int cb = 0; cb < ncob; cb++)
    for (int ca = 0; ca < ncoa; ca++) pab[cb * ncoa + ca] = 0.0;
  for (int sb = 0; sb < nsb; sb++) {
    const double *ph = &jb->sphi[(long)(sb0 + sb) * jb->maxco];
    for (int cb = 0; cb < ncob; cb++)
      for (int ca = 0; ca < ncoa; ca++) pab[cb * ncoa + ca] += ph[cb] * work[sb * ncoa + ca];
  }
  free(work);
}

/* Inverse of decontract, accumulated into the block,
 * grid_ref_task_list.c:462-499. */
static void contract_add(const oracle_basis *ib, const oracle_basis *jb, int iset,
                         int jset, bool transpose, const double *hab, double *block) {
  const int ncoa = ib->npgf[iset] * ncart(ib->lmax[iset]);
  const int ncob = jb->npgf[jset] * ncart(jb->lmax[jset]);
  const int nsa = ib->nsgf_set[iset], nsb = jb->nsgf_set[jset];
  const int sa0 = ib->first_sgf[iset] - 1, sb0 = jb->first_sgf[jset] - 1;
  double *work = calloc((size_t)nsb * ncoa + 1, sizeof(double));
  for (int sb = 0; sb < nsb; sb++) {
    const double *ph = &jb->sphi[(long)(sb0 + sb) * jb->maxco];
    for (int cb = 0; cb < ncob; cb++)
      for (int ca = 0; ca < ncoa; ca++) work[sb * ncoa + ca] += ph[cb] * hab[cb * ncoa + ca];
  }
  for (int sb = 0; sb < nsb; sb++)
    for (int sa = 0; sa < nsa; sa++) {
      const double *ph = &ib->sphi[(long)(sa0 + sa) * ib->maxco];
      double s = 0.0;
      for (int ca = 0; ca < ncoa; ca++) s += work[sb * ncoa + ca] * ph[ca];
      if (transpose)
        block[(long)(sb0 + sb) * ib->nsgf + sa0 + sa] += s;
      else
        block[(long)(sa0 + sa) * jb->nsgf + sb0 + sb] += s;
    }
  free(work);
}

static int nthreads_avail(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
static int thread_id(void) {
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}

void oracle_collocate_task_list(const oracle_task_list *tl, int func, int nlevels,
                                const double *pab_blocks, double *const *grids) {
  oracle_collocate_task_list_with(tl, func, nlevels, pab_blocks, grids,
                                  oracle_collocate_pgf_product);
}

void oracle_collocate_task_list_with(const oracle_task_list *tl, int func, int nlevels,
                                     const double *pab_blocks, double *const *grids,
                                     oracle_collocate_fn kernel) {
  assert(nlevels == tl->nlevels);
  int ld[4];
  if (!func_ldiffs(func, ld)) abort();
  const int nth = nthreads_avail();
  for (int level = 0; level < tl->nlevels; level++) {
    const olayout *y = &tl->layouts[level];
    const long npts = (long)y->npts_local[0] * y->npts_local[1] * y->npts_local[2];
    double **priv = malloc(sizeof(double *) * nth);
    for (int t = 0; t < nth; t++) priv[t] = NULL;
    const int *first = &tl->first[(long)level * tl->nblocks];
    const int *last = &tl->last[(long)level * tl->nblocks];
#pragma omp parallel num_threads(nth)
    {
      const int me = thread_id();
      double *g = calloc(npts, sizeof(double));
      priv[me] = g;
      double *pab = malloc(sizeof(double) * (tl->maxco * tl->maxco + 1));
      int old_off = -1, old_iset = -1, old_jset = -1;
#pragma omp for schedule(dynamic, 1)
      for (int blk = 0; blk < tl->nblocks; blk++) {
        for (int t = first[blk]; t <= last[blk]; t++) {
          const otask *k = &tl->tasks[t];
          const int ia = k->iatom - 1, ja = k->jatom - 1;
          const int is = k->iset - 1, js = k->jset - 1;
          const oracle_basis *ib = tl->basis[tl->atom_kinds[ia] - 1];
          const oracle_basis *jb = tl->basis[tl->atom_kinds[ja] - 1];
          const int off = tl->block_offsets[k->block_num - 1];
          if (off != old_off || is != old_iset || js != old_jset) {
            decontract(ib, jb, is, js, ia <= ja, &pab_blocks[off], pab);
            old_off = off;
            old_iset = is;
            old_jset = js;
          }
          const int nca = ncart(ib->lmax[is]), ncb = ncart(jb->lmax[js]);
          kernel(
              tl->orthorhombic, k->border_mask, func, ib->lmax[is], ib->lmin[is],
              jb->lmax[js], jb->lmin[js], ib->zet[is * ib->maxpgf + k->ipgf - 1],
              jb->zet[js * jb->maxpgf + k->jpgf - 1], ia == ja ? 1.0 : 2.0, y->dh,
              y->dh_inv, &tl->atom_positions[3 * ia], k->rab, y->npts_global,
              y->npts_local, y->shift_local, y->border_width, k->radius,
              (k->ipgf - 1) * nca, (k->jpgf - 1) * ncb, ib->npgf[is] * nca,
              jb->npgf[js] * ncb, pab, g);
        }
      }
      free(pab);
    }
    /* Sum private grids in thread order (the reference uses a tree). */
    double *out = grids[level];
#pragma omp parallel for schedule(static)
    for (long p = 0; p < npts; p++) {
      double s = 0.0;
      for (int t = 0; t < nth; t++)
        if (priv[t]) s += priv[t][p];
      out[p] = s;
    }
    for (int t = 0; t < nth; t++) free(priv[t]);
    free(priv);
  }
}

/* The oracle's integrate kernel behind the reference-style signature. */
static void integrate_adapter(bool ortho, bool tau, int bm, int la_max, int la_min,
                              int lb_max, int lb_min, double za, double zb,
                              const double dh[3][3], const double dh_inv[3][3],
                              const double ra[3], const double rab[3], const int ng[3],
                              const int nl[3], const int sh[3], const int bw[3],
                              double radius, int o1, int o2, int n1, int n2,
                              const double *grid, double *hab, const double *pab,
                              double *forces, double *virials, void *hdab, void *hadb,
                              void *a_hdab) {
  (void)hdab; (void)hadb; (void)a_hdab;
  oracle_integrate_pgf_product(ortho, tau, bm, la_max, la_min, lb_max, lb_min, za, zb, dh,
                               dh_inv, ra, rab, ng, nl, sh, bw, radius, o1, o2, n1, n2,
                               grid, hab, pab, forces, virials);
}

void oracle_integrate_task_list(const oracle_task_list *tl, bool compute_tau,
                                int natoms, int nlevels, const double *pab_blocks,
                                const double *const *grids, double *hab_blocks,
                                long hab_len, double *forces, double *virial) {
  oracle_integrate_task_list_with(tl, compute_tau, natoms, nlevels, pab_blocks, grids,
                                  hab_blocks, hab_len, forces, virial, integrate_adapter);
}

void oracle_integrate_task_list_with(const oracle_task_list *tl, bool compute_tau,
                                     int natoms, int nlevels, const double *pab_blocks,
                                     const double *const *grids, double *hab_blocks,
                                     long hab_len, double *forces, double *virial,
                                     oracle_integrate_fn kernel) {
  assert(nlevels == tl->nlevels && natoms == tl->natoms);
  assert(!(forces || virial) || pab_blocks);
  memset(hab_blocks, 0, sizeof(double) * hab_len);
  if (forces) memset(forces, 0, sizeof(double) * 3 * natoms);
  if (virial) memset(virial, 0, sizeof(double) * 9);
  const int nth = nthreads_avail();
  for (int level = 0; level < tl->nlevels; level++) {
    const olayout *y = &tl->layouts[level];
    const int *first = &tl->first[(long)level * tl->nblocks];
    const int *last = &tl->last[(long)level * tl->nblocks];
    const double *grid = grids[level];
#pragma omp parallel num_threads(nth)
    {
      const size_t mm = (size_t)tl->maxco * tl->maxco + 1;
      double *pab = malloc(sizeof(double) * mm);
      double *hab = malloc(sizeof(double) * mm);
#pragma omp for schedule(dynamic, 1)
      for (int blk = 0; blk < tl->nblocks; blk++) {
        if (last[blk] < first[blk]) continue;
        const int ia = tl->tasks[first[blk]].iatom - 1;
        const int ja = tl->tasks[first[blk]].jatom - 1;
        const oracle_basis *ib = tl->basis[tl->atom_kinds[ia] - 1];
        const oracle_basis *jb = tl->basis[tl->atom_kinds[ja] - 1];
        const int off = tl->block_offsets[blk];
        double fblk[6] = {0}, vblk[18] = {0};
        int cur_is = -1, cur_js = -1;
        for (int t = first[blk]; t <= last[blk]; t++) {
          const otask *k = &tl->tasks[t];
          assert(k->iatom - 1 == ia && k->jatom - 1 == ja);
          const int is = k->iset - 1, js = k->jset - 1;
          const int nca = ncart(ib->lmax[is]), ncb = ncart(jb->lmax[js]);
          const int ncoa = ib->npgf[is] * nca, ncob = jb->npgf[js] * ncb;
          if (is != cur_is || js != cur_js) {
            if (cur_is >= 0)
              contract_add(ib, jb, cur_is, cur_js, ia <= ja, hab, &hab_blocks[off]);
            if (forces || virial)
              decontract(ib, jb, is, js, ia <= ja, &pab_blocks[off], pab);
            memset(hab, 0, sizeof(double) * ncoa * ncob);
            cur_is = is;
            cur_js = js;
          }
          kernel(
              tl->orthorhombic, compute_tau, k->border_mask, ib->lmax[is],
              ib->lmin[is], jb->lmax[js], jb->lmin[js],
              ib->zet[is * ib->maxpgf + k->ipgf - 1],
              jb->zet[js * jb->maxpgf + k->jpgf - 1], y->dh, y->dh_inv,
              &tl->atom_positions[3 * ia], k->rab, y->npts_global, y->npts_local,
              y->shift_local, y->border_width, k->radius, (k->ipgf - 1) * nca,
              (k->jpgf - 1) * ncb, ncoa, ncob, grid, hab,
              (forces || virial) ? pab : NULL, forces ? fblk : NULL,
              virial ? vblk : NULL, NULL, NULL, NULL);
        }
        if (cur_is >= 0)
          contract_add(ib, jb, cur_is, cur_js, ia <= ja, hab, &hab_blocks[off]);
        const double sc = ia == ja ? 1.0 : 2.0;
        if (forces) {
#pragma omp critical(oracle_forces)
          for (int i = 0; i < 3; i++) {
            forces[3 * ia + i] += sc * fblk[i];
            forces[3 * ja + i] += sc * fblk[3 + i];
          }
        }
        if (virial) {
#pragma omp critical(oracle_virial)
          for (int i = 0; i < 9; i++) {
            virial[i] += sc * vblk[i];
            virial[i] += sc * vblk[9 + i];
          }
        }
      }
      free(pab);
      free(hab);
    }
  }
}

long oracle_task_list_points(const oracle_task_list *tl, int lp_extra) {
  (void)lp_extra;
  long total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
  for (int t = 0; t < tl->ntasks; t++) {
    const otask *k = &tl->tasks[t];
    const olayout *y = &tl->layouts[k->level - 1];
    const int ia = k->iatom - 1, ja = k->jatom - 1;
    const oracle_basis *ib = tl->basis[tl->atom_kinds[ia] - 1];
    const oracle_basis *jb = tl->basis[tl->atom_kinds[ja] - 1];
    total += oracle_task_points(
        tl->orthorhombic, k->border_mask, 0,
        ib->zet[(k->iset - 1) * ib->maxpgf + k->ipgf - 1],
        jb->zet[(k->jset - 1) * jb->maxpgf + k->jpgf - 1], y->dh, y->dh_inv,
        &tl->atom_positions[3 * ia], k->rab, y->npts_global, y->npts_local,
        y->shift_local, y->border_width, k->radius);
  }
  return total;
}
