/*
 * task_gen.c -- CPU restatement of CP2K's task-list generation for
 * replicated real-space grids (the part of generate_qs_task_list that turns
 * the orbital neighbour list into the arguments of grid_create_task_list).
 *
 * TEST INFRASTRUCTURE ONLY (see grid_oracle.h for who may load it).  Plain
 * loops in the reference's order:
 *
 *   oracle_exp_radius         src/aobasis/ao_util.F:95-190
 *   oracle_product_radius     src/aobasis/ao_util.F:207-290 (no pab, cutoff 1)
 *   oracle_gridlevel          src/pw_env/gaussian_gridlevels.F:146-166
 *   oracle_generate_task_list src/task_list_methods.F:246-294 (neighbour-list
 *       loop), :652-719 (task_list_inner_loop), :755-800
 *       (compute_pgf_properties), :876-943 (pgf_to_tasks), :2070-2136
 *       (get_atom_pair), :2931-2960 (tasks_less_than, stable merge sort),
 *       :2439-2484 (rs_calc_offsets), :507-618 (create_grid_task_list)
 *
 * Single-precision exponentials are expf of the argument rounded to float
 * (EXP of a REAL(sp) value); x**l with integer l is right-to-left binary
 * powering.  Build with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "grid_oracle.h"

static double exp_sp(double x) { return (double)expf((float)x); }

static double powi(double x, int n) {
  double y = (n & 1) ? x : 1.0;
  while (n >>= 1) {
    x = x * x;
    if (n & 1) y = y * x;
  }
  return y;
}

/* exp_radius with epsabs given and epsrel absent; returns the upper end of
 * the final bracket (rlow when prefactor is 0 or the start is already below
 * the threshold).  lo_out (may be NULL) receives the lower end. */
double oracle_exp_radius_bracket(int l, double alpha, double threshold, double prefactor,
                                 double epsabs, double rlow, double *lo_out) {
  const double contract[2] = {(double)0.38f, (double)0.62f};
  const double a = fabs(alpha), t = fabs(threshold);
  double radius = rlow;
  if (lo_out) *lo_out = rlow;
  if (prefactor == 0.0) return radius;
  double r = fmax(sqrt(0.5 * (double)l / a), radius);
  const double d = fabs(prefactor);
  double g = d;
  if (l != 0) g = g * exp_sp(-a * r * r) * powi(r, l);
  if (g < t) return radius;
  radius = r * 2.0 + 1.0;
  int i;
  for (i = 1; i <= 5000; i++) {
    g = d * exp_sp(-a * radius * radius) * powi(radius, l);
    if (g < t) break;
    r = radius;
    radius = r * 2.0 + 1.0;
  }
  double dr = 0.0;
  for (i = i + 1; i <= 5000; i++) {
    const double rd = radius - r;
    if (rd < epsabs || rd == dr) break;
    double s[2], gs[2];
    for (int j = 0; j < 2; j++) {
      s[j] = r + rd * contract[j];
      gs[j] = d * exp_sp(-a * s[j] * s[j]) * powi(s[j], l);
    }
    for (int j = 0; j < 2; j++) {
      if (gs[j] < t) {
        radius = s[j];
        break;
      }
      r = s[j];
    }
    dr = rd;
  }
  if (lo_out) *lo_out = r;
  return radius;
}

double oracle_exp_radius(int l, double alpha, double threshold, double prefactor,
                         double epsabs, double rlow) {
  return oracle_exp_radius_bracket(l, alpha, threshold, prefactor, epsabs, rlow, NULL);
}

/* polycoef of exp_radius_very_extended (before the prefactor): poly[i],
 * i = 0 .. la_max + lb_max. */
void oracle_product_polycoef(int la_max, int lb_max, double rad_a, double rad_b, double *poly) {
  const int lp = la_max + lb_max;
  for (int i = 0; i <= lp; i++) poly[i] = 0.0;
  for (int lxa = 0; lxa <= la_max; lxa++)
    for (int lxb = 0; lxb <= lb_max; lxb++) {
      double coef[21];
      for (int i = 0; i <= lp; i++) coef[i] = 0.0;
      double bini = 1.0, s1 = 1.0;
      for (int i = 0; i <= lxa; i++) {
        double binj = 1.0, s2 = 1.0;
        for (int j = 0; j <= lxb; j++) {
          coef[lxa + lxb - i - j] = coef[lxa + lxb - i - j] + bini * binj * s1 * s2;
          binj = (binj * (lxb - j)) / (j + 1);
          s2 = s2 * rad_b;
        }
        bini = (bini * (lxa - i)) / (i + 1);
        s1 = s1 * rad_a;
      }
      for (int i = 0; i <= lxa + lxb; i++) poly[i] = fmax(poly[i], coef[i]);
    }
}

double oracle_product_radius(int la_max, int lb_max, const double ra[3], const double rb[3],
                             const double rp[3], double zetp, double eps, double prefactor) {
  double sa = 0.0, sb = 0.0;
  for (int d = 0; d < 3; d++) {
    sa = sa + (ra[d] - rp[d]) * (ra[d] - rp[d]);
    sb = sb + (rb[d] - rp[d]) * (rb[d] - rp[d]);
  }
  double poly[21];
  oracle_product_polycoef(la_max, lb_max, sqrt(sa), sqrt(sb), poly);
  const double prefactor_local = prefactor * 1.0; /* MAX(1, cutoff) with cutoff 1 */
  for (int i = 0; i <= la_max + lb_max; i++) poly[i] = poly[i] * prefactor_local;
  double radius = 0.0;
  for (int i = 0; i <= la_max + lb_max; i++)
    radius = fmax(radius, oracle_exp_radius(i, zetp, eps, poly[i], 1.0e-2, radius));
  return radius;
}

int oracle_gridlevel(int nlevels, const double *cutoff, double rel_cutoff, double exponent) {
  int level = 1;
  const double needed = fabs(exponent) * rel_cutoff;
  for (int i = 1; i <= nlevels; i++)
    if (cutoff[i - 1] + 1e-6 >= needed) level = i;
  return level;
}

/* ---- task list ------------------------------------------------------------ */

typedef struct {
  int level, image, iatom, jatom, iset, jset, ipgf, jpgf;
  long gen;   /* position in generation order: ties keep it (stable sort) */
  int pair;   /* neighbour-list entry */
  double radius;
} gen_task;

static int gen_task_cmp(const void *pa, const void *pb) {
  const gen_task *a = pa, *b = pb;
  const int ka[8] = {a->level, a->image, a->iatom, a->jatom, a->iset, a->jset, a->ipgf, a->jpgf};
  const int kb[8] = {b->level, b->image, b->iatom, b->jatom, b->iset, b->jset, b->ipgf, b->jpgf};
  for (int i = 0; i < 8; i++)
    if (ka[i] != kb[i]) return ka[i] < kb[i] ? -1 : 1;
  return a->gen < b->gen ? -1 : (a->gen > b->gen);
}

typedef struct {
  int row, col, image;
} atom_pair;

static int atom_pair_cmp(const void *pa, const void *pb) {
  const atom_pair *a = pa, *b = pb;
  if (a->row != b->row) return a->row < b->row ? -1 : 1;
  if (a->col != b->col) return a->col < b->col ? -1 : 1;
  return a->image < b->image ? -1 : (a->image > b->image);
}

long long oracle_generate_task_list(
    int npairs, const int *pair_atoms, const double *pair_rab, const int *pair_image,
    int natoms, const int *atom_kinds, const double *atom_positions, int nkinds,
    const oracle_basis *const *basis_sets, int maxset, int maxpgf, const double *set_radius,
    const double *pgf_radius, const double *kind_radius, int nlevels, const double *cutoff,
    double rel_cutoff, double eps_rho, long long capacity, int *level_list, int *iatom_list,
    int *jatom_list, int *iset_list, int *jset_list, int *ipgf_list, int *jpgf_list,
    int *border_mask_list, int *block_num_list, double *radius_list, double *rab_list,
    int *nblocks, int *block_offsets, long long *buffer_size) {
  (void)natoms;
  long cap = 1024, n = 0;
  gen_task *tasks = malloc(sizeof(gen_task) * cap);
  for (int p = 0; p < npairs; p++) {
    const int iatom = pair_atoms[2 * p], jatom = pair_atoms[2 * p + 1];
    const int ka = atom_kinds[iatom - 1] - 1, kb = atom_kinds[jatom - 1] - 1;
    const oracle_basis *A = basis_sets[ka], *B = basis_sets[kb];
    const double *rab = pair_rab + 3 * p;
    const double *ra = atom_positions + 3 * (iatom - 1);
    double kind_radius_b = 0.0;
    if (kind_radius) {
      kind_radius_b = kind_radius[kb];
    } else {
      for (int s = 0; s < B->nset; s++)
        kind_radius_b = fmax(kind_radius_b, set_radius[kb * maxset + s]);
    }
    const double rab2 = rab[0] * rab[0] + rab[1] * rab[1] + rab[2] * rab[2];
    const double dab = sqrt(rab2);
    for (int iset = 1; iset <= A->nset; iset++) {
      const double sra = set_radius[ka * maxset + iset - 1];
      if (sra + kind_radius_b < dab) continue;
      for (int jset = 1; jset <= B->nset; jset++) {
        const double srb = set_radius[kb * maxset + jset - 1];
        if (sra + srb < dab) continue;
        for (int ipgf = 1; ipgf <= A->npgf[iset - 1]; ipgf++) {
          const double rpa = pgf_radius[((long)ka * maxset + iset - 1) * maxpgf + ipgf - 1];
          if (rpa + srb < dab) continue;
          for (int jpgf = 1; jpgf <= B->npgf[jset - 1]; jpgf++) {
            const double rpb = pgf_radius[((long)kb * maxset + jset - 1) * maxpgf + jpgf - 1];
            if (rpa + rpb < dab) continue;
            const double zeta = A->zet[(iset - 1) * A->maxpgf + ipgf - 1];
            const double zetb = B->zet[(jset - 1) * B->maxpgf + jpgf - 1];
            const double zetp = zeta + zetb;
            double rp[3], rb[3];
            for (int d = 0; d < 3; d++) {
              rp[d] = ra[d] + zetb / zetp * rab[d];
              rb[d] = ra[d] + rab[d];
            }
            const double f = zetb / zetp;
            const double prefactor = exp(-zeta * f * rab2);
            if (n == cap) {
              cap *= 2;
              tasks = realloc(tasks, sizeof(gen_task) * cap);
            }
            gen_task *t = &tasks[n];
            t->level = oracle_gridlevel(nlevels, cutoff, rel_cutoff, zetp);
            t->image = pair_image ? pair_image[p] : 1;
            t->iatom = iatom;
            t->jatom = jatom;
            t->iset = iset;
            t->jset = jset;
            t->ipgf = ipgf;
            t->jpgf = jpgf;
            t->gen = n;
            t->pair = p;
            t->radius = oracle_product_radius(A->lmax[iset - 1], B->lmax[jset - 1], ra, rb, rp,
                                              zetp, eps_rho, prefactor);
            n++;
          }
        }
      }
    }
  }
  if (n > capacity || n == 0) {
    if (n == 0) {
      *nblocks = 0;
      *buffer_size = 0;
    }
    free(tasks);
    return n;
  }
  /* get_atom_pair (receive side, symmetric): blocks numbered by sorted
   * distinct (min atom, max atom, image) */
  atom_pair *pairs = malloc(sizeof(atom_pair) * n);
  for (long i = 0; i < n; i++) {
    const gen_task *t = &tasks[i];
    pairs[i].row = t->iatom < t->jatom ? t->iatom : t->jatom;
    pairs[i].col = t->iatom < t->jatom ? t->jatom : t->iatom;
    pairs[i].image = t->image;
  }
  atom_pair *uniq = malloc(sizeof(atom_pair) * n);
  memcpy(uniq, pairs, sizeof(atom_pair) * n);
  qsort(uniq, n, sizeof(atom_pair), atom_pair_cmp);
  long nu = 0;
  for (long i = 0; i < n; i++)
    if (nu == 0 || atom_pair_cmp(&uniq[nu - 1], &uniq[i]) != 0) uniq[nu++] = uniq[i];
  /* rs_calc_offsets */
  long long total = 0;
  for (long b = 0; b < nu; b++) {
    block_offsets[b] = (int)total;
    const int kr = atom_kinds[uniq[b].row - 1] - 1, kc = atom_kinds[uniq[b].col - 1] - 1;
    total += (long long)basis_sets[kr]->nsgf * basis_sets[kc]->nsgf;
  }
  *nblocks = (int)nu;
  *buffer_size = total;
  /* tasks_sort (stable) */
  int *block_of = malloc(sizeof(int) * n);
  for (long i = 0; i < n; i++) {
    const atom_pair *hit = bsearch(&pairs[i], uniq, nu, sizeof(atom_pair), atom_pair_cmp);
    block_of[i] = (int)(hit - uniq) + 1;
  }
  gen_task *sorted = malloc(sizeof(gen_task) * n);
  memcpy(sorted, tasks, sizeof(gen_task) * n);
  qsort(sorted, n, sizeof(gen_task), gen_task_cmp);
  for (long i = 0; i < n; i++) {
    const gen_task *t = &sorted[i];
    level_list[i] = t->level;
    iatom_list[i] = t->iatom;
    jatom_list[i] = t->jatom;
    iset_list[i] = t->iset;
    jset_list[i] = t->jset;
    ipgf_list[i] = t->ipgf;
    jpgf_list[i] = t->jpgf;
    border_mask_list[i] = 0;
    block_num_list[i] = block_of[t->gen];
    radius_list[i] = t->radius;
    for (int d = 0; d < 3; d++) rab_list[3 * i + d] = pair_rab[3 * t->pair + d];
  }
  free(sorted);
  free(block_of);
  free(uniq);
  free(pairs);
  free(tasks);
  return n;
}
