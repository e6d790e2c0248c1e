/*
 * grid_oracle.h -- CPU restatement of CP2K's FP64 "ref" grid backend.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X
 * backend in cp2k_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never links
 * it and never falls back to it.
 *
 * Every routine restates the algorithm of the reference file cited next to
 * it (paths relative to the cp2k source tree).  Membership of grid points in a
 * task's truncation region follows the reference operation by operation
 * (compiled with -ffp-contract=off), so the same set of points is visited;
 * the Gaussian factor tables are evaluated with one exp() per entry instead
 * of the reference's multiplicative recurrences, which changes results only
 * at the level of rounding.
 *
 * Pinning (tests/test_oracle.py; tests/brute.py is the analytic evaluator):
 *   - single-task collocate/integrate/forces/virial, ortho + general,
 *     border masks, wrap-around, tau: the 13 sample_tasks known-answer files;
 *   - all 35 grid_func stencils and the hab/force/virial getters: a brute-
 *     force analytic evaluation of the Gaussian products and derivatives;
 *   - truncation regions, random l/exponents/geometry: the reference's own
 *     single-task sources compiled into oracle/_ref (when present);
 *   - task-list layer (sorting, load_pab/store_hab contraction, rscale,
 *     force scaling): batched replay of the fixtures with the reference's
 *     dummy-task-list conventions, a textbook triple-loop contraction, and the
 *     collocate/integrate adjoint identity.
 */
#ifndef GRID_ORACLE_H
#define GRID_ORACLE_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field layout as grid_basis_set (src/grid/common/grid_basis_set.h:14-26). */
typedef struct {
  int nset;
  int nsgf;
  int maxco;
  int maxpgf;
  int *lmin;
  int *lmax;
  int *npgf;
  int *nsgf_set;
  int *first_sgf;
  double *sphi; /* [nsgf][maxco] */
  double *zet;  /* [nset][maxpgf] */
} oracle_basis;

/* Single-task collocation, src/grid/ref/grid_ref_collocate.c:19-50.
 * pab is [n2][n1]; grid is [nz][ny][nx] (x fastest) and is accumulated into. */
void oracle_collocate_pgf_product(
    bool orthorhombic, int border_mask, int func, int la_max, int la_min,
    int lb_max, int lb_min, double zeta, double zetb, double rscale,
    const double dh[3][3], const double dh_inv[3][3], const double ra[3],
    const double rab[3], const int npts_global[3], const int npts_local[3],
    const int shift_local[3], const int border_width[3], double radius, int o1,
    int o2, int n1, int n2, const double *pab, double *grid);

/* Single-task integration, src/grid/ref/grid_ref_integrate.c:95-203.
 * hab is [n2][n1] and accumulated into.  forces is [2][3] (a then b),
 * virials is [2][3][3]; both optional (NULL) and accumulated into. */
void oracle_integrate_pgf_product(
    bool orthorhombic, bool compute_tau, int border_mask, int la_max,
    int la_min, int lb_max, int lb_min, double zeta, double zetb,
    const double dh[3][3], const double dh_inv[3][3], const double ra[3],
    const double rab[3], const int npts_global[3], const int npts_local[3],
    const int shift_local[3], const int border_width[3], double radius, int o1,
    int o2, int n1, int n2, const double *grid, double *hab, const double *pab,
    double *forces, double *virials);

/* The same with the reference's full argument list
 * (src/grid/ref/grid_ref_integrate.c:44-154): hdab, hadb ([n2][n1][3]) and
 * a_hdab ([n2][n1][3][3]) are the unweighted force / virial getters per hab
 * element, optional (NULL), accumulated into, not defined with compute_tau. */
void oracle_integrate_pgf_product_full(
    bool orthorhombic, bool compute_tau, int border_mask, int la_max,
    int la_min, int lb_max, int lb_min, double zeta, double zetb,
    const double dh[3][3], const double dh_inv[3][3], const double ra[3],
    const double rab[3], const int npts_global[3], const int npts_local[3],
    const int shift_local[3], const int border_width[3], double radius, int o1,
    int o2, int n1, int n2, const double *grid, double *hab, const double *pab,
    double *forces, double *virials, double *hdab, double *hadb,
    double *a_hdab);

/* Number of (task, grid point) pairs the single-task kernels visit. */
long oracle_task_points(bool orthorhombic, int border_mask, int lp,
                        double zeta, double zetb, const double dh[3][3],
                        const double dh_inv[3][3], const double ra[3],
                        const double rab[3], const int npts_global[3],
                        const int npts_local[3], const int shift_local[3],
                        const int border_width[3], double radius);

/* Task list, src/grid/ref/grid_ref_task_list.c:43-182 (same arguments as
 * grid_create_task_list, src/grid/grid_task_list.h:59-68). */
typedef struct oracle_task_list oracle_task_list;

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
    const double *dh_inv);

void oracle_free_task_list(oracle_task_list *tl);

/* src/grid/ref/grid_ref_task_list.c:434-456; grids are overwritten. */
void oracle_collocate_task_list(const oracle_task_list *tl, int func,
                                int nlevels, const double *pab_blocks,
                                double *const *grids);

/* src/grid/ref/grid_ref_task_list.c:659-691; hab (hab_len doubles), forces
 * [natoms][3] and virial [3][3] are overwritten (forces/virial optional). */
void oracle_integrate_task_list(const oracle_task_list *tl, bool compute_tau,
                                int natoms, int nlevels,
                                const double *pab_blocks,
                                const double *const *grids, double *hab_blocks,
                                long hab_len, double *forces, double *virial);

/* The same task-list drivers with a pluggable per-task kernel, so the CPU
 * baseline can run the reference's own cpu-backend kernels
 * (grid_cpu_collocate_pgf_product / grid_cpu_integrate_pgf_product, whose
 * signatures these are, src/grid/cpu/grid_cpu_collocate.h and
 * grid_cpu_integrate.h) under the oracle's OpenMP block loop. */
typedef void (*oracle_collocate_fn)(
    bool orthorhombic, int border_mask, int func, int la_max, int la_min,
    int lb_max, int lb_min, double zeta, double zetb, double rscale,
    const double dh[3][3], const double dh_inv[3][3], const double ra[3],
    const double rab[3], const int npts_global[3], const int npts_local[3],
    const int shift_local[3], const int border_width[3], double radius, int o1,
    int o2, int n1, int n2, const double *pab, double *grid);

typedef void (*oracle_integrate_fn)(
    bool orthorhombic, bool compute_tau, int border_mask, int la_max,
    int la_min, int lb_max, int lb_min, double zeta, double zetb,
    const double dh[3][3], const double dh_inv[3][3], const double ra[3],
    const double rab[3], const int npts_global[3], const int npts_local[3],
    const int shift_local[3], const int border_width[3], double radius, int o1,
    int o2, int n1, int n2, const double *grid, double *hab, const double *pab,
    double *forces, double *virials, void *hdab, void *hadb, void *a_hdab);

void oracle_collocate_task_list_with(const oracle_task_list *tl, int func,
                                     int nlevels, const double *pab_blocks,
                                     double *const *grids,
                                     oracle_collocate_fn kernel);

void oracle_integrate_task_list_with(const oracle_task_list *tl,
                                     bool compute_tau, int natoms, int nlevels,
                                     const double *pab_blocks,
                                     const double *const *grids,
                                     double *hab_blocks, long hab_len,
                                     double *forces, double *virial,
                                     oracle_integrate_fn kernel);

/* Sum of oracle_task_points over every task of the list, for func's l-shift
 * (collocate) -- the task-point count used by the throughput metric. */
long oracle_task_list_points(const oracle_task_list *tl, int lp_extra);

/* ---- task-list generation (oracle/task_gen.c) ----------------------------
 * src/task_list_methods.F generate_qs_task_list for replicated grids; the
 * arguments mean what they mean for grid_mi355x_generate_task_list
 * (include/grid_task_gen.h).  Pinned by closed forms of exp_radius (l = 0
 * and the bracket property), the binomial expansion of the polynomial
 * prefactor, gridlevel cases and a brute-force screening count. */
double oracle_exp_radius(int l, double alpha, double threshold, double prefactor,
                         double epsabs, double rlow);
double oracle_exp_radius_bracket(int l, double alpha, double threshold, double prefactor,
                                 double epsabs, double rlow, double *lo_out);
void oracle_product_polycoef(int la_max, int lb_max, double rad_a, double rad_b, double *poly);
double oracle_product_radius(int la_max, int lb_max, const double ra[3], const double rb[3],
                             const double rp[3], double zetp, double eps, double prefactor);
int oracle_gridlevel(int nlevels, const double *cutoff, double rel_cutoff, double exponent);
long long oracle_generate_task_list(
    int npairs, const int *pair_atoms, const double *pair_rab, const int *pair_image,
    int natoms, const int *atom_kinds, const double *atom_positions, int nkinds,
    const oracle_basis *const *basis_sets, int maxset, int maxpgf, const double *set_radius,
    const double *pgf_radius, const double *kind_radius, int nlevels, const double *cutoff,
    double rel_cutoff, double eps_rho, long long capacity, int *level_list, int *iatom_list,
    int *jatom_list, int *iset_list, int *jset_list, int *ipgf_list, int *jpgf_list,
    int *border_mask_list, int *block_num_list, double *radius_list, double *rab_list,
    int *nblocks, int *block_offsets, long long *buffer_size);

#ifdef __cplusplus
}
#endif

#endif
