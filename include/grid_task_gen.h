/*
 * grid_task_gen.h -- task-list generation on the MI355X.
 *
 * Replaces the part of CP2K's generate_qs_task_list that turns the orbital
 * neighbour list into the argument set of grid_create_task_list, for
 * replicated (not distributed) real-space grids:
 *
 *   generate_qs_task_list   src/task_list_methods.F:117-501 (the neighbour
 *                           list loop :246-294)
 *   task_list_inner_loop    :652-719  set / kind / pgf radius screening
 *   compute_pgf_properties  :755-800  product radius from
 *                           exp_radius_very_extended (src/aobasis/ao_util.F
 *                           :207-290, exp_radius :95-190), cutoff 1
 *   gaussian_gridlevel      src/pw_env/gaussian_gridlevels.F:146-166
 *   pgf_to_tasks            :876-943  one task per surviving pgf pair
 *   distribute_tasks        :1889-2058 (not distributed): get_atom_pair
 *                           :2070-2136 numbers the atom pairs (blocks),
 *                           tasks_sort orders the tasks by tasks_less_than
 *                           :2931-2960 (stable merge sort)
 *   rs_calc_offsets         :2439-2484  block offsets in the pab buffer
 *   create_grid_task_list   :507-618  the lists handed to grid_create_task_list
 *
 * The whole pipeline runs on the device: screening count, scan, the radius
 * bisections, a stable radix sort of the tasks by (grid level, image, iatom,
 * jatom, iset, jset, ipgf, jpgf) and the block numbering.  The outputs are the
 * arrays of create_grid_task_list, in its order, ready for
 * grid_create_task_list (the border masks are 0: no distributed grid).
 */
#ifndef GRID_TASK_GEN_H
#define GRID_TASK_GEN_H

#include "grid_basis_set.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Inputs (host memory):
 *   pair_atoms [npairs][2]  iatom, jatom (1-based) of each neighbour-list
 *                           entry, in neighbour-list order
 *   pair_rab   [npairs][3]  rab = rb - ra of the entry (with its image)
 *   pair_image [npairs]     cell index of the entry (1-based; NULL: all 1,
 *                           as without k-points)
 *   atom_kinds [natoms]     kind of each atom (1-based)
 *   atom_positions [natoms][3]  pbc(r) of each atom (ra of its pairs)
 *   basis_sets [nkinds]     lmin, lmax, npgf, zet and nsgf of each kind
 *   set_radius [nkinds][maxset], pgf_radius [nkinds][maxset][maxpgf]:
 *                           the basis sets' set_radius and pgf_radius
 *                           (pgf_radius(ipgf, iset) of the reference)
 *   kind_radius [nkinds]    the kinds' kind_radius (NULL: the largest set
 *                           radius of the kind, as for the orbital basis)
 *   cutoff [nlevels], rel_cutoff: the grid levels' cutoffs
 *                           (gridlevel_info%cutoff, %rel_cutoff)
 *   eps_rho                 eps_rho_rspace
 * Outputs (host memory, capacity entries; rab_list 3 per task):
 *   the task arrays of grid_create_task_list, border_mask_list all 0;
 *   block_offsets (capacity >= npairs entries) and *nblocks: the blocks
 *   (distinct (min(iatom, jatom), max(iatom, jatom), image) of the tasks, in
 *   that order) and their offsets in the pab / hab buffer
 *   (nsgf(row) * nsgf(col) doubles each), *buffer_size their total.
 * Returns the number of tasks.  If it exceeds capacity nothing is written
 * and the call can be repeated with a larger capacity.  Aborts with a
 * message on invalid input (sizes beyond 2^21 atoms, 2^18 images, 16 levels,
 * 256 sets or primitives per set, l above 10).
 */
long long grid_mi355x_generate_task_list(
    const int npairs, const int *pair_atoms, const double *pair_rab,
    const int *pair_image, const int natoms, const int *atom_kinds,
    const double *atom_positions, const int nkinds,
    const grid_basis_set *const *basis_sets, const int maxset, const int maxpgf,
    const double *set_radius, const double *pgf_radius, const double *kind_radius,
    const int nlevels, const double *cutoff, const double rel_cutoff,
    const double eps_rho, const long long capacity, int *level_list,
    int *iatom_list, int *jatom_list, int *iset_list, int *jset_list,
    int *ipgf_list, int *jpgf_list, int *border_mask_list, int *block_num_list,
    double *radius_list, double *rab_list, int *nblocks, int *block_offsets,
    long long *buffer_size);

/* Milliseconds of device time (HIP events on the library stream) of the last
 * grid_mi355x_generate_task_list call, from the upload of the inputs to the
 * last output kernel (without the copies of the outputs to the host). */
double grid_mi355x_generate_task_list_ms(void);

#ifdef __cplusplus
}
#endif

#endif
