/*
 * grid_task_list.h -- the grid task-list API (collocate / integrate).
 *
 * Exactly the entry points CP2K's Fortran binding calls through
 * ISO_C_BINDING (src/grid/grid_api.F:757-796, :981-991, :1060-1075) and that
 * src/grid/grid_task_list.h:59-127 declares.  All indices in the per-task
 * lists are one based; grids are [nz][ny][nx] with x fastest; dh[i][:] is
 * lattice step vector i.
 */
#ifndef GRID_TASK_LIST_H
#define GRID_TASK_LIST_H

#include <stdbool.h>

#include "grid_basis_set.h"
#include "grid_constants.h"
#include "offload_buffer.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef void grid_task_list;

#ifndef __cplusplus
/* src/grid/grid_task_list.h:59-68.  Copies every array; keeps the basis_sets
 * pointers.  A non-NULL *task_list is reused (rebuilt in place). */
void grid_create_task_list(
    const bool orthorhombic, const int ntasks, const int nlevels,
    const int natoms, const int nkinds, const int nblocks,
    const int block_offsets[nblocks], const double atom_positions[natoms][3],
    const int atom_kinds[natoms], const grid_basis_set *basis_sets[nkinds],
    const int level_list[ntasks], const int iatom_list[ntasks],
    const int jatom_list[ntasks], const int iset_list[ntasks],
    const int jset_list[ntasks], const int ipgf_list[ntasks],
    const int jpgf_list[ntasks], const int border_mask_list[ntasks],
    const int block_num_list[ntasks], const double radius_list[ntasks],
    const double rab_list[ntasks][3], const int npts_global[nlevels][3],
    const int npts_local[nlevels][3], const int shift_local[nlevels][3],
    const int border_width[nlevels][3], const double dh[nlevels][3][3],
    const double dh_inv[nlevels][3][3], grid_task_list **task_list);

/* src/grid/grid_task_list.h:95-100.  Overwrites every level grid. */
void grid_collocate_task_list(const grid_task_list *task_list,
                              const enum grid_func func, const int nlevels,
                              const int npts_local[nlevels][3],
                              const offload_buffer *pab_blocks,
                              offload_buffer *grids[nlevels]);

/* src/grid/grid_task_list.h:122-127.  Overwrites hab_blocks and, when not
 * NULL, forces and virial (which require pab_blocks). */
void grid_integrate_task_list(
    const grid_task_list *task_list, const bool compute_tau, const int natoms,
    const int nlevels, const int npts_local[nlevels][3],
    const offload_buffer *pab_blocks, const offload_buffer *grids[nlevels],
    offload_buffer *hab_blocks, double forces[natoms][3], double virial[3][3]);
#else
void grid_create_task_list(
    const bool orthorhombic, const int ntasks, const int nlevels,
    const int natoms, const int nkinds, const int nblocks,
    const int *block_offsets, const double (*atom_positions)[3],
    const int *atom_kinds, const grid_basis_set **basis_sets,
    const int *level_list, const int *iatom_list, const int *jatom_list,
    const int *iset_list, const int *jset_list, const int *ipgf_list,
    const int *jpgf_list, const int *border_mask_list,
    const int *block_num_list, const double *radius_list,
    const double (*rab_list)[3], const int (*npts_global)[3],
    const int (*npts_local)[3], const int (*shift_local)[3],
    const int (*border_width)[3], const double (*dh)[3][3],
    const double (*dh_inv)[3][3], grid_task_list **task_list);

void grid_collocate_task_list(const grid_task_list *task_list,
                              const enum grid_func func, const int nlevels,
                              const int (*npts_local)[3],
                              const offload_buffer *pab_blocks,
                              offload_buffer **grids);

void grid_integrate_task_list(const grid_task_list *task_list,
                              const bool compute_tau, const int natoms,
                              const int nlevels, const int (*npts_local)[3],
                              const offload_buffer *pab_blocks,
                              const offload_buffer **grids,
                              offload_buffer *hab_blocks, double (*forces)[3],
                              double (*virial)[3]);
#endif

/* src/grid/grid_task_list.h:78 -- basis sets are freed separately. */
void grid_free_task_list(grid_task_list *task_list);

#ifdef __cplusplus
}
#endif

#endif
