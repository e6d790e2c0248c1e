/*
 * grid_mi355x.h -- backend-level drop-in symbols and MI355X extensions.
 *
 * Part 1 replaces src/grid/gpu/grid_gpu_task_list.h:25-59, so the library can
 * stand in for the .cu files of src/grid/gpu inside an unmodified CP2K build whose
 * dispatcher (src/grid/grid_task_list.c) stays in place.
 *
 * Part 2 has no counterpart in the reference: precision selection, HBM-
 * resident entry points (no host copies; for callers that keep grids and
 * matrices on the device, and for measuring the kernels without PCIe), the
 * stream the library launches on, and per-kernel timing.
 */
#ifndef GRID_MI355X_H
#define GRID_MI355X_H

#include <stdbool.h>

#include "grid_task_list.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Part 1: src/grid/gpu/grid_gpu_task_list.h:25-59 -------------------- */
typedef void grid_gpu_task_list;

void grid_gpu_create_task_list(
    const bool ortho, const int ntasks, const int nlevels, const int natoms,
    const int nkinds, const int nblocks, const int *block_offsets,
    const double *atom_positions, const int *atom_kinds,
    const grid_basis_set **basis_sets, const int *level_list,
    const int *iatom_list, const int *jatom_list, const int *iset_list,
    const int *jset_list, const int *ipgf_list, const int *jpgf_list,
    const int *border_mask_list, const int *block_num_list,
    const double *radius_list, const double *rab_list, const int *npts_global,
    const int *npts_local, const int *shift_local, const int *border_width,
    const double *dh, const double *dh_inv, grid_gpu_task_list **ptr);

void grid_gpu_free_task_list(grid_gpu_task_list *ptr);

void grid_gpu_collocate_task_list(const grid_gpu_task_list *ptr,
                                  const enum grid_func func, const int nlevels,
                                  const offload_buffer *pab_blocks,
                                  offload_buffer **grids);

void grid_gpu_integrate_task_list(const grid_gpu_task_list *ptr,
                                  const bool compute_tau, const int nlevels,
                                  const offload_buffer *pab_blocks,
                                  const offload_buffer **grids,
                                  offload_buffer *hab_blocks, double *forces,
                                  double *virial);

/* ---- Part 2: extensions (no reference counterpart) ----------------------- */

/* Arithmetic of the per-point stage.  FP64: everything in double (matches
 * the reference to summation order).  FP32_MIXED: the Gaussian x polynomial
 * evaluation at each grid point runs in float (orthorhombic tasks with
 * la_max + lb_max <= 6 after the l shifts of the call: factor tables rounded
 * to float, per-column polynomials, per-point sums; integrate reads the
 * potential rounded to float; general cells: the per-point exponential in
 * float); coefficient transforms, the factor tables' computation, index math,
 * grid accumulation, hab and the force/virial reductions stay double.  Tasks
 * with larger l stay FP64.  Default FP64, or the value of environment
 * variable GRID_MI355X_PRECISION (0/1). */
enum grid_mi355x_precision {
  GRID_MI355X_FP64 = 0,
  GRID_MI355X_FP32_MIXED = 1,
};
void grid_mi355x_set_precision(int precision);
int grid_mi355x_get_precision(void);

/* HBM-resident collocate: pab_blocks and grids[level] are device pointers;
 * work is enqueued on the library stream and the call returns without
 * waiting (see grid_mi355x_synchronize). */
void grid_mi355x_collocate_task_list_dev(const grid_task_list *task_list,
                                         const enum grid_func func,
                                         const int nlevels,
                                         const double *pab_blocks,
                                         double *const *grids);

/* No counterpart in the reference.  grid_mi355x_collocate_task_list_dev for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): it calls no
 * allocator.  Its per-call buffers (level table, decontracted density blocks,
 * per-task coefficients) live in a static device array of 128 MiB that each
 * device holds for the life of the process; a list needing more is refused
 * with a message.  Calls that overlap on one device share the array and
 * race.  Not for empty or batched lists. */
void grid_mi355x_collocate_task_list_noalloc(const grid_task_list *task_list,
                                             const enum grid_func func,
                                             const int nlevels,
                                             const double *pab_blocks,
                                             double *const *grids);

/* HBM-resident integrate: hab_blocks (hab_length doubles) is overwritten;
 * forces [natoms][3] and virial [3][3] are device pointers or NULL. */
void grid_mi355x_integrate_task_list_dev(
    const grid_task_list *task_list, const bool compute_tau, const int natoms,
    const int nlevels, const double *pab_blocks, const double *const *grids,
    double *hab_blocks, const long hab_length, double *forces, double *virial);

/* HBM buffers for callers without their own device allocator (tests,
 * benchmarks).  Copies are synchronous; kind 1 host->device, 2 device->host,
 * 3 device->device.  Aborts on failure like every other entry point. */
void *grid_mi355x_device_malloc(const size_t bytes);
void grid_mi355x_device_free(void *ptr);
void grid_mi355x_memcpy(void *dst, const void *src, const size_t bytes, const int kind);

/* Stream all library work is launched on (a hipStream_t; NULL restores the
 * library's own stream). */
void grid_mi355x_set_stream(void *hip_stream);
void *grid_mi355x_get_stream(void);
void grid_mi355x_synchronize(void);

/* Per-kernel timing with HIP events on the launch stream.  Kernel ids:
 * 0 collocate (whole call: ids 6, 7, 8, 11 and 13 of it), 1 integrate (whole
 * call: ids 9, 10, 11, 13 of it), 2 decontract pab, 3 contract hab, 4 zero/fill,
 * 5 force/virial reduction, 6 cab -> cxyz pass, 7 collocation through LDS
 * tiles, 8 collocation straight into HBM, 9 integrate points, 10 cxyz ->
 * hab/forces pass, 11 per-task kernels (lp > 6), 12 rounding of the
 * potential to float (FP32-mixed integrate), 13 general-cell point kernels
 * (collocate or integrate), 14 batch graph replays. */
void grid_mi355x_set_timing(const bool enable);
void grid_mi355x_reset_timing(void);
/* Returns total milliseconds and number of launches of kernel id. */
double grid_mi355x_kernel_time_ms(const int kernel_id, long *launches);

/* Eager launches of the point kernels of kernel id (ids as above: 7, 8, 9,
 * 11, 13) since the last grid_mi355x_reset_timing, split by the arithmetic
 * of their per-point evaluation: FP64 or FP32 (FP32-mixed mode).  Kernels
 * replayed from a batch graph are not counted. */
void grid_mi355x_point_launches(const int kernel_id, long *fp64, long *fp32);

/* Number of (task, grid point) pairs the list visits, counted on the device
 * with the same membership rules as collocate/integrate. */
long long grid_mi355x_task_points(const grid_task_list *task_list);

/* Task points split by la_max + lb_max of the task (points[v] for v < n). */
void grid_mi355x_task_points_by_lp(const grid_task_list *task_list, long long *points,
                                   const int n);

/* Number of tasks that were assigned to the orthorhombic / general kernels. */
void grid_mi355x_task_counts(const grid_task_list *task_list, long *ortho,
                             long *general, long *skipped);

/* Batched task lists (many independent lists in one set of launches).
 * create: the lists (each from grid_create_task_list; they stay owned by the
 * caller and may be freed afterwards) are merged into one list whose every
 * launch covers all of them; the batch owns HBM buffers for each list:
 * pab/hab (laid out as that list's blocks, length doubles), forces [natoms][3],
 * virial [3][3], and per level a density and a potential grid.
 * run: what = 1 collocate (density from pab), 2 integrate (hab, and with
 * forces also forces and virial, from the potential), 3 both; each distinct
 * (what, func, compute_tau, forces, precision) is captured once into a
 * hipGraph and replayed.  Work is enqueued on the library stream; the call
 * returns without waiting (grid_mi355x_synchronize). */
typedef struct grid_mi355x_batch grid_mi355x_batch;
void grid_mi355x_create_batch(const int nlists, const grid_task_list *const *lists,
                              grid_mi355x_batch **batch);
void grid_mi355x_free_batch(grid_mi355x_batch *batch);
void grid_mi355x_batch_buffers(const grid_mi355x_batch *batch, const int list,
                               double **pab, double **hab, long *length,
                               double **forces, double **virial);
double *grid_mi355x_batch_grid(const grid_mi355x_batch *batch, const int list,
                               const int level, const int potential);
void grid_mi355x_batch_run(grid_mi355x_batch *batch, const int what,
                           const enum grid_func func, const bool compute_tau,
                           const bool forces);
/* Task points of all lists of the batch. */
void grid_mi355x_batch_task_points(const grid_mi355x_batch *batch, long long *points);

/* Launch shape of a batch's point kernels: enable = true runs the tile
 * collocation and the point integration as persistent launches (as many
 * workgroups as are resident at once) that take their units -- tiles of any
 * list and level, or groups of tasks, the largest first -- from a device
 * work queue (an atomic cursor, zeroed by a kernel inside the graph);
 * false (the default) launches one workgroup per unit.  Changing it drops
 * the captured graphs. */
void grid_mi355x_batch_set_work_queue(grid_mi355x_batch *batch, const bool enable);
int grid_mi355x_batch_work_queue(const grid_mi355x_batch *batch);

/* Batch autotuner: times the run (what, func, compute_tau, forces) both
 * ways (3 graph replays each) and keeps the faster; returns 1 if the work
 * queue won.  Synchronous. */
int grid_mi355x_autotune_batch(grid_mi355x_batch *batch, const int what,
                               const enum grid_func func, const bool compute_tau,
                               const bool forces);

/* Launch-shape autotuner: for every level that collocates through LDS tiles,
 * times the collocation (func, pab_blocks and grids as in
 * grid_mi355x_collocate_task_list_dev) for each workgroup count (2, 4 or 8
 * rounds per CU over the level's tile items) and cap of waves per workgroup
 * (12, 8, 6), one level after the other, and keeps the fastest in the list
 * (later collocations use it).  Synchronous.  tuned_shape reads the choice of
 * a level (general = 1: the general-cell tiles of the level). */
void grid_mi355x_autotune(const grid_task_list *task_list, const enum grid_func func,
                          const double *pab_blocks, double *const *grids);
void grid_mi355x_tuned_shape(const grid_task_list *task_list, const int level,
                             const int general, int *rounds, int *max_waves);

#ifdef __cplusplus
}
#endif

#endif
