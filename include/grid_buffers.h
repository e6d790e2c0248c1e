/*
 * grid_buffers.h -- matrix blocks <-> grid-call buffers on the device, and
 * device residency of offload_buffers.
 *
 * The pack / unpack pair replaces the Fortran helpers that move DBCSR blocks
 * into the pab buffer before grid_collocate_task_list / after
 * grid_integrate_task_list: rs_copy_to_buffer -> rs_pack_buffer and
 * rs_copy_to_matrices -> rs_unpack_buffer (src/task_list_methods.F:2561-2674).
 * A maintainer binds them through ISO_C_BINDING with the block pointers of
 * the atom pairs (atom_pair_recv) and pair_offsets_recv.  With blocks in HBM
 * the copies run on the device and the buffer's device copy becomes the
 * current one, so the grid calls read it without a host -> device copy.
 *
 * Residency has no counterpart in the reference (offload_buffer there has no
 * state; src/offload/offload_buffer.c:52-100): a resident buffer keeps the
 * results a grid call writes into it (grids of a collocation, hab of an
 * integration) in HBM; host_buffer is brought up to date only by
 * grid_mi355x_buffer_sync_host.  Buffers never named here behave as in the
 * reference (copied in and out on every call).  A caller that writes
 * host_buffer of a buffer whose device copy is current says so with
 * grid_mi355x_buffer_host_written.
 */
#ifndef GRID_BUFFERS_H
#define GRID_BUFFERS_H

#include <stdbool.h>

#include "offload_buffer.h"

#ifdef __cplusplus
extern "C" {
#endif

/* rs_pack_buffer, src/task_list_methods.F:2597-2630: for each atom pair i,
 * dest[pair_offsets[i] .. pair_offsets[i] + block_sizes[i]) = blocks[i][..]
 * (the block column-major, as RESHAPE lays it out).  blocks is a host array
 * of pointers; the blocks are all in device memory (copied on the device;
 * dest's device copy becomes current) or all in host memory (copied into
 * host_buffer; the host copy becomes current). */
void grid_mi355x_pack_buffer(const int npairs, const double *const *blocks,
                             const int *block_sizes, const int *pair_offsets,
                             offload_buffer *dest);

/* rs_unpack_buffer, src/task_list_methods.F:2636-2674: blocks[i][..] +=
 * src[pair_offsets[i] ..] from src's current copy (blocks all in device
 * memory: on the device; all in host memory: on the host). */
void grid_mi355x_unpack_buffer(const int npairs, double *const *blocks,
                               const int *block_sizes, const int *pair_offsets,
                               const offload_buffer *src);

/* Results written into a resident buffer by the grid calls stay on the
 * device.  Needs a buffer with a device copy (this library's
 * offload_create_buffer makes one). */
void grid_mi355x_buffer_set_resident(offload_buffer *buffer, const bool resident);

/* Copies the device copy to host_buffer if the device copy is current. */
void grid_mi355x_buffer_sync_host(offload_buffer *buffer);

/* The caller wrote host_buffer: the next device read copies it. */
void grid_mi355x_buffer_host_written(offload_buffer *buffer);

/* 1 if the buffer's device copy holds its current data, else 0. */
int grid_mi355x_buffer_device_current(const offload_buffer *buffer);

#ifdef __cplusplus
}
#endif

#endif
