/*
 * offload_buffer.h -- host/device double buffer handed across the grid API.
 *
 * Replaces src/offload/offload_buffer.h:16-38 (same struct layout, same entry
 * points).  offload_create_buffer allocates pinned host memory plus an HBM
 * mirror (src/offload/offload_buffer.c:52-100 with __OFFLOAD_HIP); results of
 * the grid calls are always in host_buffer on return.  When CP2K links its
 * own offload_buffer.c these definitions are interposed by it; the grid
 * library then uses whatever device_buffer the caller's buffer carries, or a
 * private HBM mirror when it is NULL.
 */
#ifndef OFFLOAD_BUFFER_H
#define OFFLOAD_BUFFER_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  size_t size;           /* bytes */
  double *host_buffer;   /* pinned host memory */
  double *device_buffer; /* HBM mirror, may be NULL */
} offload_buffer;

/* src/offload/offload_buffer.h:26 -- length in doubles; *buffer reused when
 * large enough, freed when length == 0. */
void offload_create_buffer(const int length, offload_buffer **buffer);

/* src/offload/offload_buffer.h:32 */
void offload_free_buffer(offload_buffer *buffer);

/* src/offload/offload_buffer.h:38 */
double *offload_get_buffer_host_pointer(offload_buffer *buffer);

#ifdef __cplusplus
}
#endif

#endif
