/*
 * grid_library.h -- library lifecycle, configuration and statistics.
 *
 * Same entry points as src/grid/common/grid_library.h:22-85 (the per-thread
 * sphere cache accessor of the cpu backend has no counterpart here).
 */
#ifndef GRID_LIBRARY_H
#define GRID_LIBRARY_H

#include <stdbool.h>

#include "grid_constants.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/grid/common/grid_library.h:22 -- aborts if called twice. */
void grid_library_init(void);

/* src/grid/common/grid_library.h:28 */
void grid_library_finalize(void);

typedef struct {
  enum grid_backend backend;
  bool validate;
  bool apply_cutoff;
} grid_library_config;

/* src/grid/common/grid_library.h:44 */
void grid_library_set_config(const enum grid_backend backend,
                             const bool validate, const bool apply_cutoff);

/* src/grid/common/grid_library.h:51 */
grid_library_config grid_library_get_config(void);

/* src/grid/common/grid_library.h:57 -- prints the "GRID STATISTICS" table
 * through print_func(message, length, output_unit).  Counters are local to
 * this process (one GPU); fortran_comm is accepted and not reduced over. */
void grid_library_print_stats(const int fortran_comm,
                              void (*print_func)(const char *, int, int),
                              int output_unit);

enum grid_library_kernel {
  GRID_COLLOCATE_ORTHO = 0,
  GRID_INTEGRATE_ORTHO = 1,
  GRID_COLLOCATE_GENERAL = 2,
  GRID_INTEGRATE_GENERAL = 3,
};

/* src/grid/common/grid_library.h:83 */
void grid_library_counter_add(const int lp, const enum grid_backend backend,
                              const enum grid_library_kernel kern,
                              const int increment);

#ifdef __cplusplus
}
#endif

#endif
