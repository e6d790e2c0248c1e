/*
 * grid_constants.h -- function and backend identifiers of the grid API.
 *
 * Values are part of the ABI: Fortran passes them as plain integers
 * (src/grid/grid_api.F:981-991, :1171-1177).  Same enumerators as
 * src/grid/common/grid_constants.h:10-54.
 */
#ifndef GRID_CONSTANTS_H
#define GRID_CONSTANTS_H

/* What is collocated for each density element (src/grid/common/
 * grid_prepare_pab.h documents each operator). */
enum grid_func {
  GRID_FUNC_AB = 100,
  GRID_FUNC_DADB = 200,
  GRID_FUNC_ADBmDAB_X = 301,
  GRID_FUNC_ADBmDAB_Y = 302,
  GRID_FUNC_ADBmDAB_Z = 303,
  GRID_FUNC_ARDBmDARB_XX = 411,
  GRID_FUNC_ARDBmDARB_XY = 412,
  GRID_FUNC_ARDBmDARB_XZ = 413,
  GRID_FUNC_ARDBmDARB_YX = 421,
  GRID_FUNC_ARDBmDARB_YY = 422,
  GRID_FUNC_ARDBmDARB_YZ = 423,
  GRID_FUNC_ARDBmDARB_ZX = 431,
  GRID_FUNC_ARDBmDARB_ZY = 432,
  GRID_FUNC_ARDBmDARB_ZZ = 433,
  GRID_FUNC_DABpADB_X = 501,
  GRID_FUNC_DABpADB_Y = 502,
  GRID_FUNC_DABpADB_Z = 503,
  GRID_FUNC_DX = 601,
  GRID_FUNC_DY = 602,
  GRID_FUNC_DZ = 603,
  GRID_FUNC_DXDY = 701,
  GRID_FUNC_DYDZ = 702,
  GRID_FUNC_DZDX = 703,
  GRID_FUNC_DXDX = 801,
  GRID_FUNC_DYDY = 802,
  GRID_FUNC_DZDZ = 803,
  GRID_FUNC_DAB_X = 901,
  GRID_FUNC_DAB_Y = 902,
  GRID_FUNC_DAB_Z = 903,
  GRID_FUNC_ADB_X = 904,
  GRID_FUNC_ADB_Y = 905,
  GRID_FUNC_ADB_Z = 906,
  GRID_FUNC_CORE_X = 1001,
  GRID_FUNC_CORE_Y = 1002,
  GRID_FUNC_CORE_Z = 1003,
};

/* AUTO and GPU both select this library's MI355X backend; REF, CPU and DGEMM
 * are CP2K's host backends, which this library does not provide (selecting
 * them aborts with a message instead of silently running elsewhere). */
enum grid_backend {
  GRID_BACKEND_AUTO = 10,
  GRID_BACKEND_REF = 11,
  GRID_BACKEND_CPU = 12,
  GRID_BACKEND_DGEMM = 13,
  GRID_BACKEND_GPU = 14,
};

#endif
