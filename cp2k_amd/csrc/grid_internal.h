// Internal data layout shared by the host side (grid_api.cpp) and the HIP
// kernels (grid_kernels.hip).  Everything geometric is decided once, in FP64,
// when the task list is created; a collocate/integrate call only moves the
// density/potential data and launches kernels.
#pragma once

#include <cstdint>

namespace gm {

constexpr int kMaxL = 16;            // orbital table covers l <= kMaxL
constexpr int kNumOrb = (kMaxL + 1) * (kMaxL + 2) * (kMaxL + 3) / 6;
constexpr int kTaskOutStride = 15;   // fa[3], fb[3], virial[9] per task

// Per-level grid layout, src/grid/grid_task_list.h:46-53.
struct Layout {
  int npts_global[3];
  int npts_local[3];
  int shift_local[3];
  int border_width[3];
  double dh[3][3];
  double dh_inv[3][3];
};

struct Level {
  Layout lay;
  double *grid;         // device pointer of this call's level grid
  const float *gridf;   // FP32-mixed integrate: the potential rounded to FP32
  long long npts;       // nx*ny*nz of the local grid (bounds guard)
};

enum : int { kTaskOrtho = 1, kTaskSkip = 2 };

// One task = one product of primitive Gaussians on one grid level.
// Orthorhombic tasks (orthorhombic cell and border_mask == 0) use center,
// roff, disr, lo (= lb_cube, ub = 1 - lo) and cmax; general tasks use gp,
// lo/hi (= index box) and blo/bhi (border-masked bounds of the local grid).
struct Task {
  double ra[3], rab[3], rp[3], rb[3];
  double zeta, zetb, zetp, pexp;  // pexp = exp(-zeta*zetb/zetp*|rab|^2)
  double rscale;                  // 1 on-site, 2 off-site (collocate only)
  double radius, disr;
  double roff[3];
  double gp[3];
  int center[3];
  int lo[3], hi[3];
  int blo[3], bhi[3];
  int cmax;
  int la_max, la_min, lb_max, lb_min;
  int level, flags, group;
  int o1, o2;          // Cartesian offsets of this pgf pair in the group block
  int iatom, jatom;    // zero based
  int geo;             // orthorhombic: geometry class (GeoClass index)
};

// Truncation region shared by the orthorhombic tasks of one level with the
// same discretised radius (it does not depend on the task's centre): the
// columns (i, j) of the central plane in j-major order, each packed as
// (i - lo_x) | (j - lo_y) << 8 | (kl - lo_z) << 16 with planes kl .. 1 - kl,
// and per group of 64 columns the smallest kl.
struct GeoClass {
  int col_off, ncol;   // into the column array
  int grp_off;         // into the group minimum array
  int row_off;         // into the row start array: first column of row j, j = lo_y .. 2 - lo_y
  int lo[3];
  long long npts;      // points of one task of the class
};

// Collocation tiles: a level grid is cut into tiles of nx * by * bz points
// (full rows along x) that fit into LDS.  An item is the part of one task
// whose rows j0..j1 and planes k0..k1 (cube indices) map into one tile.  A
// workgroup takes the items ibeg + first + m * stride < iend (m = 0, 1, ...)
// of one tile, so the items of a tile are dealt evenly to its workgroups.
struct TileItem {
  int task;
  short j0, j1, k0, k1;
  int pad_;
};
struct TileUnit {
  int ibeg, iend, first, stride;
  int gy0, gz0;        // tile origin in the local grid
  int level, by, bz;   // the tile's level and its rows and planes (nx * by * bz points)
};

// One (matrix block, iset, jset) sub-block: the unit of (de)contraction.
struct Group {
  long long cart_off;  // [ncob][ncoa] Cartesian block in scratch
  long long sphi_a;    // offset of sphi row sa0 of the i-kind (row stride maxcoa)
  long long sphi_b;
  int block_off;       // offset of the matrix block in the pab/hab buffer
  int transpose;       // iatom <= jatom: block stored [sgf_b][sgf_a]
  int ncoa, ncob;
  int nsa, nsb;        // spherical functions of the two sets
  int sa0, sb0;        // first spherical function of the sets (zero based)
  int nsgfa, nsgfb;    // block dimensions
  int maxcoa, maxcob;
  int iatom, jatom;
};

// LDS carve-up of the per-task kernels, decided on the host from the
// maxima over the launched tasks (offsets in doubles / ints).
struct LdsPlan {
  int np;        // stride for polynomial index: max lp + 1
  int ext;       // stride of the per-axis tables: 2*max cmax + 2
  int mapn;      // ints per map array
  int rowcap, plcap;
  int o_cab, o_alpha, o_cxyz, o_cijk, o_tab, o_cxy, o_cxrow, o_red;  // doubles
  int o_stage;   // per-column moment staging of the orthorhombic integrate
  int o_row4;    // per-row int4 {first x, length, grid row base, plane<<16 | y+32768}
  int n_doubles;
  int i_map, i_row, i_plane, i_misc;  // ints, after the doubles
  int n_ints;
  int bytes;
};

// Slots per level of the general change-of-variables matrices (lp 0..11).
constexpr int kGmatSlots = 12;

// Tasks of the lane-per-task cab -> cxyz kernel: la_max, lb_max <= 3 with
// la_max + lb_max <= 4 (larger pairs unroll into too much code).
__host__ __device__ constexpr bool lane_cxyz_task(int la, int lb) {
  return la <= 3 && lb <= 3 && la + lb <= 4;
}

// Register-array bound of the point kernels for a maximum lp.
inline int lpb_bucket(int lp) { return lp <= 4 ? 4 : lp <= 8 ? 8 : lp <= 12 ? 12 : 16; }

// Wave-per-task kernels (orthorhombic tasks of one grid level): each wave of
// a workgroup owns an LDS region of wave_bytes; collocation into a level
// whose whole grid fits into LDS keeps that grid ("window", in doubles) at
// the start of LDS and adds it to HBM once per workgroup.
constexpr int kWaveMaxExt = 256;  // cube extent per axis (8-bit packed columns)
struct WavePlan {
  int level;       // grid level of the launch
  int np;          // polynomial stride (max lp + 1)
  int ext;         // per-axis table stride (max cube extent 2 - 2 lo)
  int nwaves;      // waves per workgroup
  int wave_bytes;  // LDS bytes per wave
  int window;      // doubles of the LDS tile window (0: accumulate in HBM)
  int by, bz;      // tile rows and planes (window = nx * by * bz)
  int o_cxyz, o_u; // doubles: cxyz [np^3]; {cab, alpha} (setup/finish) or tables [3][ext][np]
  int bytes;       // dynamic LDS per workgroup
  // general-cell column walk (integrate), byte offsets in the wave region:
  // ratio tables and plane offsets (doubles, q_cap), row table of a slab of planes (rows_cap
  // ints), per row j its footprint (int4, j_cap) and column start (j_cap + 1)
  int o_q, q_cap, o_rows, rows_cap, o_jinfo, o_cstart, j_cap;
};

struct PointParams {
  const Task *tasks;
  int task_base;          // first task of the launched range
  int ntasks;             // tasks in the launched range
  const int *task_idx;    // if set: the launch covers tasks task_idx[0 .. ntasks)
  const Group *groups;
  const Level *levels;
  const double *cart;     // decontracted pab (collocate; integrate with forces)
  double *hab_cart;       // integrate: Cartesian hab accumulation
  double *task_out;       // integrate: per-task forces/virial
  long long cart_len;
  int func;
  int ld[4];              // l-range shifts: la_max, la_min, lb_max, lb_min
  int compute_tau, do_forces, do_virial;
  int count_only;         // count task points into task_out[t*15]
  LdsPlan plan;
  WavePlan wp;
  // wave-per-task kernels
  const GeoClass *geo;
  const int *geo_cols;
  const int *geo_kmin;
  double *cxyz_all;       // per task: cxyz packed (lz, ly, lx), lx fastest, lx+ly+lz <= lp
  int cxyz_stride;        // doubles per task
  const int *geo_rows;
  const TileUnit *units;  // collocation tiles
  const TileItem *items;
  int unit_base;          // first unit of the launch (one unit per workgroup)
  int lane_cxyz;          // cab -> cxyz of GRID_FUNC_AB tasks with la, lb <= 2 by lane
  int fp32;               // FP32-mixed: point evaluation in FP32, accumulation in FP64
  // general cells: per level and lp the matrix M[o][i] of the change of
  // variables cxyz (i, packed (lx, ly, lz)) -> cijk (o, packed (il, jl, kl)),
  // followed by its transpose; block (level, lp) at
  // gmat + gmat_off[level * kGmatSlots + lp]
  const double *gmat;
  const int *gmat_off;
  // single-task integration (per-task kernel, one task): if pgf_out is set,
  // forces a[3], b[3], virials a[9], b[9] go there unscaled, and the hab
  // derivatives into hdab / hadb ([cart_len][3]) and a_hdab ([cart_len][9])
  // where set
  double *pgf_out;
  double *hdab, *hadb, *a_hdab;
  // device work queue (persistent launches of batched lists): if set, the
  // launch's workgroups (waves) take units (groups of qchunk tasks) from the
  // atomic cursor *queue, zeroed before the launch, until nunits (ntasks)
  // are taken
  int *queue;
  int qchunk;
  int nunits;
};

// Task of position t of a launch.
__device__ __forceinline__ long long launch_task(const PointParams &P, long long t) {
  return P.task_idx ? (long long)P.task_idx[t] : P.task_base + t;
}

}  // namespace gm
