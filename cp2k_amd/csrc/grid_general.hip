// Wave-per-task kernels for general tasks: non-orthorhombic cells or tasks
// with a border mask (gfx950, wave64).
//
// A general task covers the lattice index box lo..hi (from the 27 corners of
// the cube around its sphere) and, in each row (j, k) of the box whose grid
// indices lie inside the border-masked local grid, the points i of the
// quadratic bound of src/grid/ref/grid_ref_collint.h:446-464.  One wave
// takes one task (or one tile item of a task).  Along a line of points the
// Gaussian follows a recurrence, e.g. along a row r(i) = r0 + i h0
// (h0 = dh[0]):
//   g(i + 1) = g(i) q(i),   q(i + 1) = q(i) exp(-2 zetp |h0|^2),
// restarted from exp() every kRestart points; a point costs two multiplies
// instead of the reference's three-table product (:358-367).
//
// Collocate: cxyz (from the cab -> cxyz pass) -> cijk, the polynomial in the
// lattice offsets (di, dj, dk) (:697-762); a lane walks whole rows, per row
// ci[il] = sum cijk dj^jl dk^kl, per point a Horner sum in di times the
// Gaussian, added into an LDS tile of the level grid (or, without tiles, 64
// rows x 8 points staged in LDS and added to HBM point-major).  Integrate: a
// lane walks a column (i, j) along k (the rows tabulated in LDS say which
// points belong to the task), per column the moments sum_k V g dk^kl folded
// into lane-private cijk accumulators, one cross-lane reduction per task,
// cijk -> cxyz, then the shared cxyz -> hab/forces pass.
//
// Compiled with -ffp-contract=off; fused multiply-adds are explicit fma().

#include <hip/hip_runtime.h>

#include "grid_device.h"
#include "grid_internal.h"

namespace gm {
void count_launch(int id, bool fp32);

namespace {

constexpr int kRestart = 32;  // points between exact exp() evaluations of the recurrence

struct GenLds {
  double *cxyz;  // dense [np][np][np]
  double *cijk;  // dense [np][np][np], then (collocate) the staging of GenStage
};

// Wave regions follow the LDS tile window (wp.window doubles, 0 without tiles).
__device__ __forceinline__ GenLds gen_lds(double *lds, const WavePlan &wp, int wave) {
  double *b = reinterpret_cast<double *>(reinterpret_cast<char *>(lds) + 8 * (size_t)wp.window +
                                         (size_t)wave * wp.wave_bytes);
  return GenLds{b + wp.o_cxyz, b + wp.o_u};
}

// Row (j, k) of the task's box: grid row base and the point range i0..i1;
// false for a row outside the border-masked grid or outside the sphere.
__device__ __forceinline__ bool gen_row(const Task &T, const Layout &lay, int j, int k, int &i0,
                                        int &i1, long long &base) {
  const int gk = wrap(k - lay.shift_local[2], lay.npts_global[2]);
  if (gk < T.blo[2] || gk > T.bhi[2]) return false;
  const int gj = wrap(j - lay.shift_local[1], lay.npts_global[1]);
  if (gj < T.blo[1] || gj > T.bhi[1]) return false;
  if (!general_row(lay.dh, T.gp, j, k, T.radius, &i0, &i1)) return false;
  // the sphere lies inside the index box; the clamp guards rounding
  i0 = max(i0, T.lo[0]);
  i1 = min(i1, T.hi[0]);
  if (i0 > i1) return false;
  base = ((long long)gk * lay.npts_local[1] + gj) * lay.npts_local[0];
  return true;
}

// Gaussian and its ratio at point i of row (j, k).
__device__ __forceinline__ void gen_gauss(const Task &T, const Layout &lay, int i, int j, int k,
                                          double &g, double &q) {
  const double di = i - T.gp[0], dj = j - T.gp[1], dk = k - T.gp[2];
  double r2 = 0.0, rh = 0.0, hh = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double r = di * lay.dh[0][d] + dj * lay.dh[1][d] + dk * lay.dh[2][d];
    r2 += r * r;
    rh += r * lay.dh[0][d];
    hh += lay.dh[0][d] * lay.dh[0][d];
  }
  g = exp(-T.zetp * r2);
  q = exp(-T.zetp * (2.0 * rh + hh));
}

// Points i0..i1 of a row as runs a..b of consecutive grid indices gx(a) ..
// gx(b) inside the border mask: gx(i) = wrap(i - shift, ngx) restarts at 0
// after ngx - 1, and only blo[0] <= gx <= bhi[0] is kept.  f(a, b, gx(a))
// per run, so that the per-point loops carry no wrap or mask test.
template <typename F>
__device__ __forceinline__ void gen_runs(const Task &T, const Layout &lay, int i0, int i1, F &&f) {
  const int ngx = lay.npts_global[0];
  const int top = min(T.bhi[0], ngx - 1);
  int i = i0, gx = wrap(i0 - lay.shift_local[0], ngx);
  while (i <= i1) {
    const int a = i + max(T.blo[0] - gx, 0);
    const int b = min(i1, i + (top - gx));
    if (a <= b) f(a, b, gx + (a - i));
    i += ngx - gx;
    gx = 0;
  }
}

// Rows r = lane, lane + 64, ... of an nj x nk block as offsets (j, k) =
// (r % nj, r / nj), advanced without a division per row.
struct RowWalk {
  int j, k, dj, dk, nj;
  __device__ __forceinline__ RowWalk(int lane, int nj_) : nj(nj_) {
    j = lane % nj;
    k = lane / nj;
    dj = 64 % nj;
    dk = 64 / nj;
  }
  __device__ __forceinline__ void next() {
    j += dj;
    k += dk;
    if (j >= nj) {
      j -= nj;
      k++;
    }
  }
};

// The task's cijk (packed: kl slowest, il fastest, written over its cxyz by
// general_cijk_kernel) -> the wave's LDS.
__device__ __forceinline__ void gen_cijk(const PointParams &P, long long task, int lp,
                                         const GenLds &W, int lane) {
  const int nc = (lp + 1) * (lp + 2) * (lp + 3) / 6;
  const double *src = P.cxyz_all + task * P.cxyz_stride;
  for (int e = lane; e < nc; e += 64) W.cijk[e] = src[e];
  wave_sync();
}

// ci[il] = sum_{jl, kl} cijk[kl][jl][il] dj^jl dk^kl over the packed cijk.
template <int LP>
__device__ __forceinline__ void gen_row_poly(const double *cijk, double dj, double dk,
                                             double *ci) {
#pragma unroll
  for (int il = 0; il <= LP; il++) ci[il] = 0.0;
  double dkp = 1.0;
  int e = 0;
#pragma unroll
  for (int kl = 0; kl <= LP; kl++) {
    double w = dkp;
#pragma unroll
    for (int jl = 0; jl <= LP - kl; jl++) {
#pragma unroll
      for (int il = 0; il <= LP - kl - jl; il++, e++) ci[il] = fma(cijk[e], w, ci[il]);
      w *= dj;
    }
    dkp *= dk;
  }
}

// Collocation staging of one wave: 64 rows x kChunk points, stored lane
// (row) major and read back point major, so that each atomic instruction
// covers kChunk consecutive points of 64 / kChunk rows instead of one point
// of 64 rows.
constexpr int kChunk = 8;
struct GenStage {
  double *val;  // [64][kChunk + 1]
  int *idx;     // [64][kChunk + 1] grid index, -1 for no point
};

__device__ __forceinline__ GenStage gen_stage(const GenLds &W, int NP) {
  GenStage g;
  g.val = W.cijk + NP * NP * NP;
  g.idx = reinterpret_cast<int *>(g.val + 64 * (kChunk + 1));
  return g;
}

template <int LP, typename R>
__device__ __forceinline__ void gen_collocate_rows(const Task &T, const Layout &lay,
                                                   const double *cijk, int NP,
                                                   const GenStage &S, double *grid, int lane) {
  const int nj = T.hi[1] - T.lo[1] + 1, nk = T.hi[2] - T.lo[2] + 1;
  const int ngx = lay.npts_global[0];
  double hh = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) hh += lay.dh[0][d] * lay.dh[0][d];
  const double r2step = exp(-2.0 * T.zetp * hh);
  for (int r0 = 0; r0 < nj * nk; r0 += 64) {
    const int r = r0 + lane;
    const int j = T.lo[1] + r % nj, k = T.lo[2] + r / nj;
    int i0 = 0, i1 = -1;
    long long base = 0;
    if (r < nj * nk && !gen_row(T, lay, j, k, i0, i1, base)) i1 = i0 - 1;
    if (r >= nj * nk) i1 = i0 - 1;
    double cd[LP + 1];
    gen_row_poly<LP>(cijk, j - T.gp[1], k - T.gp[2], cd);
    R ci[LP + 1];
#pragma unroll
    for (int il = 0; il <= LP; il++) ci[il] = (R)cd[il];
    const int len = i1 - i0 + 1;
    const int nch = -wave_min(-((len + kChunk - 1) / kChunk));  // chunks of the longest row
    int gx = wrap(i0 - lay.shift_local[0], ngx);
    const int rbase = (int)base;  // level grids hold fewer than 2^31 points
    double g = 0.0, q = 0.0, di = i0 - T.gp[0];
    for (int c = 0; c < nch; c++) {
      const int ib = i0 + c * kChunk;
      if (c % (kRestart / kChunk) == 0 && ib <= i1) gen_gauss(T, lay, ib, j, k, g, q);
#pragma unroll
      for (int s = 0; s < kChunk; s++) {
        const R dr = (R)di;
        R p = ci[LP];
#pragma unroll
        for (int il = LP - 1; il >= 0; il--) p = fmar(p, dr, ci[il]);
        const bool in = ib + s <= i1 && gx >= T.blo[0] && gx <= T.bhi[0];
        S.val[lane * (kChunk + 1) + s] = (double)(p * (R)g);
        S.idx[lane * (kChunk + 1) + s] = in ? rbase + gx : -1;
        g *= q;
        q *= r2step;
        di += 1.0;
        gx = (gx + 1 == ngx) ? 0 : gx + 1;
      }
      wave_sync();
#pragma unroll
      for (int h = 0; h < kChunk; h++) {
        const int e = (h * (64 / kChunk) + lane / kChunk) * (kChunk + 1) + lane % kChunk;
        const int x = S.idx[e];
        if (x >= 0) atomicAdd(&grid[x], S.val[e]);
      }
      wave_sync();
    }
  }
}

// Row stride (doubles) of a general LDS tile: odd, so that lanes on
// consecutive rows hit different LDS banks.
__device__ __forceinline__ int gen_tile_stride(int nx) { return nx | 1; }

// Collocate the rows j0..j1, planes k0..k1 of a task (an item) into the LDS
// tile of unit U: win[((gz - gz0) * by + gy - gy0) * sx + gx].
template <int LP, typename R>
__device__ __forceinline__ void gen_tile_rows(const Task &Tg, const Layout &layg, const double *cijk,
                                              int NP, const TileItem &I, const TileUnit &U,
                                              double *win, int lane) {
  // the task and layout in registers: the tile's atomics could otherwise
  // alias them, and every row would reload them and recompute what does not
  // depend on the row
  const Task T = Tg;
  const Layout lay = layg;
  const int nj = I.j1 - I.j0 + 1, nk = I.k1 - I.k0 + 1;
  const int sx = gen_tile_stride(lay.npts_local[0]);
  double hh = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) hh += lay.dh[0][d] * lay.dh[0][d];
  const double r2step = exp(-2.0 * T.zetp * hh);
  RowWalk rw(lane, nj);
  for (int r = lane; r < nj * nk; r += 64, rw.next()) {
    const int j = I.j0 + rw.j, k = I.k0 + rw.k;
    int i0, i1;
    long long base;
    if (!gen_row(T, lay, j, k, i0, i1, base)) continue;
    const int ly = wrap(j - lay.shift_local[1], lay.npts_global[1]) - U.gy0;
    const int lz = wrap(k - lay.shift_local[2], lay.npts_global[2]) - U.gz0;
    if ((unsigned)ly >= (unsigned)U.by || (unsigned)lz >= (unsigned)U.bz) continue;  // not for host-built items
    double *row = win + (lz * U.by + ly) * sx;
    double cd[LP + 1];
    gen_row_poly<LP>(cijk, j - T.gp[1], k - T.gp[2], cd);
    R ci[LP + 1];
#pragma unroll
    for (int il = 0; il <= LP; il++) ci[il] = (R)cd[il];
    gen_runs(T, lay, i0, i1, [&](int a, int b, int gxa) {
      double *dst = row + gxa;
      for (int s0 = a; s0 <= b; s0 += kRestart) {
        double g, q;
        gen_gauss(T, lay, s0, j, k, g, q);
        const int n = min(b - s0 + 1, kRestart);
        double di = s0 - T.gp[0];
        for (int s = 0; s < n; s++) {
          const R dr = (R)di;
          R p = ci[LP];
#pragma unroll
          for (int il = LP - 1; il >= 0; il--) p = fmar(p, dr, ci[il]);
          atomicAdd(dst + s, (double)(p * (R)g));
          g *= q;
          q *= r2step;
          di += 1.0;
        }
        dst += n;
      }
    });
  }
}

// ---------------------------------------------------------------------------
// Integrate: column walk.  A wave takes one task.  Its lanes take
// consecutive columns (i, j) of the task's footprint, x fastest, so that a
// load instruction reads consecutive grid points of one or two rows; the
// planes k are stepped together by the whole wave.  A point (i, j, k) belongs
// to the task exactly when the reference's row (j, k) holds it
// (grid_ref_collint.h:416-464 via gen_row: border mask, quadratic bounds);
// the rows of a slab of planes are tabulated in LDS first:
//   rows[k][j]   (i0 - lo0) | n << 16 for the points i0 .. i0 + n - 1
//   jinfo[j]     footprint of row j over the slab {x0, width, k_lo, k_hi}
//   cstart[j]    columns before row j (the footprint rows, j-major)
// Along a column the Gaussian follows g(k + 1) = g(k) q(k), q(k + 1) =
// q(k) exp(-2 zetp |h2|^2), with q(i, j, k) = qi[i] qj[j] qk[k] from three
// tables; g is restarted from exp() every kColRestart planes.  Per point:
// one potential value times g into the z moments M[kl] = sum V g dk^kl;
// per column the moments times di^il dj^jl into the task's cijk sums
// (lane-private registers for lp <= 6, the matrix cores for lp 7..11).
// ---------------------------------------------------------------------------
constexpr int kColRestart = 64;  // planes between exact exp() evaluations of g (FP64)
constexpr int kColBatch = 8;     // potential loads in flight per lane
constexpr int kColStage = 40;    // MFMA staging per column: di^il[16], dj^jl[12], M[12]

struct ColLds {
  double *cijk;   // [NC] the task's reduced sums (packed, il fastest)
  double *stage;  // [64 * kColStage] (lp > 6)
  double *q;      // qi[ni], qj[nj], qk[nk]
  int *rows;      // [planes of the slab][nj]
  int4 *jinfo;    // [nj]
  int *cstart;    // [nj + 1]
};

__device__ __forceinline__ ColLds col_lds(double *lds, const WavePlan &wp, int wave) {
  char *b = reinterpret_cast<char *>(lds) + (size_t)wave * wp.wave_bytes;
  ColLds w;
  w.cijk = reinterpret_cast<double *>(b) + wp.o_cxyz;
  w.stage = reinterpret_cast<double *>(b) + wp.o_u;
  w.q = reinterpret_cast<double *>(b + wp.o_q);
  w.rows = reinterpret_cast<int *>(b + wp.o_rows);
  w.jinfo = reinterpret_cast<int4 *>(b + wp.o_jinfo);
  w.cstart = reinterpret_cast<int *>(b + wp.o_cstart);
  return w;
}

// Buffer-resource load of the value at byte offset off of a grid plane of
// `bytes` bytes at `plane` (wave-uniform).  Offsets at or past `bytes` return
// zero (the hardware's range check), kPlaneOutside among them: planes are
// below 2^31 bytes.
constexpr unsigned kPlaneOutside = 0x80000000u;
constexpr int kBufferWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format

template <typename R>
__device__ __forceinline__ R plane_load(const char *plane, unsigned bytes, unsigned off) {
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(plane), (short)0, (int)bytes, kBufferWord3);
  if constexpr (sizeof(R) == 4) {
    return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
  } else {
    return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
  }
}

// One column's moments -> the task's sums, lane-private (lp <= 6):
// acc[(kl, jl, il)] += M[kl] dj^jl di^il.
template <int LP, typename R>
__device__ __forceinline__ void col_fold_regs(const R (&M)[LP + 1], double di, double dj,
                                              double *acc) {
  double pi[LP + 1], pj[LP + 1];
  pi[0] = pj[0] = 1.0;
#pragma unroll
  for (int l = 1; l <= LP; l++) {
    pi[l] = pi[l - 1] * di;
    pj[l] = pj[l - 1] * dj;
  }
  int e = 0;
#pragma unroll
  for (int kl = 0; kl <= LP; kl++)
#pragma unroll
    for (int jl = 0; jl <= LP - kl; jl++) {
      const double w = pj[jl] * (double)M[kl];
#pragma unroll
      for (int il = 0; il <= LP - kl - jl; il++, e++) acc[e] = fma(pi[il], w, acc[e]);
    }
}

typedef double col_double4 __attribute__((ext_vector_type(4)));

// One group of 64 columns -> the task's sums on the matrix cores (lp 7..11):
//   cijk[il][(jl, kl)] += sum_c di[c]^il (dj[c]^jl M[c][kl]),
// v_mfma_f64_16x16x4f64 over 4 columns per step, M dimension il (padded to
// 16), N dimension (jl, kl) in blocks of 16 (lz-major).  Operand layout:
// A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k, D[i][j] in lane
// j + 16 (i % 4), register i / 4.
template <int LP, typename R>
__device__ __forceinline__ void col_fold_mfma(const R (&M)[LP + 1], double di, double dj,
                                              double *stage, const int *qy, const int *qz,
                                              col_double4 *acc, int lane) {
  constexpr int NQ = (LP + 1) * (LP + 2) / 2;
  constexpr int NB = (NQ + 15) / 16;
  double *sc = stage + lane * kColStage;
  double p = 1.0;
#pragma unroll
  for (int l = 0; l < 16; l++) {
    sc[l] = l <= LP ? p : 0.0;
    p *= di;
  }
  p = 1.0;
#pragma unroll
  for (int l = 0; l <= LP; l++) {
    sc[16 + l] = p;
    sc[28 + l] = (double)M[l];
    p *= dj;
  }
  wave_sync();
#pragma unroll 4
  for (int kk = 0; kk < 16; kk++) {
    const double *sk = stage + (4 * kk + (lane >> 4)) * kColStage;
    const double a = sk[lane & 15];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
      const double b = qz[nb] <= LP ? sk[16 + qy[nb]] * sk[28 + qz[nb]] : 0.0;
      acc[nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[nb], 0, 0, 0);
    }
  }
  wave_sync();
}

// The task's points: slabs of planes, per slab the row table and the
// footprint, then groups of 64 columns.  fold(M, di, dj) takes each
// column's moments (every lane calls it, lanes without a column with M = 0).
template <int LP, typename R, typename Fold>
__device__ __forceinline__ void col_walk(const Task &T, const Layout &lay, const R *grid,
                                         const WavePlan &wp, const ColLds &W, int lane,
                                         Fold &&fold) {
  const int ni = T.hi[0] - T.lo[0] + 1, nj = T.hi[1] - T.lo[1] + 1, nk = T.hi[2] - T.lo[2] + 1;
  if (nj > wp.j_cap || nj > wp.rows_cap || ni + nj + 2 * nk > wp.q_cap) return;  // host-sized: never
  const int ks = min(nk, wp.rows_cap / nj);  // planes per slab
  const double *h0 = lay.dh[0], *h1 = lay.dh[1], *h2 = lay.dh[2];
  const double g02 = h0[0] * h2[0] + h0[1] * h2[1] + h0[2] * h2[2];
  const double g12 = h1[0] * h2[0] + h1[1] * h2[1] + h1[2] * h2[2];
  const double g22 = h2[0] * h2[0] + h2[1] * h2[1] + h2[2] * h2[2];
  // q(i, j, k) = exp(-zetp (2 r.h2 + h2.h2)), r.h2 = di g02 + dj g12 + dk g22
  for (int e = lane; e < ni + nj + nk; e += 64) {
    double arg;
    if (e < ni)
      arg = 2.0 * g02 * (T.lo[0] + e - T.gp[0]);
    else if (e < ni + nj)
      arg = 2.0 * g12 * (T.lo[1] + (e - ni) - T.gp[1]);
    else
      arg = 2.0 * g22 * (T.lo[2] + (e - ni - nj) - T.gp[2]) + g22;
    W.q[e] = exp(-T.zetp * arg);
  }
  // dk of every plane in R (the same for all columns: a broadcast LDS read
  // per point instead of an FP64 add and a conversion)
  R *dkt = reinterpret_cast<R *>(W.q + ni + nj + nk);
  for (int e = lane; e < nk; e += 64) dkt[e] = (R)(T.lo[2] + e - T.gp[2]);
  const double qq = exp(-2.0 * T.zetp * g22);
  const int nx = lay.npts_local[0], nxy = nx * lay.npts_local[1];
  const int ngx = lay.npts_global[0], ngy = lay.npts_global[1], ngz = lay.npts_global[2];
  // potential loads through a buffer resource per z plane (base address in
  // scalar registers, the point's byte offset in the plane per lane): a
  // point outside the task reads at kPlaneOutside, past the plane, which the
  // buffer range check returns as zero.  Buffer loads count on the vector
  // memory counter only, so a wait for a row-table read does not drain them.
  const char *gbase = reinterpret_cast<const char *>(grid);
  const unsigned plane_bytes = (unsigned)nxy * (unsigned)sizeof(R);  // < 2^31 (host check)
  for (int kb = 0; kb < nk; kb += ks) {
    const int kn = min(ks, nk - kb);
    wave_sync();  // the previous slab's readers are done
    for (int r = lane; r < nj * kn; r += 64) {
      const int jj = r % nj, kk = r / nj;
      int i0, i1;
      long long base;
      int v = 0;
      if (gen_row(T, lay, T.lo[1] + jj, T.lo[2] + kb + kk, i0, i1, base))
        v = (i0 - T.lo[0]) | ((i1 - i0 + 1) << 16);
      W.rows[r] = v;
    }
    wave_sync();
    for (int jj = lane; jj < nj; jj += 64) {
      int a = 0x7fffffff, b = -1, klo = kn, khi = -1;
      for (int kk = 0; kk < kn; kk++) {
        const int v = W.rows[kk * nj + jj];
        const int n = v >> 16;
        if (n > 0) {
          const int x0 = v & 0xffff;
          a = min(a, x0);
          b = max(b, x0 + n - 1);
          klo = min(klo, kk);
          khi = kk;
        }
      }
      W.jinfo[jj] = b >= a ? make_int4(a, b - a + 1, klo, khi) : make_int4(0, 0, kn, -1);
    }
    wave_sync();
    int ncol = 0;
    for (int j0 = 0; j0 < nj; j0 += 64) {
      const int jj = j0 + lane;
      const int w = jj < nj ? W.jinfo[jj].y : 0;
      int incl = w;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
      }
      if (jj < nj) W.cstart[jj] = ncol + incl - w;
      ncol += __shfl(incl, 63, 64);
    }
    if (lane == 0) W.cstart[nj] = ncol;
    wave_sync();
    int jj = 0;
    for (int c0 = 0; c0 < ncol; c0 += 64) {
      const int c = c0 + lane;
      while (jj + 1 < nj && W.cstart[jj + 1] <= c) jj++;
      const int4 J = W.jinfo[jj];
      const int x = min(J.x + (c - W.cstart[jj]), ni - 1);
      const int i = T.lo[0] + x, j = T.lo[1] + jj;
      const int gx = wrap(i - lay.shift_local[0], ngx);
      const int gy = wrap(j - lay.shift_local[1], ngy);
      const bool act = c < ncol && gx >= T.blo[0] && gx <= T.bhi[0];
      // wave-uniform plane range (scalar registers: the plane loop and the
      // plane base addresses stay off the vector ALU)
      const int klo = __builtin_amdgcn_readfirstlane(wave_min(act ? J.z : kn));
      const int khi = __builtin_amdgcn_readfirstlane(-wave_min(act ? -J.w : 1));
      const double di = i - T.gp[0], dj = j - T.gp[1];
      R M[LP + 1];
#pragma unroll
      for (int l = 0; l <= LP; l++) M[l] = R(0);
      if (klo <= khi) {
        const double qij = W.q[x] * W.q[ni + jj];
        const unsigned cb = (unsigned)(gy * nx + gx) * (unsigned)sizeof(R);  // bytes into a plane
        // the Gaussian recurrence stays FP64 in both modes (in FP32 its
        // relative drift grows with the square of the planes walked)
        for (int k0 = klo; k0 <= khi; k0 += kColRestart) {
          const int k1 = min(khi, k0 + kColRestart - 1);
          const int kabs = T.lo[2] + kb + k0;
          const double dk0 = kabs - T.gp[2];
          double r2 = 0.0;
#pragma unroll
          for (int d = 0; d < 3; d++) {
            const double rd = di * h0[d] + dj * h1[d] + dk0 * h2[d];
            r2 += rd * rd;
          }
          double g = exp(-T.zetp * r2);
          double q = qij * W.q[ni + nj + kb + k0];
          int gz = wrap(kabs - lay.shift_local[2], ngz);
          const char *plane = gbase + (size_t)gz * plane_bytes;
          const int *rowk = W.rows + k0 * nj + jj;
          const R *dkp = dkt + kb + k0;
          auto step = [&](R val, R dkr) {
            R t = val * (R)g;
#pragma unroll
            for (int l = 0; l <= LP; l++) {
              M[l] += t;
              t *= dkr;
            }
            g *= q;
            q *= qq;
          };
          // row-table entry v -> is the column's point in the row; the table
          // read is unconditional (every lane's rowk lies inside the table),
          // so that no branch and no wait separates the loads of a batch
          auto inside = [&](int v) {
            return act & ((unsigned)(x - (v & 0xffff)) < (unsigned)(v >> 16));
          };
          // NB planes at a time, their loads issued together with no branch
          // between them (a point outside the row reads the plane's first
          // value and drops it): batches of kColBatch, then at most one
          // batch each of 4, 2 and 1 for the planes left
          int kk = k0;
          auto batch = [&](auto nb) {
            constexpr int NB = decltype(nb)::value;
            R val[NB];
#pragma unroll
            for (int u = 0; u < NB; u++) {
              val[u] = plane_load<R>(plane, plane_bytes, inside(rowk[u * nj]) ? cb : kPlaneOutside);
              if (++gz == ngz) {
                gz = 0;
                plane = gbase;
              } else {
                plane += plane_bytes;
              }
            }
            rowk += NB * nj;
#pragma unroll
            for (int u = 0; u < NB; u++) step(val[u], dkp[u]);
            dkp += NB;
            kk += NB;
          };
          while (kk + kColBatch - 1 <= k1) batch(std::integral_constant<int, kColBatch>());
          if (kk + 3 <= k1) batch(std::integral_constant<int, 4>());
          if (kk + 1 <= k1) batch(std::integral_constant<int, 2>());
          if (kk <= k1) batch(std::integral_constant<int, 1>());
        }
      }
      fold(M, act ? di : 0.0, act ? dj : 0.0);
    }
  }
}

// cxyz = M^T cijk (the adjoint of the change of variables), one output per
// lane, packed to the task's cxyz slot for the cxyz -> hab pass.
__device__ __forceinline__ void col_cijk_to_cxyz(const PointParams &P, const Task &T, int lp,
                                                 const double *cijk, long long task, int lane) {
  const int nc = (lp + 1) * (lp + 2) * (lp + 3) / 6;
  const double *M = P.gmat + P.gmat_off[T.level * kGmatSlots + lp];
  double *out = P.cxyz_all + task * P.cxyz_stride;
  for (int i = lane; i < nc; i += 64) {
    double s = 0.0;
    for (int o = 0; o < nc; o++) s = fma(M[o * nc + i], cijk[o], s);
    out[i] = s;
  }
}

template <int LP, typename R>
__device__ __forceinline__ void col_integrate_task(const PointParams &P, const Task &T,
                                                   const Layout &lay, const R *grid,
                                                   const ColLds &W, long long task, int lane) {
  constexpr int NC = (LP + 1) * (LP + 2) * (LP + 3) / 6;
  if constexpr (LP <= 6) {
    double acc[NC];
#pragma unroll
    for (int e = 0; e < NC; e++) acc[e] = 0.0;
    col_walk<LP, R>(T, lay, grid, P.wp, W, lane,
                    [&](const R(&M)[LP + 1], double di, double dj) {
                      col_fold_regs<LP, R>(M, di, dj, acc);
                    });
    wave_sum_store<NC>(acc, W.cijk, lane);
  } else {
    constexpr int NQ = (LP + 1) * (LP + 2) / 2;
    constexpr int NB = (NQ + 15) / 16;
    int qy[NB], qz[NB];  // (jl, kl) of this lane's B column in each block
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
      int n = 16 * nb + (lane & 15), z = 0;
      while (z <= LP && n >= LP - z + 1) {
        n -= LP - z + 1;
        z++;
      }
      qy[nb] = n;
      qz[nb] = z;  // z > LP: padding column
    }
    col_double4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) acc[nb] = col_double4{0.0, 0.0, 0.0, 0.0};
    col_walk<LP, R>(T, lay, grid, P.wp, W, lane,
                    [&](const R(&M)[LP + 1], double di, double dj) {
                      col_fold_mfma<LP, R>(M, di, dj, W.stage, qy, qz, acc, lane);
                    });
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int il = 4 * r + (lane >> 4);
        if (qz[nb] <= LP && il <= LP - qy[nb] - qz[nb])
          W.cijk[packed_index(il, qy[nb], qz[nb], LP)] = acc[nb][r];
      }
  }
  wave_sync();
  col_cijk_to_cxyz(P, T, LP, W.cijk, task, lane);
}

}  // namespace

// cxyz -> cijk of general tasks with lp <= 6, in place (packed, kl slowest,
// il fastest): cijk[o] = sum_i MT[i][o] cxyz[i] with the level's transposed
// matrix (grid_ref_collint.h:697-762 as a matrix-vector product).  One task
// per wave; every read of the cxyz precedes the wave barrier, every write
// follows it.
__global__ __launch_bounds__(256) void general_cijk_kernel(const Task *__restrict__ tasks,
                                                           PointParams P) {
  __shared__ double cx[4][84];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long t = (long long)blockIdx.x * 4 + wave;
  if (t >= P.ntasks) return;
  const long long task = launch_task(P, t);
  const Task &T = tasks[task];
  const int lp = T.la_max + T.lb_max + P.ld[0] + P.ld[2];
  if (lp > 6) return;  // the per-task kernels take these
  const int nc = (lp + 1) * (lp + 2) * (lp + 3) / 6;
  double *io = P.cxyz_all + task * P.cxyz_stride;
  for (int e = lane; e < nc; e += 64) cx[wave][e] = io[e];
  wave_sync();
  const double *MT = P.gmat + P.gmat_off[T.level * kGmatSlots + lp] + nc * nc;
  for (int o = lane; o < nc; o += 64) {
    double s = 0.0;
    for (int i = 0; i < nc; i++) s = fma(MT[i * nc + o], cx[wave][i], s);
    io[o] = s;
  }
}

// The same change of variables with a task per lane, for lp <= 4 (the
// matrix rows are the same for every lane of a level and lp: broadcast
// loads); lanes of larger lp return (general_cijk_kernel takes them).
template <int LP>
__device__ __forceinline__ void lane_cijk(const PointParams &P, const Task &T, double *io) {
  constexpr int NC = (LP + 1) * (LP + 2) * (LP + 3) / 6;
  double cx[NC];
#pragma unroll
  for (int i = 0; i < NC; i++) cx[i] = io[i];
  const double *MT = P.gmat + P.gmat_off[T.level * kGmatSlots + LP] + NC * NC;
#pragma unroll
  for (int o = 0; o < NC; o++) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NC; i++) s = fma(MT[i * NC + o], cx[i], s);
    io[o] = s;
  }
}

__global__ __launch_bounds__(256) void general_cijk_lane_kernel(const Task *__restrict__ tasks,
                                                                PointParams P) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P.ntasks) return;
  const long long task = launch_task(P, t);
  const Task &T = tasks[task];
  const int lp = T.la_max + T.lb_max + P.ld[0] + P.ld[2];
  double *io = P.cxyz_all + task * P.cxyz_stride;
  switch (lp) {
    case 0: lane_cijk<0>(P, T, io); break;
    case 1: lane_cijk<1>(P, T, io); break;
    case 2: lane_cijk<2>(P, T, io); break;
    case 3: lane_cijk<3>(P, T, io); break;
    case 4: lane_cijk<4>(P, T, io); break;
    default: break;
  }
}

// Collocation of general tasks, one task per wave, into the HBM level grids.
template <int LPB, typename R>
__global__ __launch_bounds__(256) void general_collocate_kernel(const Task *__restrict__ tasks,
                                                                const Level *__restrict__ levels,
                                                                PointParams P) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long t = (long long)blockIdx.x * (blockDim.x >> 6) + wave;
  if (t >= P.ntasks) return;
  const long long task = launch_task(P, t);
  const Task &T = tasks[task];
  const Level &L = levels[T.level];
  const Layout &lay = L.lay;
  const int NP = P.wp.np;
  const GenLds W = gen_lds(lds, P.wp, wave);
  const int lp = T.la_max + T.lb_max + P.ld[0] + P.ld[2];
  gen_cijk(P, task, lp, W, lane);
  const GenStage S = gen_stage(W, NP);
  switch (lp) {
    case 0: gen_collocate_rows<0, R>(T, lay, W.cijk, NP, S, L.grid, lane); break;
    case 1: gen_collocate_rows<1, R>(T, lay, W.cijk, NP, S, L.grid, lane); break;
    case 2: gen_collocate_rows<2, R>(T, lay, W.cijk, NP, S, L.grid, lane); break;
    case 3: gen_collocate_rows<3, R>(T, lay, W.cijk, NP, S, L.grid, lane); break;
    case 4: gen_collocate_rows<4, R>(T, lay, W.cijk, NP, S, L.grid, lane); break;
    case 5: if (LPB >= 5) gen_collocate_rows<(LPB >= 5 ? 5 : 0), R>(T, lay, W.cijk, NP, S, L.grid, lane); break;
    case 6: if (LPB >= 6) gen_collocate_rows<(LPB >= 6 ? 6 : 0), R>(T, lay, W.cijk, NP, S, L.grid, lane); break;
    default: break;  // the host routes larger lp to the per-task kernels
  }
}

// Collocation of general tasks through LDS tiles: workgroup = unit
// P.unit_base + blockIdx.x, the items U.ibeg + U.first + m * U.stride < U.iend
// of one tile of level U.level; the tile is added to HBM once.
template <int LPB, typename R>
__global__ __launch_bounds__(LPB <= 4 ? 1024 : 512) void general_tile_kernel(
    const Task *__restrict__ tasks, const Level *__restrict__ levels,
    const TileUnit *__restrict__ units, const TileItem *__restrict__ items, PointParams P) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  const TileUnit U = units[P.unit_base + blockIdx.x];
  const Level &L = levels[U.level];
  const Layout &lay = L.lay;
  const int nx = lay.npts_local[0], ny = lay.npts_local[1], nz = lay.npts_local[2];
  const int sx = gen_tile_stride(nx);
  const int window = sx * U.by * U.bz;  // <= P.wp.window
  for (int e = threadIdx.x; e < window; e += blockDim.x) lds[e] = 0.0;
  __syncthreads();
  const int NP = P.wp.np;
  const GenLds W = gen_lds(lds, P.wp, wave);
  for (long long it = U.ibeg + U.first + (long long)wave * U.stride; it < U.iend;
       it += (long long)nw * U.stride) {
    const TileItem I = items[it];
    const Task &T = tasks[I.task];
    const int lp = T.la_max + T.lb_max + P.ld[0] + P.ld[2];
    gen_cijk(P, I.task, lp, W, lane);
    switch (lp) {
      case 0: gen_tile_rows<0, R>(T, lay, W.cijk, NP, I, U, lds, lane); break;
      case 1: gen_tile_rows<1, R>(T, lay, W.cijk, NP, I, U, lds, lane); break;
      case 2: gen_tile_rows<2, R>(T, lay, W.cijk, NP, I, U, lds, lane); break;
      case 3: gen_tile_rows<3, R>(T, lay, W.cijk, NP, I, U, lds, lane); break;
      case 4: gen_tile_rows<4, R>(T, lay, W.cijk, NP, I, U, lds, lane); break;
      case 5: if (LPB >= 5) gen_tile_rows<(LPB >= 5 ? 5 : 0), R>(T, lay, W.cijk, NP, I, U, lds, lane); break;
      case 6: if (LPB >= 6) gen_tile_rows<(LPB >= 6 ? 6 : 0), R>(T, lay, W.cijk, NP, I, U, lds, lane); break;
      default: break;  // the host routes larger lp to the per-task kernels
    }
    wave_sync();
  }
  __syncthreads();
  // one tile row per wave and pass: coalesced adds, no division per point
  for (int r = wave; r < U.by * U.bz; r += nw) {
    const int y = U.gy0 + r % U.by, z = U.gz0 + r / U.by;
    if (y >= ny || z >= nz) continue;
    double *dst = L.grid + ((long long)z * ny + y) * nx;
    const double *src = lds + r * sx;
    for (int x = lane; x < nx; x += 64) {
      const double v = src[x];
      if (v != 0.0) atomicAdd(&dst[x], v);
    }
  }
}

// Integration of general tasks, one task per wave (the column walk); writes
// the task's cxyz (packed) for the cxyz -> hab pass.  R = float reads
// Level::gridf and sums each column's moments in FP32.  The kernel covers
// lp LO..HI: one exact lp keeps the register count (lane sums of lp <= 6)
// at what that lp needs; LO >= 7 sums on the matrix cores.  FP64 asks for 4
// waves per SIMD up to lp 4 and 3 at lp 5 (a few spilled registers, measured
// faster); FP32 keeps its registers (there the spills only added traffic).
template <int LO, int HI, typename R>
__global__ __launch_bounds__(256, (sizeof(R) == 8 ? (HI <= 4 ? 4 : (HI == 5 ? 3 : 2)) : 1)) void general_integrate_kernel(const Task *__restrict__ tasks,
                                                                const Level *__restrict__ levels,
                                                                PointParams P) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long t = (long long)blockIdx.x * (blockDim.x >> 6) + wave;
  if (t >= P.ntasks) return;
  const long long task = launch_task(P, t);
  const Task &T = tasks[task];
  const Level &L = levels[T.level];
  const Layout &lay = L.lay;
  const ColLds W = col_lds(lds, P.wp, wave);
  const int lp = T.la_max + T.lb_max + P.ld[0] + P.ld[2];
  const R *grid;
  if constexpr (sizeof(R) == 4)
    grid = L.gridf;
  else
    grid = L.grid;
#define GEN_INT_CASE(V)                                                            \
  case V:                                                                          \
    if constexpr (V >= LO && V <= HI) col_integrate_task<V, R>(P, T, lay, grid, W, task, lane); \
    break;
  switch (lp) {
    GEN_INT_CASE(0) GEN_INT_CASE(1) GEN_INT_CASE(2) GEN_INT_CASE(3) GEN_INT_CASE(4)
    GEN_INT_CASE(5) GEN_INT_CASE(6) GEN_INT_CASE(7) GEN_INT_CASE(8) GEN_INT_CASE(9)
    GEN_INT_CASE(10) GEN_INT_CASE(11)
    default: break;  // the host routes larger lp to the per-task kernels
  }
#undef GEN_INT_CASE
}

template <typename K>
static hipError_t launch_gen(K kernel, const PointParams &P, hipStream_t s) {
  const size_t lds = (size_t)P.wp.bytes;
  if (lds > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  const int blocks = (P.ntasks + P.wp.nwaves - 1) / P.wp.nwaves;
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(64 * P.wp.nwaves), lds, s, P.tasks, P.levels, P);
  return hipGetLastError();
}

// General tasks [task_base, task_base + ntasks) with lp <= lpb (4 or 6);
// P.wp: np = lp bound + 1, two dense np^3 arrays per wave.
hipError_t launch_general_cijk(const PointParams &P, hipStream_t s) {
  if (P.ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(general_cijk_kernel, dim3((P.ntasks + 3) / 4), dim3(256), 0, s, P.tasks, P);
  return hipGetLastError();
}

// General tasks with lp <= 4 (the others in the range are left alone).
hipError_t launch_general_cijk_lane(const PointParams &P, hipStream_t s) {
  if (P.ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(general_cijk_lane_kernel, dim3((P.ntasks + 255) / 256), dim3(256), 0, s,
                     P.tasks, P);
  return hipGetLastError();
}

hipError_t launch_general_collocate(const PointParams &P, int lpb, hipStream_t s) {
  if (P.ntasks <= 0) return hipSuccess;
  if (P.wp.nwaves < 1 || P.wp.nwaves > 4 || P.wp.np > lpb + 1) return hipErrorInvalidValue;
  count_launch(13, P.fp32);
  if (lpb == 4)
    return P.fp32 ? launch_gen(general_collocate_kernel<4, float>, P, s)
                  : launch_gen(general_collocate_kernel<4, double>, P, s);
  if (lpb == 6)
    return P.fp32 ? launch_gen(general_collocate_kernel<6, float>, P, s)
                  : launch_gen(general_collocate_kernel<6, double>, P, s);
  return hipErrorInvalidValue;
}

// nunits workgroups from unit P.unit_base; P.wp.window: the largest tile.
hipError_t launch_general_tiles(const PointParams &P, int lpb, int nunits, hipStream_t s) {
  if (nunits <= 0) return hipSuccess;
  if (P.wp.nwaves < 1 || P.wp.nwaves > (lpb <= 4 ? 16 : 8) || P.wp.window <= 0 ||
      P.wp.np > lpb + 1)
    return hipErrorInvalidValue;
  count_launch(13, P.fp32);
  auto go = [&](auto kernel) {
    const size_t lds = (size_t)P.wp.bytes;
    if (lds > 65536) {
      const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kernel, dim3(nunits), dim3(64 * P.wp.nwaves), lds, s, P.tasks, P.levels,
                       P.units, P.items, P);
    return hipGetLastError();
  };
  if (lpb == 4) return P.fp32 ? go(general_tile_kernel<4, float>) : go(general_tile_kernel<4, double>);
  if (lpb == 6) return P.fp32 ? go(general_tile_kernel<6, float>) : go(general_tile_kernel<6, double>);
  return hipErrorInvalidValue;
}

// General tasks with lp lo..hi: (v, v) for v <= 6, (0, 4), (5, 6), (7, 11).
hipError_t launch_general_integrate(const PointParams &P, int lo, int hi, hipStream_t s) {
  if (P.ntasks <= 0) return hipSuccess;
  if (P.wp.nwaves < 1 || P.wp.nwaves > 4 || P.wp.rows_cap < 1 || P.wp.j_cap < 1 ||
      P.wp.np > hi + 1)
    return hipErrorInvalidValue;
  count_launch(13, P.fp32);
  auto go = [&](auto kd, auto kf) { return P.fp32 ? launch_gen(kf, P, s) : launch_gen(kd, P, s); };
#define GEN_INT_RANGE(A, B)                                                          \
  if (lo == A && hi == B)                                                           \
    return go(general_integrate_kernel<A, B, double>, general_integrate_kernel<A, B, float>);
  GEN_INT_RANGE(0, 0) GEN_INT_RANGE(1, 1) GEN_INT_RANGE(2, 2) GEN_INT_RANGE(3, 3)
  GEN_INT_RANGE(4, 4) GEN_INT_RANGE(5, 5) GEN_INT_RANGE(6, 6)
  GEN_INT_RANGE(0, 4) GEN_INT_RANGE(5, 6) GEN_INT_RANGE(7, 11)
#undef GEN_INT_RANGE
  return hipErrorInvalidValue;
}

}  // namespace gm
