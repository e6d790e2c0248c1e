// Wave-per-task kernels for orthorhombic tasks (gfx950, wave64).
//
// One wave processes one task at a time, with no workgroup barriers on the
// task path; task records are read through __restrict__ pointers into scalar
// registers.  A collocate or integrate call runs three kernels:
//
//   cxyz_wave_kernel      cab (prepare_pab of the decontracted density) ->
//                         cxyz about rp (src/grid/ref/grid_ref_collint.h
//                         :877-910), stored per task in HBM (collocate)
//   *_points kernels      the truncation region, point by point
//   finish_wave_kernel    cxyz -> cab -> hab, forces, virial (integrate;
//                         grid_ref_integrate.c:132-171)
//
// Point walk.  The truncation region of a task depends only on its level and
// discretised radius, so the host builds it once per such geometry class:
// the columns (i, j) of the central plane, each with its plane interval
// [kl, 1 - kl], computed with the reference's floating-point bounds
// (grid_ref_collint.h:139-167).  Lanes take consecutive columns (x fastest,
// so loads and adds are contiguous along x) and step through the planes k
// together, reading per-axis tables (c h - roff)^l exp(-zetp (c h - roff)^2)
// from the wave's LDS region.
//
// Collocate: per column D[lz] = sum_{lx,ly} cxyz px py, then per point
// (lp + 1) fused multiply-adds and one atomic add, into an LDS tile of the
// level grid (full rows along x) that a workgroup fills from many tasks and
// adds to HBM once; each task contributes the rows and planes of its cube
// that fall into the tile (an item).  Levels with rows too long for any tile
// add straight into HBM.  Integrate: per point (lp + 1) multiply-adds of the z
// moments, per column their outer product with px, py into lane-private cxyz
// accumulators, one cross-lane reduction per task.
//
// Compiled with -ffp-contract=off; fused multiply-adds are explicit fma().

#include <hip/hip_runtime.h>

#include "grid_device.h"
#include "grid_internal.h"

namespace gm {
void count_launch(int id, bool fp32);

namespace {

// Per-wave LDS region.
struct WaveLds {
  double *cxyz;  // dense [np][np][np] (setup, finish) or packed (collocate points)
  double *u;     // {cab, alpha} (setup/finish) or per-axis tables (point kernels)
};

__device__ __forceinline__ WaveLds wave_lds(double *lds, const WavePlan &wp, int wave) {
  char *b = reinterpret_cast<char *>(lds) + 8 * (size_t)wp.window + (size_t)wave * wp.wave_bytes;
  WaveLds w;
  w.cxyz = reinterpret_cast<double *>(b) + wp.o_cxyz;
  w.u = reinterpret_cast<double *>(b) + wp.o_u;
  return w;
}

// Per-axis tables, tab[(d * ext + c - lo_d) * np + l], computed in FP64 and
// stored in the evaluation type R.
template <typename R>
__device__ void wave_tables(const Task &T, const Layout &lay, int lp, const WavePlan &wp,
                            R *tab, int lane) {
  const int n0 = 2 - 2 * T.lo[0], n1 = 2 - 2 * T.lo[1], n2 = 2 - 2 * T.lo[2];
  for (int e = lane; e < n0 + n1 + n2; e += 64) {
    const int d = e < n0 ? 0 : (e < n0 + n1 ? 1 : 2);
    const int q = e - (d == 0 ? 0 : (d == 1 ? n0 : n0 + n1));
    const int lo = d == 0 ? T.lo[0] : (d == 1 ? T.lo[1] : T.lo[2]);
    const double h = d == 0 ? lay.dh[0][0] : (d == 1 ? lay.dh[1][1] : lay.dh[2][2]);
    const double roff = d == 0 ? T.roff[0] : (d == 1 ? T.roff[1] : T.roff[2]);
    const int c = lo + q;
    const double t = c * h - roff;
    double v = exp(-T.zetp * t * t);
    R *row = tab + (d * wp.ext + q) * wp.np;
    for (int l = 0; l <= lp; l++) {
      row[l] = (R)v;
      v *= t;
    }
  }
  wave_sync();
}

// Column c0 + lane of the task's geometry class.
struct Column {
  int i, j;  // cube indices
  int kl;    // planes kl .. 1 - kl; kl = 1 marks a lane without a column
  int base;  // y * nx + x of the column in the local grid
  int gy;    // y of the column in the local grid
};

__device__ __forceinline__ Column load_column(const PointParams &P, const GeoClass &G,
                                              const Layout &lay, const int cw[3], int c) {
  Column C{0, 0, 1, 0, 0};
  if (c >= G.ncol) return C;
  const int rec = P.geo_cols[G.col_off + c];
  C.i = G.lo[0] + (rec & 255);
  C.j = G.lo[1] + ((rec >> 8) & 255);
  const int gx = wrap(cw[0] + C.i, lay.npts_global[0]);
  const int gy = wrap(cw[1] + C.j, lay.npts_global[1]);
  // a cube that fits the local grid (checked on the host) never fails this
  if (gx < lay.npts_local[0] && gy < lay.npts_local[1]) {
    C.kl = G.lo[2] + (rec >> 16);
    C.base = gy * lay.npts_local[0] + gx;
    C.gy = gy;
  }
  return C;
}

template <int LP, typename R>
__device__ __forceinline__ void load_pxy(const Task &T, const WavePlan &wp, const R *tab,
                                         const Column &C, R px[LP + 1], R py[LP + 1]) {
  const int np = wp.np;
  const R *rx = tab + clampi(C.i - T.lo[0], 0, wp.ext - 1) * np;
  const R *ry = tab + (wp.ext + clampi(C.j - T.lo[1], 0, wp.ext - 1)) * np;
#pragma unroll
  for (int l = 0; l <= LP; l++) {
    px[l] = rx[l];
    py[l] = ry[l];
  }
}

// cab (prepare_pab of the decontracted density) -> dense cxyz in the wave's LDS.
__device__ void wave_cxyz(const PointParams &P, const Task &T, const Group &G, int la, int lb,
                          int la_min, int lb_min, int NP, const WaveLds &W, int lane) {
  const int n1 = ncoset(la), n2 = ncoset(lb);
  double *cab = W.u, *alpha = W.u + n1 * n2;
  for (int e = lane; e < n1 * n2; e += 64) cab[e] = 0.0;
  for (int e = lane; e < NP * NP * NP; e += 64) W.cxyz[e] = 0.0;
  wave_sync();
  const int a0 = ncoset(T.la_min - 1), na = ncoset(T.la_max) - a0;
  const int b0 = ncoset(T.lb_min - 1), nb = ncoset(T.lb_max) - b0;
  const CabRef cr{cab, n1};
  for (int e = lane; e < na * nb; e += 64) {
    const int ia = a0 + e % na, jb = b0 + e / na;
    const long long q = G.cart_off + (long long)(T.o2 + jb) * G.ncoa + T.o1 + ia;
    const double p = (q < P.cart_len) ? P.cart[q] : 0.0;
    if (P.func == 100)
      cab[jb * n1 + ia] = p;
    else
      apply_func(P.func, orb(ia), orb(jb), T.zeta, T.zetb, p, cr);
  }
  build_alpha(T, la, lb, alpha, lane, 64);
  wave_sync();
  const int np = la + lb + 1, nA = la + 1, nB = lb + 1;
  const double pref = T.rscale * T.pexp;
  const int a1 = ncoset(la_min - 1), ma = ncoset(la) - a1;
  const int b1 = ncoset(lb_min - 1), mb = ncoset(lb) - b1;
  for (int e = lane; e < ma * mb; e += 64) {
    const int ia = a1 + e % ma, jb = b1 + e / ma;
    const double cv = cab[jb * n1 + ia];
    if (cv == 0.0) continue;
    const Ang A = orb(ia), B = orb(jb);
    const double *ax = &alpha[((0 * nB + B.x) * nA + A.x) * np];
    const double *ay = &alpha[((1 * nB + B.y) * nA + A.y) * np];
    const double *az = &alpha[((2 * nB + B.z) * nA + A.z) * np];
    // the (lz, ly, lx) terms of the pair, started at a lane-dependent term
    // and wrapped around: lanes that all began at (0, 0, 0) would add to the
    // same LDS address in the same instruction
    const int nxl = A.x + B.x + 1, nyl = A.y + B.y + 1, nzl = A.z + B.z + 1;
    const int cnt = nxl * nyl * nzl;
    int m0 = lane % cnt;
    int lx = m0 % nxl, ly = (m0 / nxl) % nyl, lz = m0 / (nxl * nyl);
    for (int m = 0; m < cnt; m++) {
      atomicAdd(&W.cxyz[(lz * NP + ly) * NP + lx], cv * (ax[lx] * ay[ly] * az[lz] * pref));
      if (++lx == nxl) {
        lx = 0;
        if (++ly == nyl) {
          ly = 0;
          if (++lz == nzl) lz = 0;
        }
      }
    }
  }
  wave_sync();
}

__device__ __forceinline__ void axis_offsets(const Task &T, const Layout &lay, int cw[3]) {
  cw[0] = pm(T.center[0] - lay.shift_local[0], lay.npts_global[0]);
  cw[1] = pm(T.center[1] - lay.shift_local[1], lay.npts_global[1]);
  cw[2] = pm(T.center[2] - lay.shift_local[2], lay.npts_global[2]);
}

// High lp (7, 8): the same sums with the ly loop kept rolled and py read
// from the column's LDS table row ry, which keeps the register count below
// the limit of the tile kernels.
template <int LP, typename R>
__device__ __forceinline__ void column_poly_ry(const R *cx, const R *px, const R *ry, R *D) {
  int e = 0;
#pragma unroll
  for (int lz = 0; lz <= LP; lz++) {
    R s = 0;
#pragma unroll 1
    for (int ly = 0; ly <= LP - lz; ly++) {
      R t = 0;
#pragma unroll
      for (int lx = 0; lx <= LP - lz; lx++)
        if (lx <= LP - lz - ly) t = fmar(cx[e + lx], px[lx], t);
      e += LP - lz - ly + 1;
      s = fmar(t, ry[ly], s);
    }
    D[lz] = s;
  }
}

// D[lz] = sum_ly py[ly] sum_lx cxyz[(lz, ly, lx)] px[lx] of one column.
template <int LP, typename R>
__device__ __forceinline__ void column_poly(const R *cx, const R *px, const R *py, R *D) {
  int e = 0;
#pragma unroll
  for (int lz = 0; lz <= LP; lz++) {
    R s = 0;
#pragma unroll
    for (int ly = 0; ly <= LP - lz; ly++) {
      R t = 0;
#pragma unroll
      for (int lx = 0; lx <= LP - lz - ly; lx++, e++) t = fmar(cx[e], px[lx], t);
      s = fmar(t, py[ly], s);
    }
    D[lz] = s;
  }
}

// Collocate the task's points straight into the HBM level grid.
template <int LP, typename R>
__device__ __forceinline__ void collocate_columns(const PointParams &P, const Task &T, const GeoClass &G,
                                  const Layout &lay, const WavePlan &wp, const R *tab,
                                  const R *cx, double *dst, int lane) {
  const int NP = wp.np;
  const R *tabz = tab + 2 * wp.ext * NP;
  const int nxy = lay.npts_local[0] * lay.npts_local[1];
  const int ngz = lay.npts_global[2], nz = lay.npts_local[2];
  int cw[3];
  axis_offsets(T, lay, cw);
  for (int c0 = 0; c0 < G.ncol; c0 += 64) {
    const Column C = load_column(P, G, lay, cw, c0 + lane);
    R D[LP + 1];
    if constexpr (LP >= 7) {
      R px[LP + 1];
      const R *rx = tab + clampi(C.i - T.lo[0], 0, wp.ext - 1) * NP;
      const R *ry = tab + (wp.ext + clampi(C.j - T.lo[1], 0, wp.ext - 1)) * NP;
#pragma unroll
      for (int l = 0; l <= LP; l++) px[l] = rx[l];
      column_poly_ry<LP>(cx, px, ry, D);
    } else {
      R px[LP + 1], py[LP + 1];
      load_pxy<LP>(T, wp, tab, C, px, py);
      column_poly<LP>(cx, px, py, D);
    }
    const int kmin = P.geo_kmin[G.grp_off + (c0 >> 6)];
    const unsigned span = (unsigned)(2 - 2 * C.kl);  // 0 for an idle lane
    int gz = wrap(cw[2] + kmin, ngz);
#pragma unroll 2
    for (int k = kmin; k <= 1 - kmin; k++) {
      const R *pz = tabz + (k - T.lo[2]) * NP;
      R v = 0;
#pragma unroll
      for (int lz = 0; lz <= LP; lz++) v = fmar(D[lz], pz[lz], v);
      if (gz < nz && (unsigned)(k - C.kl) < span) atomicAdd(&dst[gz * nxy + C.base], (double)v);
      gz = (gz + 1 == ngz) ? 0 : gz + 1;
    }
  }
}

// NCL groups of 64 columns at once (columns c0 + 64 u + lane, u < NCL): the
// planes loop, its pz reads and its scalar bookkeeping serve NCL points per
// lane and step.
template <int LP, int NCL, typename R>
__device__ __forceinline__ void tile_group(const PointParams &P, const Task &T, const GeoClass &G,
                                           const Layout &lay, const WavePlan &wp,
                                           const R *tab, const R *cx,
                                           const TileItem &I, const TileUnit &U, const int cw[3],
                                           int c0, int c_end, double *win, int lane) {
  const int NP = wp.np;
  const R *tabz = tab + 2 * wp.ext * NP;
  const int nx = lay.npts_local[0];
  const int nxby = nx * U.by;
  const int ngz = lay.npts_global[2];
  Column C[NCL];
  int lbase[NCL];
  unsigned span[NCL];
  R D[NCL][LP + 1];
  int klmin = 1;
#pragma unroll
  for (int u = 0; u < NCL; u++) {
    const int c = c0 + 64 * u + lane;
    C[u] = load_column(P, G, lay, cw, c < c_end ? c : G.ncol);
    lbase[u] = 0;
    if (C[u].kl <= 0) {
      const int ly = C[u].gy - U.gy0;
      if ((unsigned)ly < (unsigned)U.by)
        lbase[u] = C[u].base - U.gy0 * nx;
      else
        C[u].kl = 1;  // outside the tile: never for items built on the host
    }
    if constexpr (LP >= 7) {
      R px[LP + 1];
      const R *rx = tab + clampi(C[u].i - T.lo[0], 0, wp.ext - 1) * wp.np;
      const R *ry = tab + (wp.ext + clampi(C[u].j - T.lo[1], 0, wp.ext - 1)) * wp.np;
#pragma unroll
      for (int l = 0; l <= LP; l++) px[l] = rx[l];
      column_poly_ry<LP>(cx, px, ry, D[u]);
    } else {
      R px[LP + 1], py[LP + 1];
      load_pxy<LP>(T, wp, tab, C[u], px, py);
      column_poly<LP>(cx, px, py, D[u]);
    }
    span[u] = (unsigned)(2 - 2 * C[u].kl);  // 0 for an idle lane
    klmin = min(klmin, C[u].kl);
  }
  const int kmin = __builtin_amdgcn_readfirstlane(wave_min(klmin));
  const int ka = max(kmin, (int)I.k0), kb = min(1 - kmin, (int)I.k1);
  int gz = wrap(cw[2] + ka, ngz);
  const int tza = gz - U.gz0, nk = kb - ka + 1;
  if (tza >= 0 && tza + nk <= U.bz && gz + nk <= ngz) {
    // every plane inside the tile and no periodic wrap on the way: no test
    // or wrap per plane
    const R *pz = tabz + (ka - T.lo[2]) * NP;
    double *wz = win + tza * nxby;
#pragma unroll 2
    for (int k = ka; k <= kb; k++, pz += NP, wz += nxby) {
#pragma unroll
      for (int u = 0; u < NCL; u++) {
        R v = 0;
#pragma unroll
        for (int lz = 0; lz <= LP; lz++) v = fmar(D[u][lz], pz[lz], v);
        if ((unsigned)(k - C[u].kl) < span[u]) atomicAdd(&wz[lbase[u]], (double)v);
      }
    }
    return;
  }
#pragma unroll 2
  for (int k = ka; k <= kb; k++) {
    const R *pz = tabz + (k - T.lo[2]) * NP;
    const int tz = gz - U.gz0;
    if ((unsigned)tz < (unsigned)U.bz) {
#pragma unroll
      for (int u = 0; u < NCL; u++) {
        R v = 0;
#pragma unroll
        for (int lz = 0; lz <= LP; lz++) v = fmar(D[u][lz], pz[lz], v);
        if ((unsigned)(k - C[u].kl) < span[u]) atomicAdd(&win[tz * nxby + lbase[u]], (double)v);
      }
    }
    gz = (gz + 1 == ngz) ? 0 : gz + 1;
  }
}

// Collocate the rows j0..j1 and planes k0..k1 of a task into an LDS tile
// win[(z - gz0) * nx * by + (y - gy0) * nx + x] of unit U.
template <int LP, typename R>
__device__ __forceinline__ void tile_columns(const PointParams &P, const Task &T,
                                             const GeoClass &G, const Layout &lay,
                                             const WavePlan &wp, const R *tab,
                                             const R *cx, const TileItem &I,
                                             const TileUnit &U, double *win, int lane) {
  // two columns per lane fit the registers for lp <= 4 (FP64) or lp <= 6 (FP32)
  constexpr int L2 = sizeof(R) == 4 ? 6 : 4;
  int cw[3];
  axis_offsets(T, lay, cw);
  const int c_beg = P.geo_rows[G.row_off + (I.j0 - G.lo[1])];
  const int c_end = P.geo_rows[G.row_off + (I.j1 + 1 - G.lo[1])];
  for (int c0 = c_beg; c0 < c_end;) {
    if (LP <= L2 && c_end - c0 > 64) {
      tile_group<LP, (LP <= L2 ? 2 : 1), R>(P, T, G, lay, wp, tab, cx, I, U, cw, c0, c_end, win, lane);
      c0 += (LP <= L2 ? 128 : 64);
    } else {
      tile_group<LP, 1, R>(P, T, G, lay, wp, tab, cx, I, U, cw, c0, c_end, win, lane);
      c0 += 64;
    }
  }
}

// Integrate NCL groups of 64 columns at once into acc[(lz, ly, lx)]: the z
// moments of a column in the evaluation type R, their outer product with px,
// py accumulated in FP64.
template <int LP, int NCL, typename R>
__device__ __forceinline__ void integrate_group(const PointParams &P, const Task &T,
                                                const GeoClass &G, const Layout &lay,
                                                const WavePlan &wp, const R *tab,
                                                const R *grid, const int cw[3], int c0,
                                                double *acc, int lane) {
  const int NP = wp.np;
  const R *tabz = tab + 2 * wp.ext * NP;
  const int nxy = lay.npts_local[0] * lay.npts_local[1];
  const int ngz = lay.npts_global[2], nz = lay.npts_local[2];
  Column C[NCL];
  unsigned span[NCL];
  R M[NCL][LP + 1];
  int kmin = 1;
#pragma unroll
  for (int u = 0; u < NCL; u++) {
    C[u] = load_column(P, G, lay, cw, c0 + 64 * u + lane);
    span[u] = (unsigned)(2 - 2 * C[u].kl);
    kmin = min(kmin, P.geo_kmin[G.grp_off + (c0 >> 6) + u]);
#pragma unroll
    for (int l = 0; l <= LP; l++) M[u][l] = 0;
  }
  int gz = wrap(cw[2] + kmin, ngz);
  const int nk = 2 - 2 * kmin;
  // the grid through a global-memory pointer: generic (flat) loads would also
  // count on the LDS counter, so every wait for a table read would drain them
  typedef const __attribute__((address_space(1))) R GVal;
  GVal *gridg = (GVal *)grid;
  if (sizeof(R) == 4 && gz + nk <= nz && gz + nk <= ngz) {
    // FP32, every plane inside the local grid and no periodic wrap on the
    // way: no test or wrap per plane (FP64 measured slower with this copy)
    const R *pz = tabz + (kmin - T.lo[2]) * NP;
    GVal *gp = gridg + gz * nxy;
#pragma unroll 4  // several grid loads in flight per wave
    for (int k = kmin; k <= 1 - kmin; k++, pz += NP, gp += nxy) {
      R g[NCL];
#pragma unroll
      for (int u = 0; u < NCL; u++) {
        const bool in = (unsigned)(k - C[u].kl) < span[u];
        const R v = gp[in ? C[u].base : 0];
        g[u] = in ? v : R(0);
      }
#pragma unroll
      for (int u = 0; u < NCL; u++)
#pragma unroll
        for (int l = 0; l <= LP; l++) M[u][l] = fmar(g[u], pz[l], M[u][l]);
    }
  } else {
#pragma unroll 4  // several grid loads in flight per wave
    for (int k = kmin; k <= 1 - kmin; k++) {
      const R *pz = tabz + (k - T.lo[2]) * NP;
      R g[NCL];
#pragma unroll
      for (int u = 0; u < NCL; u++) {
        const bool in = gz < nz && (unsigned)(k - C[u].kl) < span[u];
        if constexpr (sizeof(R) == 4) {
          // FP32: unconditional loads (index 0 stands in for a point outside
          // the column) keep the unrolled iterations' loads in flight; FP64
          // measured faster with the loads under the branch
          const R v = gridg[in ? gz * nxy + C[u].base : 0];
          g[u] = in ? v : R(0);
        } else {
          g[u] = 0;
          if (in) g[u] = gridg[gz * nxy + C[u].base];
        }
      }
#pragma unroll
      for (int u = 0; u < NCL; u++)
#pragma unroll
        for (int l = 0; l <= LP; l++) M[u][l] = fmar(g[u], pz[l], M[u][l]);
      gz = (gz + 1 == ngz) ? 0 : gz + 1;
    }
  }
#pragma unroll
  for (int u = 0; u < NCL; u++) {
    R px[LP + 1], py[LP + 1];
    load_pxy<LP>(T, wp, tab, C[u], px, py);
    int e = 0;
#pragma unroll
    for (int lz = 0; lz <= LP; lz++)
#pragma unroll
      for (int ly = 0; ly <= LP - lz; ly++) {
        const double yz = (double)py[ly] * (double)M[u][lz];
#pragma unroll
        for (int lx = 0; lx <= LP - lz - ly; lx++, e++) acc[e] = fma((double)px[lx], yz, acc[e]);
      }
  }
}

// Integrate the task's points into lane-private accumulators acc[(lz, ly, lx)].
template <int LP, typename R>
__device__ __forceinline__ void integrate_columns(const PointParams &P, const Task &T,
                                                  const GeoClass &G, const Layout &lay,
                                                  const WavePlan &wp, const R *tab,
                                                  const R *grid, double *acc, int lane) {
  int cw[3];
  axis_offsets(T, lay, cw);
  // one column per lane: two (integrate_group<LP, 2>) cost occupancy (156
  // VGPRs) and measured slower on h2o64
  for (int c0 = 0; c0 < G.ncol; c0 += 64)
    integrate_group<LP, 1, R>(P, T, G, lay, wp, tab, grid, cw, c0, acc, lane);
}

template <int LP, typename R>
__device__ __forceinline__ void integrate_task(const PointParams &P, const Task &T, const GeoClass &G,
                               const Layout &lay, const WavePlan &wp, const R *tab,
                               const R *grid, double *out, int lane) {
  constexpr int NC = (LP + 1) * (LP + 2) * (LP + 3) / 6;
  double acc[NC];
#pragma unroll
  for (int e = 0; e < NC; e++) acc[e] = 0.0;
  integrate_columns<LP, R>(P, T, G, lay, wp, tab, grid, acc, lane);
  wave_sum_store<NC>(acc, out, lane);
}

// cxyz -> cab -> hab (Cartesian, atomics into the group's slice), forces and
// virial per task (grid_ref_integrate.c:132-171).
// NT threads per task: 64 (one wave, wave_sync) or a whole workgroup
// (__syncthreads; red: LDS for the per-wave force/virial partial sums).
template <int NT>
__device__ __forceinline__ void task_sync() {
  if constexpr (NT == 64)
    wave_sync();
  else
    __syncthreads();
}

// KIND: 0 hab only, 1 hab with forces / virial (no tau: the plain getters
// only, which keeps the register count low), 2 anything (tau getters).
template <int KIND, int NT>
__device__ __forceinline__ void task_finish(const PointParams &P, const Task &T, const Group &G, int la, int lb,
                            int la_min, int lb_min, const WaveLds &W, double *red,
                            long long task, int tid) {
  const int lane = tid & 63;
  const int n1 = ncoset(la), n2 = ncoset(lb);
  const int np = la + lb + 1, nA = la + 1, nB = lb + 1, lp = la + lb;
  double *cab = W.u, *alpha = W.u + n1 * n2, *X = alpha + 3 * nA * nB * np;
  build_alpha(T, la, lb, alpha, tid, NT);
  task_sync<NT>();
  // separable adjoint of cab -> cxyz: first along x for every (ax, bx),
  //   X[ax][bx][(lz, ly)] = sum_lx alpha_x[bx][ax][lx] cxyz[lx, ly, lz],
  // (lz, ly) packed with ly + lz <= lp (row lz holds np - lz entries; W.cxyz
  // holds the packed cxyz), then per cab element the y and z sums
  const int nq = np * (np + 1) / 2;
  for (int e = tid; e < nA * nB * nq; e += NT) {
    int q = e % nq, lz = 0;
    const int ab = e / nq, ax = ab / nB, bx = ab % nB;
    while (q >= np - lz) {
      q -= np - lz;
      lz++;
    }
    const int ly = q;
    const double *al = &alpha[(bx * nA + ax) * np];
    const double *c = W.cxyz + packed_index(0, ly, lz, lp);
    const int top = min(ax + bx, lp - ly - lz);
    double s = 0.0;
    for (int lx = 0; lx <= top; lx++) s = fma(al[lx], c[lx], s);
    X[e] = s;
  }
  task_sync<NT>();
  const int a1 = ncoset(la_min - 1), ma = ncoset(la) - a1;
  const int b1 = ncoset(lb_min - 1), mb = ncoset(lb) - b1;
  for (int e = tid; e < n1 * n2; e += NT) {
    const int ia = e % n1, jb = e / n1;
    double s = 0.0;
    if (ia >= a1 && ia < a1 + ma && jb >= b1 && jb < b1 + mb) {
      const Ang A = orb(ia), B = orb(jb);
      const double *ay = &alpha[((1 * nB + B.y) * nA + A.y) * np];
      const double *az = &alpha[((2 * nB + B.z) * nA + A.z) * np];
      const double *x = X + (A.x * nB + B.x) * nq;
      for (int lz = 0; lz <= A.z + B.z; lz++) {
        const double *xr = x + lz * np - lz * (lz - 1) / 2;  // row lz of the packed (lz, ly)
        double t = 0.0;
        for (int ly = 0; ly <= A.y + B.y; ly++) t = fma(ay[ly], xr[ly], t);
        s = fma(az[lz], t, s);
      }
      s *= T.pexp;
    }
    cab[e] = s;
  }
  task_sync<NT>();
  const CabGet cg{cab, n1};
  constexpr bool FULL = KIND > 0;
  const bool tau = KIND == 2 && P.compute_tau != 0;
  const double za = T.zeta, zb = T.zetb;
  const int a0 = ncoset(T.la_min - 1), na = ncoset(T.la_max) - a0;
  const int b0 = ncoset(T.lb_min - 1), nb = ncoset(T.lb_max) - b0;
  double f[6] = {0, 0, 0, 0, 0, 0};
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int e = tid; e < na * nb; e += NT) {
    const int ia = a0 + e % na, jb = b0 + e / na;
    const Ang A = orb(ia), B = orb(jb);
    const long long q = G.cart_off + (long long)(T.o2 + jb) * G.ncoa + T.o1 + ia;
    if (q < P.cart_len) atomicAdd(&P.hab_cart[q], get_hab(A, B, za, zb, cg, tau));
    if (FULL && P.do_forces && !tau) {
      const double pv = (q < P.cart_len) ? P.cart[q] : 0.0;
      pair_forces_plain(A, B, za, zb, T.rab, cg, pv, P.do_virial != 0, f, v);
    } else if (KIND == 2 && P.do_forces) {
      const double pv = (q < P.cart_len) ? P.cart[q] : 0.0;
      for (int i = 0; i < 3; i++) {
        f[i] += pv * get_fa(A, B, i, za, zb, cg, tau);
        f[3 + i] += pv * get_fb(A, B, i, za, zb, T.rab, cg, tau);
      }
      if (P.do_virial) {
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            v[i * 3 + j] += pv * get_va(A, B, i, j, za, zb, cg, tau);
            v[i * 3 + j] += pv * get_vb(A, B, i, j, za, zb, T.rab, cg, tau);
          }
      }
    }
  }
  if (FULL && P.do_forces) {
    double *out = P.task_out + task * kTaskOutStride;
    const double scale = (T.iatom == T.jatom) ? 1.0 : 2.0;
    // without virial only the 6 force sums (the 9 virial slots are zero)
    const int nsum = P.do_virial ? kTaskOutStride : 6;
    if (NT == 64 && !P.do_virial && lane >= 6 && lane < kTaskOutStride) out[lane] = 0.0;
    if (NT > 64 && !P.do_virial && tid >= 6 && tid < kTaskOutStride) out[tid] = 0.0;
    for (int c = 0; c < nsum; c++) {
      const double s = wave_total(c < 6 ? f[c] : v[c - 6]);
      if (NT == 64 && lane == 63) out[c] = scale * s;
      if (NT > 64 && lane == 63) red[(tid >> 6) * kTaskOutStride + c] = s;
    }
    if constexpr (NT > 64) {
      __syncthreads();
      if (tid < nsum) {
        double s = 0.0;
        for (int w = 0; w < NT / 64; w++) s += red[w * kTaskOutStride + tid];
        out[tid] = scale * s;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Lane-per-task cab -> cxyz for GRID_FUNC_AB (cab = the density sub-block),
// la, lb <= 3 with la + lb <= 4 (lane_cxyz_task): every loop over orbitals,
// binomials and (lx, ly, lz) is
// unrolled at compile time; accumulators live in registers.
// (x - a)^ax (x - b)^bx = sum_l alpha[l] (x - p)^l with
// alpha[k + m] += C(ax, k) C(bx, m) (p - a)^(ax - k) (p - b)^(bx - m),
// src/grid/ref/grid_ref_collint.h:840-861.
// ---------------------------------------------------------------------------
constexpr OrbTable kOrb = make_orb_table();

constexpr int binom(int n, int k) {
  int r = 1;
  for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
  return r;
}

constexpr int ncoset_c(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6; }

constexpr int pidx_c(int lx, int ly, int lz, int lp) {
  int off = 0;
  for (int z = 0; z < lz; z++) off += (lp - z + 1) * (lp - z + 2) / 2;
  for (int y = 0; y < ly; y++) off += lp - lz - y + 1;
  return off + lx;
}

// alpha of one axis for exponents (A, B): coefficients of (x - p)^l, l <= A + B
template <int A, int B>
__device__ __forceinline__ void axis_alpha(const double *pa, const double *pb, double *al) {
#pragma unroll
  for (int l = 0; l <= A + B; l++) al[l] = 0.0;
#pragma unroll
  for (int k = 0; k <= A; k++)
#pragma unroll
    for (int m = 0; m <= B; m++)
      al[k + m] = fma((double)(binom(A, k) * binom(B, m)), pa[A - k] * pb[B - m], al[k + m]);
}

template <int LA, int LB, int IA, int IB>
__device__ __forceinline__ void lane_pair(const double (&pa)[3][LA + 1], const double (&pb)[3][LB + 1],
                                          double c, double *acc) {
  constexpr int LP = LA + LB;
  constexpr int AX = kOrb.v[IA][0], AY = kOrb.v[IA][1], AZ = kOrb.v[IA][2];
  constexpr int BX = kOrb.v[IB][0], BY = kOrb.v[IB][1], BZ = kOrb.v[IB][2];
  double ax[AX + BX + 1], ay[AY + BY + 1], az[AZ + BZ + 1];
  axis_alpha<AX, BX>(pa[0], pb[0], ax);
  axis_alpha<AY, BY>(pa[1], pb[1], ay);
  axis_alpha<AZ, BZ>(pa[2], pb[2], az);
#pragma unroll
  for (int lz = 0; lz <= AZ + BZ; lz++)
#pragma unroll
    for (int ly = 0; ly <= AY + BY; ly++) {
      const double cyz = c * (ay[ly] * az[lz]);
#pragma unroll
      for (int lx = 0; lx <= AX + BX; lx++)
        acc[pidx_c(lx, ly, lz, LP)] = fma(cyz, ax[lx], acc[pidx_c(lx, ly, lz, LP)]);
    }
}

template <int LA, int LB, int IA, int IB>
struct PairLoop {
  static __device__ __forceinline__ void run(const Task &T, const Group &G, const double *cart,
                                             long long cart_len, double pref,
                                             const double (&pa)[3][LA + 1],
                                             const double (&pb)[3][LB + 1], double *acc) {
    constexpr int la_i = kOrb.v[IA][0] + kOrb.v[IA][1] + kOrb.v[IA][2];
    constexpr int lb_i = kOrb.v[IB][0] + kOrb.v[IB][1] + kOrb.v[IB][2];
    if (la_i >= T.la_min && lb_i >= T.lb_min) {
      const long long q = G.cart_off + (long long)(T.o2 + IB) * G.ncoa + T.o1 + IA;
      const double p = (q < cart_len) ? cart[q] : 0.0;
      if (p != 0.0) lane_pair<LA, LB, IA, IB>(pa, pb, p * pref, acc);
    }
    if constexpr (IA + 1 < ncoset_c(LA))
      PairLoop<LA, LB, IA + 1, IB>::run(T, G, cart, cart_len, pref, pa, pb, acc);
    else if constexpr (IB + 1 < ncoset_c(LB))
      PairLoop<LA, LB, 0, IB + 1>::run(T, G, cart, cart_len, pref, pa, pb, acc);
  }
};

template <int LA, int LB>
__device__ __forceinline__ void lane_cxyz(const PointParams &P, const Task &T, const Group &G,
                                          double *out) {
  constexpr int LP = LA + LB;
  constexpr int NC = (LP + 1) * (LP + 2) * (LP + 3) / 6;
  double acc[NC];
#pragma unroll
  for (int e = 0; e < NC; e++) acc[e] = 0.0;
  double pa[3][LA + 1], pb[3][LB + 1];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double da = T.rp[d] - T.ra[d], db = T.rp[d] - T.rb[d];
    pa[d][0] = 1.0;
    pb[d][0] = 1.0;
#pragma unroll
    for (int k = 1; k <= LA; k++) pa[d][k] = pa[d][k - 1] * da;
#pragma unroll
    for (int k = 1; k <= LB; k++) pb[d][k] = pb[d][k - 1] * db;
  }
  PairLoop<LA, LB, 0, 0>::run(T, G, P.cart, P.cart_len, T.rscale * T.pexp, pa, pb, acc);
#pragma unroll
  for (int e = 0; e < NC; e++) out[e] = acc[e];
}

// ---------------------------------------------------------------------------
// Lane-per-task cxyz -> cab -> hab (and forces) for small l: the adjoint of
// lane_cxyz.  LA, LB: the shifted l of the call (forces raise la by one and
// lower both l_min by one, src/grid/ref/grid_ref_integrate.c); cab[A][B] for
// every orbital pair below LA, LB is a compile-time unrolled sum
//   cab(A, B) = pexp sum_{lx,ly,lz} ax[lx] ay[ly] az[lz] cxyz[lx, ly, lz]
// with the per-axis expansions of lane_cxyz; hab and the force getters of
// grid_process_vab.h:29-205 then read the needed entries.  Each lane adds
// its task's hab sub-block with global atomics (images of one atom pair
// share it) and writes its own forces (no cross-lane reduction).
// ---------------------------------------------------------------------------
constexpr int cidx_c(int x, int y, int z) {
  return x + y + z == 0 ? 0
                        : ncoset_c(x + y + z - 1) + ((y + z) * (y + z + 1)) / 2 + z;
}

template <int LA, int LB, int IA, int IB>
__device__ __forceinline__ double lane_cab(const double (&pa)[3][LA + 1],
                                           const double (&pb)[3][LB + 1], const double *cx) {
  constexpr int LP = LA + LB;
  constexpr int AX = kOrb.v[IA][0], AY = kOrb.v[IA][1], AZ = kOrb.v[IA][2];
  constexpr int BX = kOrb.v[IB][0], BY = kOrb.v[IB][1], BZ = kOrb.v[IB][2];
  double ax[AX + BX + 1], ay[AY + BY + 1], az[AZ + BZ + 1];
  axis_alpha<AX, BX>(pa[0], pb[0], ax);
  axis_alpha<AY, BY>(pa[1], pb[1], ay);
  axis_alpha<AZ, BZ>(pa[2], pb[2], az);
  double s = 0.0;
#pragma unroll
  for (int lz = 0; lz <= AZ + BZ; lz++) {
    double u = 0.0;
#pragma unroll
    for (int ly = 0; ly <= AY + BY; ly++) {
      double t = 0.0;
#pragma unroll
      for (int lx = 0; lx <= AX + BX; lx++) t = fma(ax[lx], cx[pidx_c(lx, ly, lz, LP)], t);
      u = fma(ay[ly], t, u);
    }
    s = fma(az[lz], u, s);
  }
  return s;
}

// The fields of a task and its group that the lane finish reads, copied to
// registers once (the hab atomics could otherwise alias them, and every
// scattered reload costs the address unit a cache line per lane).
struct LaneTask {
  long long base;   // cart offset of the task's (o1, o2) corner
  int ncoa, la_min, lb_min;
  double pexp, zeta, zetb, rab[3];
};

// hab and forces of orbital pair (IA, IB) (l up to the task's la_max,
// lb_max), each cab entry evaluated where it is used.
template <int LA, int LB, int IA, int IB, bool FORCES>
struct GetLoop {
  static __device__ __forceinline__ void run(const PointParams &P, const LaneTask &L,
                                             const double (&pa)[3][LA + 1],
                                             const double (&pb)[3][LB + 1], const double *cx,
                                             double *f) {
    constexpr int AX = kOrb.v[IA][0], AY = kOrb.v[IA][1], AZ = kOrb.v[IA][2];
    constexpr int BX = kOrb.v[IB][0], BY = kOrb.v[IB][1], BZ = kOrb.v[IB][2];
    if (AX + AY + AZ >= L.la_min && BX + BY + BZ >= L.lb_min) {
      const long long q = L.base + (long long)IB * L.ncoa + IA;
      if (q < P.cart_len) {
        const double c0 = L.pexp * lane_cab<LA, LB, IA, IB>(pa, pb, cx);
        __hip_atomic_fetch_add(&P.hab_cart[q], c0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if constexpr (FORCES) {
          const double pv = __builtin_nontemporal_load(&P.cart[q]);
          const double u0 = L.pexp * lane_cab<LA, LB, cidx_c(AX + 1, AY, AZ), IB>(pa, pb, cx);
          const double u1 = L.pexp * lane_cab<LA, LB, cidx_c(AX, AY + 1, AZ), IB>(pa, pb, cx);
          const double u2 = L.pexp * lane_cab<LA, LB, cidx_c(AX, AY, AZ + 1), IB>(pa, pb, cx);
          double d0 = 0.0, d1 = 0.0, d2 = 0.0, e0 = 0.0, e1 = 0.0, e2 = 0.0;
          if constexpr (AX > 0) d0 = AX * L.pexp * lane_cab<LA, LB, cidx_c(AX - 1, AY, AZ), IB>(pa, pb, cx);
          if constexpr (AY > 0) d1 = AY * L.pexp * lane_cab<LA, LB, cidx_c(AX, AY - 1, AZ), IB>(pa, pb, cx);
          if constexpr (AZ > 0) d2 = AZ * L.pexp * lane_cab<LA, LB, cidx_c(AX, AY, AZ - 1), IB>(pa, pb, cx);
          if constexpr (BX > 0) e0 = BX * L.pexp * lane_cab<LA, LB, IA, cidx_c(BX - 1, BY, BZ)>(pa, pb, cx);
          if constexpr (BY > 0) e1 = BY * L.pexp * lane_cab<LA, LB, IA, cidx_c(BX, BY - 1, BZ)>(pa, pb, cx);
          if constexpr (BZ > 0) e2 = BZ * L.pexp * lane_cab<LA, LB, IA, cidx_c(BX, BY, BZ - 1)>(pa, pb, cx);
          f[0] += pv * (2.0 * L.zeta * u0 - d0);
          f[1] += pv * (2.0 * L.zeta * u1 - d1);
          f[2] += pv * (2.0 * L.zeta * u2 - d2);
          f[3] += pv * (2.0 * L.zetb * (u0 - L.rab[0] * c0) - e0);
          f[4] += pv * (2.0 * L.zetb * (u1 - L.rab[1] * c0) - e1);
          f[5] += pv * (2.0 * L.zetb * (u2 - L.rab[2] * c0) - e2);
        }
      }
    }
    constexpr int LA0 = FORCES ? LA - 1 : LA;
    if constexpr (IA + 1 < ncoset_c(LA0))
      GetLoop<LA, LB, IA + 1, IB, FORCES>::run(P, L, pa, pb, cx, f);
    else if constexpr (IB + 1 < ncoset_c(LB))
      GetLoop<LA, LB, 0, IB + 1, FORCES>::run(P, L, pa, pb, cx, f);
  }
};

template <int LA0, int LB0, bool FORCES>
__device__ __forceinline__ void lane_finish(const PointParams &P, const Task &T, const Group &G,
                                            long long task) {
  constexpr int LA = LA0 + (FORCES ? 1 : 0), LB = LB0;
  constexpr int LP = LA + LB;
  constexpr int NC = (LP + 1) * (LP + 2) * (LP + 3) / 6;
  LaneTask L;
  L.ncoa = G.ncoa;
  L.base = G.cart_off + (long long)T.o2 * L.ncoa + T.o1;
  L.la_min = T.la_min;
  L.lb_min = T.lb_min;
  L.pexp = T.pexp;
  L.zeta = T.zeta;
  L.zetb = T.zetb;
  double pa[3][LA + 1], pb[3][LB + 1];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    L.rab[d] = T.rab[d];
    const double da = T.rp[d] - T.ra[d], db = T.rp[d] - T.rb[d];
    pa[d][0] = 1.0;
    pb[d][0] = 1.0;
#pragma unroll
    for (int k = 1; k <= LA; k++) pa[d][k] = pa[d][k - 1] * da;
#pragma unroll
    for (int k = 1; k <= LB; k++) pb[d][k] = pb[d][k - 1] * db;
  }
  const double scale = (T.iatom == T.jatom) ? 1.0 : 2.0;
  double cx[NC];
  const double *in = P.cxyz_all + task * P.cxyz_stride;
#pragma unroll
  for (int e = 0; e < NC; e++) cx[e] = in[e];
  double f[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  GetLoop<LA, LB, 0, 0, FORCES>::run(P, L, pa, pb, cx, f);
  if constexpr (FORCES) {
    double *out = P.task_out + task * kTaskOutStride;
#pragma unroll
    for (int c = 0; c < 6; c++) out[c] = scale * f[c];
#pragma unroll
    for (int c = 6; c < kTaskOutStride; c++) out[c] = 0.0;
  }
}

// A task per lane (la_max + lb_max <= 3, no virial, no tau: the host
// routes only those here).
template <bool FORCES>
__global__ __launch_bounds__(256) void finish_lane_kernel(const Task *__restrict__ tasks,
                                                          const Group *__restrict__ groups,
                                                          PointParams P) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P.ntasks) return;
  const long long task = launch_task(P, t);
  const Task &T = tasks[task];
  const Group &G = groups[T.group];
  switch (T.la_max * 4 + T.lb_max) {
    case 0: lane_finish<0, 0, FORCES>(P, T, G, task); break;
    case 1: lane_finish<0, 1, FORCES>(P, T, G, task); break;
    case 2: lane_finish<0, 2, FORCES>(P, T, G, task); break;
    case 4: lane_finish<1, 0, FORCES>(P, T, G, task); break;
    case 5: lane_finish<1, 1, FORCES>(P, T, G, task); break;
    case 8: lane_finish<2, 0, FORCES>(P, T, G, task); break;
    case 3: lane_finish<0, 3, FORCES>(P, T, G, task); break;
    case 6: lane_finish<1, 2, FORCES>(P, T, G, task); break;
    case 9: lane_finish<2, 1, FORCES>(P, T, G, task); break;
    case 12: lane_finish<3, 0, FORCES>(P, T, G, task); break;
    default: break;
  }
}

// ---------------------------------------------------------------------------
// Integrate for lp 7..11 on the matrix cores.  Per task the coefficients
//   cxyz[lx][(ly, lz)] = sum over columns c of px[c][lx] * (py[c][ly] M[c][lz])
// (M = the column's z moments) are a (lp+1) x NQ x ncol product; the
// 120..364 accumulators do not fit in lane registers, but on the matrix
// cores they are 4 x NB doubles per lane.  Columns are walked 64 at a time
// (one per lane, as in integrate_group), px, py, M staged in LDS, then 16
// steps of v_mfma_f64_16x16x4f64 over 4 columns each.  Operand layout on
// gfx950 (checked on the device): A[i][k] and B[k][j] in lane i + 16 k
// (j + 16 k), D[i][j] in lane j + 16 (i % 4), register i / 4.
// ---------------------------------------------------------------------------
typedef double double4_t __attribute__((ext_vector_type(4)));
constexpr int kMfmaStage = 40;  // LDS doubles per column: px[16] (zero padded), py[12], M[12]

template <int LP>
__device__ __forceinline__ void integrate_mfma_task(const PointParams &P, const Task &T,
                                                    const GeoClass &G, const Layout &lay,
                                                    const WavePlan &wp, const double *tab,
                                                    double *stage, const double *grid,
                                                    double *out, int lane) {
  constexpr int NQ = (LP + 1) * (LP + 2) / 2;  // (ly, lz) with ly + lz <= LP, lz-major
  constexpr int NB = (NQ + 15) / 16;
  const int NP = wp.np;
  const double *tabz = tab + 2 * wp.ext * NP;
  const int nxy = lay.npts_local[0] * lay.npts_local[1];
  const int ngz = lay.npts_global[2], nz = lay.npts_local[2];
  // (ly, lz) of this lane's B column in each block
  int qy[NB], qz[NB];
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
  double4_t acc[NB];
#pragma unroll
  for (int nb = 0; nb < NB; nb++) acc[nb] = double4_t{0.0, 0.0, 0.0, 0.0};
  int cw[3];
  axis_offsets(T, lay, cw);
  double *sc = stage + lane * kMfmaStage;
  for (int c0 = 0; c0 < G.ncol; c0 += 64) {
    const Column C = load_column(P, G, lay, cw, c0 + lane);
    const unsigned span = (unsigned)(2 - 2 * C.kl);
    const int kmin = P.geo_kmin[G.grp_off + (c0 >> 6)];
    double M[LP + 1];
#pragma unroll
    for (int l = 0; l <= LP; l++) M[l] = 0.0;
    int gz = wrap(cw[2] + kmin, ngz);
#pragma unroll 2
    for (int k = kmin; k <= 1 - kmin; k++) {
      const double *pz = tabz + (k - T.lo[2]) * NP;
      double g = 0.0;
      if (gz < nz && (unsigned)(k - C.kl) < span) g = grid[gz * nxy + C.base];
#pragma unroll
      for (int l = 0; l <= LP; l++) M[l] = fma(g, pz[l], M[l]);
      gz = (gz + 1 == ngz) ? 0 : gz + 1;
    }
    double px[LP + 1], py[LP + 1];
    load_pxy<LP>(T, wp, tab, C, px, py);
#pragma unroll
    for (int l = 0; l < 16; l++) sc[l] = l <= LP ? px[l < LP ? l : LP] : 0.0;
#pragma unroll
    for (int l = 0; l <= LP; l++) {
      sc[16 + l] = py[l];
      sc[28 + l] = M[l];
    }
    wave_sync();
#pragma unroll 4
    for (int kk = 0; kk < 16; kk++) {
      const double *sk = stage + (4 * kk + (lane >> 4)) * kMfmaStage;
      const double a = sk[lane & 15];
#pragma unroll
      for (int nb = 0; nb < NB; nb++) {
        const double b = qz[nb] <= LP ? sk[16 + qy[nb]] * sk[28 + qz[nb]] : 0.0;
        acc[nb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[nb], 0, 0, 0);
      }
    }
    wave_sync();
  }
#pragma unroll
  for (int nb = 0; nb < NB; nb++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int lx = 4 * r + (lane >> 4);
      if (qz[nb] <= LP && lx <= LP - qy[nb] - qz[nb])
        out[packed_index(lx, qy[nb], qz[nb], LP)] = acc[nb][r];
    }
}

struct TaskL {
  int la, lb, la_min, lb_min, lp;
};

__device__ __forceinline__ TaskL task_l(const PointParams &P, const Task &T) {
  TaskL l;
  l.la = T.la_max + P.ld[0];
  l.lb = T.lb_max + P.ld[2];
  l.la_min = max(T.la_min + P.ld[1], 0);
  l.lb_min = max(T.lb_min + P.ld[3], 0);
  l.lp = l.la + l.lb;
  return l;
}

}  // namespace

// ---------------------------------------------------------------------------
// Kernels.  Launch ranges hold orthorhombic tasks of level wp.level that
// belong to a geometry class, with lp <= LPB (point kernels).
// ---------------------------------------------------------------------------

// cab -> cxyz of one task per lane (GRID_FUNC_AB, la_max, lb_max <= 2).
__global__ __launch_bounds__(256) void cxyz_lane_kernel(const Task *__restrict__ tasks,
                                                        const Group *__restrict__ groups,
                                                        PointParams P) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P.ntasks) return;
  const long long task = launch_task(P, t);
  const Task &T = tasks[task];
  if (!lane_cxyz_task(T.la_max, T.lb_max)) return;  // cxyz_wave_kernel takes these
  const Group &G = groups[T.group];
  double *out = P.cxyz_all + task * P.cxyz_stride;
  switch (T.la_max * 4 + T.lb_max) {
    case 0: lane_cxyz<0, 0>(P, T, G, out); break;
    case 1: lane_cxyz<0, 1>(P, T, G, out); break;
    case 2: lane_cxyz<0, 2>(P, T, G, out); break;
    case 3: lane_cxyz<0, 3>(P, T, G, out); break;
    case 4: lane_cxyz<1, 0>(P, T, G, out); break;
    case 5: lane_cxyz<1, 1>(P, T, G, out); break;
    case 6: lane_cxyz<1, 2>(P, T, G, out); break;
    case 7: lane_cxyz<1, 3>(P, T, G, out); break;
    case 8: lane_cxyz<2, 0>(P, T, G, out); break;
    case 9: lane_cxyz<2, 1>(P, T, G, out); break;
    case 10: lane_cxyz<2, 2>(P, T, G, out); break;
    case 12: lane_cxyz<3, 0>(P, T, G, out); break;
    case 13: lane_cxyz<3, 1>(P, T, G, out); break;
    default: break;
  }
}

// cab -> cxyz of one task per wave, stored packed at cxyz_all[task * stride];
// with lane_cxyz set, tasks the lane kernel takes are skipped.
__global__ __launch_bounds__(256) void cxyz_wave_kernel(const Task *__restrict__ tasks,
                                                        const Group *__restrict__ groups,
                                                        PointParams P) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long t = (long long)blockIdx.x * (blockDim.x >> 6) + wave;
  if (t >= P.ntasks) return;
  const long long task = launch_task(P, t);
  const Task &T = tasks[task];
  if (P.lane_cxyz && lane_cxyz_task(T.la_max, T.lb_max)) return;
  const TaskL l = task_l(P, T);
  const WaveLds W = wave_lds(lds, P.wp, wave);
  const int NP = P.wp.np;
  wave_cxyz(P, T, groups[T.group], l.la, l.lb, l.la_min, l.lb_min, NP, W, lane);
  double *out = P.cxyz_all + task * P.cxyz_stride;
  for (int e = lane; e < NP * NP * NP; e += 64) {
    const int lz = e / (NP * NP), ly = (e / NP) % NP, lx = e % NP;
    if (lx + ly + lz <= l.lp) out[packed_index(lx, ly, lz, l.lp)] = W.cxyz[e];
  }
}

// Packed FP64 cxyz of a task -> the wave's LDS, in the evaluation type R.
template <typename R>
__device__ __forceinline__ const R *stage_cxyz(const PointParams &P, long long task, int lp,
                                               const WaveLds &W, int lane) {
  const int nc = (lp + 1) * (lp + 2) * (lp + 3) / 6;
  const double *src = P.cxyz_all + task * P.cxyz_stride;
  R *cx = reinterpret_cast<R *>(W.cxyz);
  for (int e = lane; e < nc; e += 64) cx[e] = (R)src[e];
  return cx;
}

// Collocation straight into HBM, one task per wave.
template <int LPB, typename R>
__global__ __launch_bounds__(256) void collocate_wave_kernel(const Task *__restrict__ tasks,
                                                             const Level *__restrict__ levels,
                                                             PointParams P) {
  extern __shared__ double lds[];
  const WavePlan &wp = P.wp;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long t = (long long)blockIdx.x * (blockDim.x >> 6) + wave;
  if (t >= P.ntasks) return;
  const long long task = launch_task(P, t);
  const Task &T = tasks[task];
  const GeoClass &G = P.geo[T.geo];
  const Level &L = levels[T.level];
  const Layout &lay = L.lay;
  const WaveLds W = wave_lds(lds, wp, wave);
  const int lp = T.la_max + T.lb_max + P.ld[0] + P.ld[2];
  const R *cx = stage_cxyz<R>(P, task, lp, W, lane);
  R *tab = reinterpret_cast<R *>(W.u);
  wave_tables(T, lay, lp, wp, tab, lane);  // ends with wave_sync
  double *dst = L.grid;
  switch (lp) {
    case 0: collocate_columns<0>(P, T, G, lay, wp, tab, cx, dst, lane); break;
    case 1: collocate_columns<1>(P, T, G, lay, wp, tab, cx, dst, lane); break;
    case 2: collocate_columns<2>(P, T, G, lay, wp, tab, cx, dst, lane); break;
    case 3: collocate_columns<3>(P, T, G, lay, wp, tab, cx, dst, lane); break;
    case 4: collocate_columns<4>(P, T, G, lay, wp, tab, cx, dst, lane); break;
    case 5: if (LPB >= 5) collocate_columns<(LPB >= 5 ? 5 : 0)>(P, T, G, lay, wp, tab, cx, dst, lane); break;
    case 6: if (LPB >= 6) collocate_columns<(LPB >= 6 ? 6 : 0)>(P, T, G, lay, wp, tab, cx, dst, lane); break;
    case 7: if (LPB >= 7) collocate_columns<(LPB >= 7 ? 7 : 0)>(P, T, G, lay, wp, tab, cx, dst, lane); break;
    case 8: if (LPB >= 8) collocate_columns<(LPB >= 8 ? 8 : 0)>(P, T, G, lay, wp, tab, cx, dst, lane); break;
    default: break;  // the host routes larger lp to the per-task kernels
  }
}

// Collocation through LDS tiles: workgroup = unit P.unit_base + blockIdx.x,
// the items U.ibeg + U.first + m * U.stride < U.iend of one tile.  The tile
// accumulates in FP64 whatever the evaluation type.
template <int LPB, typename R>
__device__ __forceinline__ void collocate_tile_unit(const Task *__restrict__ tasks,
                                                    const Level *__restrict__ levels,
                                                    const TileUnit *__restrict__ units,
                                                    const TileItem *__restrict__ items,
                                                    const PointParams &P, int u) {
  extern __shared__ double lds[];
  const WavePlan &wp = P.wp;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  const TileUnit U = units[P.unit_base + u];
  const Level &L = levels[U.level];
  const Layout &lay = L.lay;
  const int window = lay.npts_local[0] * U.by * U.bz;  // <= wp.window
  for (int e = threadIdx.x; e < window; e += blockDim.x) lds[e] = 0.0;
  __syncthreads();
  const WaveLds W = wave_lds(lds, wp, wave);
  R *tab = reinterpret_cast<R *>(W.u);
  for (long long it = U.ibeg + U.first + (long long)wave * U.stride; it < U.iend;
       it += (long long)nw * U.stride) {
    const TileItem I = items[it];
    const Task &T = tasks[I.task];
    const GeoClass &G = P.geo[T.geo];
    const int lp = T.la_max + T.lb_max + P.ld[0] + P.ld[2];
    const R *cx = stage_cxyz<R>(P, I.task, lp, W, lane);
    wave_tables(T, lay, lp, wp, tab, lane);  // ends with wave_sync
    switch (lp) {
      case 0: tile_columns<0>(P, T, G, lay, wp, tab, cx, I, U, lds, lane); break;
      case 1: tile_columns<1>(P, T, G, lay, wp, tab, cx, I, U, lds, lane); break;
      case 2: tile_columns<2>(P, T, G, lay, wp, tab, cx, I, U, lds, lane); break;
      case 3: tile_columns<3>(P, T, G, lay, wp, tab, cx, I, U, lds, lane); break;
      case 4: tile_columns<4>(P, T, G, lay, wp, tab, cx, I, U, lds, lane); break;
      case 5: if (LPB >= 5) tile_columns<(LPB >= 5 ? 5 : 0)>(P, T, G, lay, wp, tab, cx, I, U, lds, lane); break;
      case 6: if (LPB >= 6) tile_columns<(LPB >= 6 ? 6 : 0)>(P, T, G, lay, wp, tab, cx, I, U, lds, lane); break;
      case 7: if (LPB >= 7) tile_columns<(LPB >= 7 ? 7 : 0)>(P, T, G, lay, wp, tab, cx, I, U, lds, lane); break;
      case 8: if (LPB >= 8) tile_columns<(LPB >= 8 ? 8 : 0)>(P, T, G, lay, wp, tab, cx, I, U, lds, lane); break;
      default: break;  // the host routes larger lp to the per-task kernels
    }
    wave_sync();
  }
  __syncthreads();
  // one tile row (nx points) per wave and pass: coalesced adds, no division
  // per point
  const int nx = lay.npts_local[0], ny = lay.npts_local[1], nz = lay.npts_local[2];
  for (int r = wave; r < U.by * U.bz; r += nw) {
    const int y = U.gy0 + r % U.by, z = U.gz0 + r / U.by;
    if (y >= ny || z >= nz) continue;
    double *dst = L.grid + ((long long)z * ny + y) * nx;
    const double *src = lds + r * nx;
    for (int x = lane; x < nx; x += 64) {
      const double v = src[x];
      if (v != 0.0) atomicAdd(&dst[x], v);
    }
  }
}

// One unit per workgroup; QUEUE (persistent launch of a batch): units from
// the cursor until all are taken.
template <int LPB, typename R, bool QUEUE>
__global__ __launch_bounds__(LPB <= 4 ? 768 : 512) void collocate_tile_kernel(
    const Task *__restrict__ tasks, const Level *__restrict__ levels,
    const TileUnit *__restrict__ units, const TileItem *__restrict__ items, PointParams P) {
  if constexpr (!QUEUE) {
    collocate_tile_unit<LPB, R>(tasks, levels, units, items, P, blockIdx.x);
  } else {
    __shared__ int next_unit;
    if (threadIdx.x == 0) next_unit = atomicAdd(P.queue, 1);
    __syncthreads();
    int u = next_unit;
    __syncthreads();
    while (u < P.nunits) {
      collocate_tile_unit<LPB, R>(tasks, levels, units, items, P, u);
      __syncthreads();  // the window is cleared for the next unit
      if (threadIdx.x == 0) next_unit = atomicAdd(P.queue, 1);
      __syncthreads();
      u = next_unit;
      __syncthreads();
    }
  }
}

// R = float reads the FP32 copy of the potential (Level::gridf).
template <int LPB, typename R>
__device__ __forceinline__ void integrate_wave_one(const Task *__restrict__ tasks,
                                                   const Level *__restrict__ levels,
                                                   const PointParams &P, R *tab, long long t,
                                                   int lane) {
  const WavePlan &wp = P.wp;
  const long long task = launch_task(P, t);
  const Task &T = tasks[task];
  const GeoClass &G = P.geo[T.geo];
  const Level &L = levels[T.level];
  const Layout &lay = L.lay;
  const int lp = T.la_max + T.lb_max + P.ld[0] + P.ld[2];
  wave_tables(T, lay, lp, wp, tab, lane);
  double *out = P.cxyz_all + task * P.cxyz_stride;
  const R *grid;
  if constexpr (sizeof(R) == 4)
    grid = L.gridf;
  else
    grid = L.grid;
  switch (lp) {
    case 0: integrate_task<0>(P, T, G, lay, wp, tab, grid, out, lane); break;
    case 1: integrate_task<1>(P, T, G, lay, wp, tab, grid, out, lane); break;
    case 2: integrate_task<2>(P, T, G, lay, wp, tab, grid, out, lane); break;
    case 3: integrate_task<3>(P, T, G, lay, wp, tab, grid, out, lane); break;
    case 4: integrate_task<4>(P, T, G, lay, wp, tab, grid, out, lane); break;
    case 5: if (LPB >= 5) integrate_task<(LPB >= 5 ? 5 : 0)>(P, T, G, lay, wp, tab, grid, out, lane); break;
    case 6: if (LPB >= 6) integrate_task<(LPB >= 6 ? 6 : 0)>(P, T, G, lay, wp, tab, grid, out, lane); break;
    default: break;  // the host routes larger lp to the per-task kernels
  }
}

// One task per wave; QUEUE (persistent launch of a batch): groups of qchunk
// tasks from the cursor until all are taken.  Two instantiations so that the
// one-task launch keeps the straight-line code of a single task.
template <int LPB, typename R, bool QUEUE>
__global__ __launch_bounds__(256, (LPB <= 4 ? 3 : 2)) void integrate_wave_kernel(
    const Task *__restrict__ tasks, const Level *__restrict__ levels, PointParams P) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const WaveLds W = wave_lds(lds, P.wp, wave);
  R *tab = reinterpret_cast<R *>(W.u);
  if constexpr (!QUEUE) {
    const long long t = (long long)blockIdx.x * (blockDim.x >> 6) + wave;
    if (t >= P.ntasks) return;
    integrate_wave_one<LPB, R>(tasks, levels, P, tab, t, lane);
  } else {
    int base = 0;
    if (lane == 0) base = atomicAdd(P.queue, P.qchunk);
    long long t0 = __builtin_amdgcn_readfirstlane(base);
    while (t0 < P.ntasks) {
      for (long long t = t0; t < t0 + P.qchunk && t < P.ntasks; t++) {
        integrate_wave_one<LPB, R>(tasks, levels, P, tab, t, lane);
        wave_sync();  // the tables are rewritten for the next task
      }
      base = 0;
      if (lane == 0) base = atomicAdd(P.queue, P.qchunk);
      t0 = __builtin_amdgcn_readfirstlane(base);
    }
  }
}

// Integration of orthorhombic tasks with lp 7..11, one task per wave, the
// column contraction on the matrix cores (integrate_mfma_task).  FP64 in both
// precision modes.  LDS per wave: tables at o_u, staging at o_cxyz.
__global__ __launch_bounds__(256) void integrate_mfma_kernel(const Task *__restrict__ tasks,
                                                             const Level *__restrict__ levels,
                                                             PointParams P) {
  extern __shared__ double lds[];
  const WavePlan &wp = P.wp;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long t = (long long)blockIdx.x * (blockDim.x >> 6) + wave;
  if (t >= P.ntasks) return;
  const long long task = launch_task(P, t);
  const Task &T = tasks[task];
  const GeoClass &G = P.geo[T.geo];
  const Level &L = levels[T.level];
  const Layout &lay = L.lay;
  const WaveLds W = wave_lds(lds, wp, wave);
  const int lp = T.la_max + T.lb_max + P.ld[0] + P.ld[2];
  wave_tables(T, lay, lp, wp, W.u, lane);
  double *out = P.cxyz_all + task * P.cxyz_stride;
  switch (lp) {
    case 7: integrate_mfma_task<7>(P, T, G, lay, wp, W.u, W.cxyz, L.grid, out, lane); break;
    case 8: integrate_mfma_task<8>(P, T, G, lay, wp, W.u, W.cxyz, L.grid, out, lane); break;
    case 9: integrate_mfma_task<9>(P, T, G, lay, wp, W.u, W.cxyz, L.grid, out, lane); break;
    case 10: integrate_mfma_task<10>(P, T, G, lay, wp, W.u, W.cxyz, L.grid, out, lane); break;
    case 11: integrate_mfma_task<11>(P, T, G, lay, wp, W.u, W.cxyz, L.grid, out, lane); break;
    default: break;  // the host routes other lp elsewhere
  }
}

// Packed cxyz of one task per wave -> hab, forces, virial.
template <int KIND>
__global__ __launch_bounds__(256, (KIND == 2 ? 2 : (KIND == 1 ? 3 : 4))) void finish_wave_kernel(
    const Task *__restrict__ tasks, const Group *__restrict__ groups, PointParams P) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long t = (long long)blockIdx.x * (blockDim.x >> 6) + wave;
  if (t >= P.ntasks) return;
  const long long task = launch_task(P, t);
  const Task &T = tasks[task];
  const TaskL l = task_l(P, T);
  const WaveLds W = wave_lds(lds, P.wp, wave);
  const double *in = P.cxyz_all + task * P.cxyz_stride;
  const int nc = (l.lp + 1) * (l.lp + 2) * (l.lp + 3) / 6;
  for (int e = lane; e < nc; e += 64) W.cxyz[e] = in[e];  // packed
  wave_sync();
  task_finish<KIND, 64>(P, T, groups[T.group], l.la, l.lb, l.la_min, l.lb_min, W, nullptr, task,
                        lane);
}

// The same per task with a whole workgroup (256 threads) on one task: for
// high l, where one task's tables fill most of LDS.
template <int KIND>
__global__ __launch_bounds__(256) void finish_block_kernel(const Task *__restrict__ tasks,
                                                           const Group *__restrict__ groups,
                                                           PointParams P) {
  extern __shared__ double lds[];
  const long long task = launch_task(P, blockIdx.x);
  const Task &T = tasks[task];
  const TaskL l = task_l(P, T);
  const WaveLds W = wave_lds(lds, P.wp, 0);
  double *red = lds + P.wp.wave_bytes / 8;
  const double *in = P.cxyz_all + task * P.cxyz_stride;
  const int nc = (l.lp + 1) * (l.lp + 2) * (l.lp + 3) / 6;
  for (int e = threadIdx.x; e < nc; e += 256) W.cxyz[e] = in[e];  // packed
  __syncthreads();
  task_finish<KIND, 256>(P, T, groups[T.group], l.la, l.lb, l.la_min, l.lb_min, W, red, task,
                         threadIdx.x);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <typename K, typename... A>
static hipError_t launch_wave(K kernel, const PointParams &P, int blocks, hipStream_t s,
                              A... args) {
  const size_t lds = (size_t)P.wp.bytes;
  if (lds > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(64 * P.wp.nwaves), lds, s, args..., P);
  return hipGetLastError();
}

static int wave_blocks(const PointParams &P) { return (P.ntasks + P.wp.nwaves - 1) / P.wp.nwaves; }

// Workgroups of a persistent launch (work queue): as many as are resident at
// once on the device (occupancy of the kernel with its LDS), at most `cap`.
template <typename K>
static int persistent_blocks(K kernel, int threads, size_t lds, int cap) {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 1;
  }
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(kernel),
                                                   threads, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const long long n = (long long)per_cu * cus;
  return (int)(n < cap ? n : cap);
}

hipError_t launch_cxyz_lane(const PointParams &P, hipStream_t s) {
  if (P.ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(cxyz_lane_kernel, dim3((P.ntasks + 255) / 256), dim3(256), 0, s, P.tasks,
                     P.groups, P);
  return hipGetLastError();
}

hipError_t launch_cxyz_wave(const PointParams &P, hipStream_t s) {
  if (P.ntasks <= 0) return hipSuccess;
  if (P.wp.nwaves < 1 || P.wp.nwaves > 4 || P.wp.window != 0) return hipErrorInvalidValue;
  return launch_wave(cxyz_wave_kernel, P, wave_blocks(P), s, P.tasks, P.groups);
}

template <typename R>
static hipError_t collocate_wave_r(const PointParams &P, int lpb, hipStream_t s) {
  if (lpb == 4) return launch_wave(collocate_wave_kernel<4, R>, P, wave_blocks(P), s, P.tasks, P.levels);
  if (lpb == 6) return launch_wave(collocate_wave_kernel<6, R>, P, wave_blocks(P), s, P.tasks, P.levels);
  if (lpb == 8) return launch_wave(collocate_wave_kernel<8, R>, P, wave_blocks(P), s, P.tasks, P.levels);
  return hipErrorInvalidValue;
}

hipError_t launch_collocate_wave(const PointParams &P, int lpb, hipStream_t s) {
  if (P.ntasks <= 0) return hipSuccess;
  if (P.wp.nwaves < 1 || P.wp.nwaves > 4 || P.wp.window != 0) return hipErrorInvalidValue;
  count_launch(8, P.fp32);
  return P.fp32 ? collocate_wave_r<float>(P, lpb, s) : collocate_wave_r<double>(P, lpb, s);
}

template <typename R>
static hipError_t collocate_tiles_r(const PointParams &P, int lpb, int nunits, hipStream_t s) {
  auto go = [&](auto kernel) {
    const int blocks = P.queue ? persistent_blocks(kernel, 64 * P.wp.nwaves, P.wp.bytes, nunits)
                               : nunits;
    return launch_wave(kernel, P, blocks, s, P.tasks, P.levels, P.units, P.items);
  };
  if (P.queue) {
    if (lpb == 4) return go(collocate_tile_kernel<4, R, true>);
    if (lpb == 6) return go(collocate_tile_kernel<6, R, true>);
    if (lpb == 8) return go(collocate_tile_kernel<8, R, true>);
  } else {
    if (lpb == 4) return go(collocate_tile_kernel<4, R, false>);
    if (lpb == 6) return go(collocate_tile_kernel<6, R, false>);
    if (lpb == 8) return go(collocate_tile_kernel<8, R, false>);
  }
  return hipErrorInvalidValue;
}

// nunits workgroups from unit P.unit_base.
hipError_t launch_collocate_tiles(const PointParams &Pin, int lpb, int nunits, hipStream_t s) {
  if (nunits <= 0) return hipSuccess;
  if (Pin.wp.nwaves < 1 || Pin.wp.nwaves > (lpb <= 4 ? 12 : 8) || Pin.wp.window <= 0)
    return hipErrorInvalidValue;
  PointParams P = Pin;
  P.nunits = nunits;
  count_launch(7, P.fp32);
  return P.fp32 ? collocate_tiles_r<float>(P, lpb, nunits, s)
                : collocate_tiles_r<double>(P, lpb, nunits, s);
}

template <typename R>
static hipError_t integrate_wave_r(const PointParams &P, int lpb, hipStream_t s) {
  auto go = [&](auto kernel) {
    const int blocks = P.queue ? persistent_blocks(kernel, 64 * P.wp.nwaves, P.wp.bytes,
                                                   wave_blocks(P))
                               : wave_blocks(P);
    return launch_wave(kernel, P, blocks, s, P.tasks, P.levels);
  };
  if (P.queue) {
    if (lpb == 4) return go(integrate_wave_kernel<4, R, true>);
    if (lpb == 6) return go(integrate_wave_kernel<6, R, true>);
  } else {
    if (lpb == 4) return go(integrate_wave_kernel<4, R, false>);
    if (lpb == 6) return go(integrate_wave_kernel<6, R, false>);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_integrate_wave(const PointParams &P, int lpb, hipStream_t s) {
  if (P.ntasks <= 0) return hipSuccess;
  if (P.wp.nwaves < 1 || P.wp.nwaves > 4 || P.wp.window != 0) return hipErrorInvalidValue;
  count_launch(9, P.fp32);
  return P.fp32 ? integrate_wave_r<float>(P, lpb, s) : integrate_wave_r<double>(P, lpb, s);
}

hipError_t launch_integrate_mfma(const PointParams &P, hipStream_t s) {
  if (P.ntasks <= 0) return hipSuccess;
  if (P.wp.nwaves < 1 || P.wp.nwaves > 4 || P.wp.window != 0 || P.wp.np > 12)
    return hipErrorInvalidValue;
  count_launch(9, false);  // FP64 in both modes
  return launch_wave(integrate_mfma_kernel, P, wave_blocks(P), s, P.tasks, P.levels);
}

// One workgroup per task; P.wp: the setup plan (one region of wave_bytes).
hipError_t launch_finish_block(const PointParams &P, hipStream_t s) {
  if (P.ntasks <= 0) return hipSuccess;
  if (P.wp.window != 0) return hipErrorInvalidValue;
  const bool full = P.do_forces || P.compute_tau;
  const size_t lds = (size_t)P.wp.wave_bytes + 8 * 4 * kTaskOutStride;
  auto go = [&](auto kernel) {
    if (lds > 65536) {
      const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kernel, dim3(P.ntasks), dim3(256), lds, s, P.tasks, P.groups, P);
    return hipGetLastError();
  };
  if (P.compute_tau) return go(finish_block_kernel<2>);
  return full ? go(finish_block_kernel<1>) : go(finish_block_kernel<0>);
}

// The lane kernel serves calls with forces, without virial and tau, for
// la_max + lb_max <= 3 (lp: after the call's shift); false: use the wave
// kernel.
bool finish_lane_ok(const PointParams &P, int lp) {
  const int lp0 = lp - P.ld[0] - P.ld[2];
  // without forces the wave kernel measured faster (hab only: few loads per
  // task, the lane kernel's scattered atomics dominate)
  return P.do_forces && !P.do_virial && !P.compute_tau && lp0 <= 3 && P.ld[2] == 0 &&
         P.ld[0] == 1;
}

hipError_t launch_finish_lane(const PointParams &P, hipStream_t s) {
  if (P.ntasks <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((P.ntasks + 255) / 256);
  if (P.do_forces)
    hipLaunchKernelGGL(finish_lane_kernel<true>, dim3(blocks), dim3(256), 0, s, P.tasks, P.groups, P);
  else
    hipLaunchKernelGGL(finish_lane_kernel<false>, dim3(blocks), dim3(256), 0, s, P.tasks, P.groups, P);
  return hipGetLastError();
}

hipError_t launch_finish_wave(const PointParams &P, hipStream_t s) {
  if (P.ntasks <= 0) return hipSuccess;
  if (P.wp.nwaves < 1 || P.wp.nwaves > 4 || P.wp.window != 0) return hipErrorInvalidValue;
  if (P.compute_tau) return launch_wave(finish_wave_kernel<2>, P, wave_blocks(P), s, P.tasks, P.groups);
  return P.do_forces ? launch_wave(finish_wave_kernel<1>, P, wave_blocks(P), s, P.tasks, P.groups)
                     : launch_wave(finish_wave_kernel<0>, P, wave_blocks(P), s, P.tasks, P.groups);
}

}  // namespace gm
