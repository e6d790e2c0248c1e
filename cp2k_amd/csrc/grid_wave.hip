// Wave-per-task kernels for orthorhombic tasks (gfx950, wave64).
//
// One wave processes one task at a time, with no workgroup barriers: the
// task record is read through __restrict__ pointers into scalar registers
// and every per-task table lives in the wave's own LDS region:
//   cab -> cxyz about rp (collocate; src/grid/ref/grid_ref_collint.h:877-910)
//   tab[d][c][l] = (c h_d - roff_d)^l exp(-zetp (c h_d - roff_d)^2)
//   i0[|k'|][|j'|] (int8): first x of row (j, k) of the truncation region
//   column starts of the central plane
// The points are walked column by column.  Lanes take consecutive (i, j)
// columns of the central plane (x fastest, so a wave's loads and adds are
// contiguous along x) and step through the planes k together.  Membership
// repeats the reference's bounds (grid_ref_collint.h:139-167): plane k keeps
// the rows with j' >= js(k) and row (j, k) the points with i' >= i0(j, k),
// where i' = i for i <= 0 and 1 - i otherwise.  Both bounds only grow with
// |k'| (every operation in them is monotone), so the planes of a column form
// the interval [kl, 1 - kl]; kl is found by bisection in the i0 table.
//
// Collocate: per column D[lz] = sum_{lx,ly} cxyz px py, then per point
// (lp + 1) fused multiply-adds and one atomic add, either into an LDS copy of
// the whole level grid shared by the workgroup (flushed to HBM once at the
// end) or straight into HBM.  Integrate: per point (lp + 1) multiply-adds of
// the z moments, per column the outer product with px, py into lane-private
// cxyz accumulators, and one cross-lane reduction per task.
//
// Compiled with -ffp-contract=off; fused multiply-adds are explicit fma().

#include <hip/hip_runtime.h>

#include "grid_device.h"
#include "grid_internal.h"

namespace gm {
namespace {

// LDS written by one lane and read by another lane of the same wave: LDS
// requests of a wave complete in order, so only compiler reordering must be
// prevented.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, true);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes (all active); the total is valid in lane 63.
__device__ __forceinline__ double wave_total(double v) {
  v += dpp_mov<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += dpp_mov<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0x141, 0xf>(v);  // row_half_mirror
  v += dpp_mov<0x140, 0xf>(v);  // row_mirror
  v += dpp_mov<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_mov<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ int pm(int a, int m) {
  const int r = a % m;
  return r < 0 ? r + m : r;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Per-wave LDS region.
struct WaveLds {
  double *cxyz, *u;  // u: {cab, alpha} during setup/finish, tables while walking
  signed char *i0;
  int *col;
};

__device__ __forceinline__ WaveLds wave_lds(double *lds, const WavePlan &wp, int wave) {
  char *b = reinterpret_cast<char *>(lds) + 8 * (size_t)wp.window + (size_t)wave * wp.wave_bytes;
  WaveLds w;
  w.cxyz = reinterpret_cast<double *>(b) + wp.o_cxyz;
  w.u = reinterpret_cast<double *>(b) + wp.o_u;
  w.i0 = reinterpret_cast<signed char *>(b + wp.o_i0);
  w.col = reinterpret_cast<int *>(b + wp.o_col);
  return w;
}

// Row-start table and the column starts of the central plane (k' = 0, the
// plane that contains every column).  Returns the number of columns.
__device__ int wave_geometry(const Task &T, const Layout &lay, const WavePlan &wp,
                             const WaveLds &W, int lane) {
  const int nkz = 1 - T.lo[2], njy = 1 - T.lo[1];
  const double dz = lay.dh[2][2], dy = lay.dh[1][1];
  const double dyinv = lay.dh_inv[1][1], dxinv = lay.dh_inv[0][0];
  for (int e = lane; e < nkz * njy; e += 64) {
    const int kz = e / njy, jy = e - kz * njy;
    const int k1 = T.lo[2] + kz, j1 = T.lo[1] + jy;
    const double kr = k1 * dz;
    const double krem = T.disr * T.disr - kr * kr;
    int i0 = 127;
    if (j1 >= ortho_start(krem, dyinv)) {
      const double jr = j1 * dy;
      const double jrem = krem - jr * jr;
      i0 = ortho_start(jrem, dxinv);
    }
    W.i0[kz * wp.i0n + jy] = (signed char)i0;
  }
  wave_sync();
  const int ny_ext = 2 * njy;  // <= 64, checked on the host
  int len = 0;
  if (lane < ny_ext) {
    const int j = T.lo[1] + lane;
    const int j1 = j <= 0 ? j : 1 - j;
    const int i0 = W.i0[(nkz - 1) * wp.i0n + (j1 - T.lo[1])];
    len = i0 == 127 ? 0 : 2 - 2 * i0;
  }
  int incl = len;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane < ny_ext) W.col[lane + 1] = incl;
  if (lane == 0) W.col[0] = 0;
  const int ncol = __shfl(incl, ny_ext - 1, 64);
  wave_sync();
  return ncol;
}

// Per-axis tables, tab[(d * ext + c - lo_d) * np + l].
__device__ void wave_tables(const Task &T, const Layout &lay, int lp, const WavePlan &wp,
                            double *tab, int lane) {
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
    double *row = tab + (d * wp.ext + q) * wp.np;
    for (int l = 0; l <= lp; l++) {
      row[l] = v;
      v *= t;
    }
  }
  wave_sync();
}

// Column c of the central plane -> cube indices, plane interval, grid base.
struct Column {
  int i, j;  // cube indices
  int kl;    // planes kl .. 1 - kl; kl = 1 marks a lane without a column
  int base;  // y * nx + x of the column in the local grid
};

__device__ __forceinline__ Column decode_column(const Task &T, const Layout &lay,
                                                const WavePlan &wp, const WaveLds &W,
                                                int ncol, int c) {
  Column C{0, 0, 1, 0};
  if (c >= ncol) return C;
  const int nkz = 1 - T.lo[2], ny_ext = 2 - 2 * T.lo[1];
  int lo = 0, hi = ny_ext;  // col[ny_ext] = ncol > c
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (W.col[mid] <= c) lo = mid; else hi = mid;
  }
  const int j = T.lo[1] + lo;
  const int jy = (j <= 0 ? j : 1 - j) - T.lo[1];
  const int i = W.i0[(nkz - 1) * wp.i0n + jy] + (c - W.col[lo]);
  const int i1 = i <= 0 ? i : 1 - i;
  int a = 0, b = nkz - 1;  // the central plane always holds the column
  while (a < b) {
    const int m = (a + b) >> 1;
    if (W.i0[m * wp.i0n + jy] <= i1) b = m; else a = m + 1;
  }
  C.i = i;
  C.j = j;
  C.kl = T.lo[2] + a;
  const int gx = pm(T.center[0] + i - lay.shift_local[0], lay.npts_global[0]);
  const int gy = pm(T.center[1] + j - lay.shift_local[1], lay.npts_global[1]);
  C.base = gy * lay.npts_local[0] + gx;
  return C;
}

// cab (prepare_pab of the decontracted density) -> cxyz, in the wave's LDS.
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
    for (int lz = 0; lz <= A.z + B.z; lz++)
      for (int ly = 0; ly <= A.y + B.y; ly++)
        for (int lx = 0; lx <= A.x + B.x; lx++)
          atomicAdd(&W.cxyz[(lz * NP + ly) * NP + lx], cv * (ax[lx] * ay[ly] * az[lz] * pref));
  }
  wave_sync();
}

template <int LP>
__device__ __forceinline__ void load_pxy(const Task &T, const WavePlan &wp, const double *tab,
                                         const Column &C, double px[LP + 1], double py[LP + 1]) {
  const int np = wp.np;
  const double *rx = tab + clampi(C.i - T.lo[0], 0, wp.ext - 1) * np;
  const double *ry = tab + (wp.ext + clampi(C.j - T.lo[1], 0, wp.ext - 1)) * np;
#pragma unroll
  for (int l = 0; l <= LP; l++) {
    px[l] = rx[l];
    py[l] = ry[l];
  }
}

// Collocate the task's points; dst is the LDS window or the HBM level grid
// (npts points either way).
template <int LP, typename Dst>
__device__ void collocate_columns(const Task &T, const Layout &lay, const WavePlan &wp,
                                  const WaveLds &W, int ncol, Dst *dst, int npts, int lane) {
  const int NP = wp.np;
  const double *tab = W.u;
  const double *tabz = tab + 2 * wp.ext * NP;
  const int nxy = lay.npts_local[0] * lay.npts_local[1];
  const int ngz = lay.npts_global[2];
  for (int c0 = 0; c0 < ncol; c0 += 64) {
    const Column C = decode_column(T, lay, wp, W, ncol, c0 + lane);
    double px[LP + 1], py[LP + 1], D[LP + 1];
    load_pxy<LP>(T, wp, tab, C, px, py);
#pragma unroll
    for (int lz = 0; lz <= LP; lz++) {
      double s = 0.0;
#pragma unroll
      for (int ly = 0; ly <= LP - lz; ly++) {
        double t = 0.0;
#pragma unroll
        for (int lx = 0; lx <= LP - lz - ly; lx++)
          t = fma(W.cxyz[(lz * NP + ly) * NP + lx], px[lx], t);
        s = fma(t, py[ly], s);
      }
      D[lz] = s;
    }
    const int kmin = wave_min(C.kl);
    int gz = pm(T.center[2] + kmin - lay.shift_local[2], ngz);
    for (int k = kmin; k <= 1 - kmin; k++) {
      const double *pz = tabz + (k - T.lo[2]) * NP;
      double v = 0.0;
#pragma unroll
      for (int lz = 0; lz <= LP; lz++) v = fma(D[lz], pz[lz], v);
      const int idx = gz * nxy + C.base;
      if (k >= C.kl && k <= 1 - C.kl && idx < npts) atomicAdd(&dst[idx], v);
      gz = (gz + 1 == ngz) ? 0 : gz + 1;
    }
  }
}

// Integrate: lane-private cxyz accumulators acc[(lz, ly, lx)], lx fastest.
template <int LP>
__device__ void integrate_columns(const Task &T, const Layout &lay, const WavePlan &wp,
                                  const WaveLds &W, int ncol, const double *grid, int npts,
                                  double *acc, int lane) {
  const int NP = wp.np;
  const double *tab = W.u;
  const double *tabz = tab + 2 * wp.ext * NP;
  const int nxy = lay.npts_local[0] * lay.npts_local[1];
  const int ngz = lay.npts_global[2];
  for (int c0 = 0; c0 < ncol; c0 += 64) {
    const Column C = decode_column(T, lay, wp, W, ncol, c0 + lane);
    double px[LP + 1], py[LP + 1], M[LP + 1];
    load_pxy<LP>(T, wp, tab, C, px, py);
#pragma unroll
    for (int l = 0; l <= LP; l++) M[l] = 0.0;
    const int kmin = wave_min(C.kl);
    int gz = pm(T.center[2] + kmin - lay.shift_local[2], ngz);
    for (int k = kmin; k <= 1 - kmin; k++) {
      const double *pz = tabz + (k - T.lo[2]) * NP;
      const int idx = gz * nxy + C.base;
      double g = 0.0;
      if (k >= C.kl && k <= 1 - C.kl && idx < npts) g = grid[idx];
#pragma unroll
      for (int l = 0; l <= LP; l++) M[l] = fma(g, pz[l], M[l]);
      gz = (gz + 1 == ngz) ? 0 : gz + 1;
    }
    int e = 0;
#pragma unroll
    for (int lz = 0; lz <= LP; lz++)
#pragma unroll
      for (int ly = 0; ly <= LP - lz; ly++) {
        const double yz = py[ly] * M[lz];
#pragma unroll
        for (int lx = 0; lx <= LP - lz - ly; lx++, e++) acc[e] = fma(px[lx], yz, acc[e]);
      }
  }
}

template <int LP>
__device__ void integrate_moments(const Task &T, const Layout &lay, const WavePlan &wp,
                                  const WaveLds &W, int ncol, const double *grid, int npts,
                                  int lane) {
  constexpr int NC = (LP + 1) * (LP + 2) * (LP + 3) / 6;
  double acc[NC];
#pragma unroll
  for (int e = 0; e < NC; e++) acc[e] = 0.0;
  integrate_columns<LP>(T, lay, wp, W, ncol, grid, npts, acc, lane);
  const int NP = wp.np;
  for (int e = lane; e < NP * NP * NP; e += 64) W.cxyz[e] = 0.0;
  wave_sync();
  int e = 0;
#pragma unroll
  for (int lz = 0; lz <= LP; lz++)
#pragma unroll
    for (int ly = 0; ly <= LP - lz; ly++)
#pragma unroll
      for (int lx = 0; lx <= LP - lz - ly; lx++, e++) {
        const double s = wave_total(acc[e]);
        if (lane == 63) W.cxyz[(lz * NP + ly) * NP + lx] = s;
      }
  wave_sync();
}

// cxyz -> cab -> hab (Cartesian, atomics into the group's slice), forces and
// virial per task (grid_ref_integrate.c:132-171).
template <bool FULL>
__device__ void wave_finish(const PointParams &P, const Task &T, const Group &G, int la, int lb,
                            int la_min, int lb_min, const WavePlan &wp, const WaveLds &W,
                            long long task, int lane) {
  const int NP = wp.np;
  const int n1 = ncoset(la), n2 = ncoset(lb);
  double *cab = W.u, *alpha = W.u + n1 * n2;
  build_alpha(T, la, lb, alpha, lane, 64);
  wave_sync();
  const int np = la + lb + 1, nA = la + 1, nB = lb + 1;
  const int a1 = ncoset(la_min - 1), ma = ncoset(la) - a1;
  const int b1 = ncoset(lb_min - 1), mb = ncoset(lb) - b1;
  for (int e = lane; e < n1 * n2; e += 64) {
    const int ia = e % n1, jb = e / n1;
    double s = 0.0;
    if (ia >= a1 && ia < a1 + ma && jb >= b1 && jb < b1 + mb) {
      const Ang A = orb(ia), B = orb(jb);
      const double *ax = &alpha[((0 * nB + B.x) * nA + A.x) * np];
      const double *ay = &alpha[((1 * nB + B.y) * nA + A.y) * np];
      const double *az = &alpha[((2 * nB + B.z) * nA + A.z) * np];
      for (int lz = 0; lz <= A.z + B.z; lz++)
        for (int ly = 0; ly <= A.y + B.y; ly++)
          for (int lx = 0; lx <= A.x + B.x; lx++)
            s += W.cxyz[(lz * NP + ly) * NP + lx] * (ax[lx] * ay[ly] * az[lz] * T.pexp);
    }
    cab[e] = s;
  }
  wave_sync();
  const CabGet cg{cab, n1};
  const bool tau = FULL && P.compute_tau != 0;
  const double za = T.zeta, zb = T.zetb;
  const int a0 = ncoset(T.la_min - 1), na = ncoset(T.la_max) - a0;
  const int b0 = ncoset(T.lb_min - 1), nb = ncoset(T.lb_max) - b0;
  double f[6] = {0, 0, 0, 0, 0, 0};
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int e = lane; e < na * nb; e += 64) {
    const int ia = a0 + e % na, jb = b0 + e / na;
    const Ang A = orb(ia), B = orb(jb);
    const long long q = G.cart_off + (long long)(T.o2 + jb) * G.ncoa + T.o1 + ia;
    if (q < P.cart_len) atomicAdd(&P.hab_cart[q], get_hab(A, B, za, zb, cg, tau));
    if (FULL && P.do_forces) {
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
    for (int c = 0; c < kTaskOutStride; c++) {
      const double s = wave_total(c < 6 ? f[c] : v[c - 6]);
      if (lane == 63) out[c] = scale * s;
    }
  }
  wave_sync();
}

}  // namespace

// ---------------------------------------------------------------------------
// Kernels.  Launch ranges hold orthorhombic tasks of level wp.level with
// lp <= LPB and cube extents <= kWaveMaxExt.
// ---------------------------------------------------------------------------
template <int LPB, bool WINDOW>
__global__ __launch_bounds__(512) void collocate_wave_kernel(const Task *__restrict__ tasks,
                                                             const Level *__restrict__ levels,
                                                             const Group *__restrict__ groups,
                                                             PointParams P) {
  extern __shared__ double lds[];
  const WavePlan &wp = P.wp;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  const Level &L = levels[wp.level];
  const Layout &lay = L.lay;
  const int npts = (int)L.npts;
  if (WINDOW) {
    for (int e = threadIdx.x; e < wp.window; e += blockDim.x) lds[e] = 0.0;
    __syncthreads();
  }
  const WaveLds W = wave_lds(lds, wp, wave);
  const long long stride = (long long)gridDim.x * nw;
  for (long long t = (long long)blockIdx.x * nw + wave; t < P.ntasks; t += stride) {
    const long long task = P.task_base + t;
    const Task &T = tasks[task];
    const int la = T.la_max + P.ld[0], lb = T.lb_max + P.ld[2];
    const int la_min = max(T.la_min + P.ld[1], 0), lb_min = max(T.lb_min + P.ld[3], 0);
    const int lp = la + lb;
    if (!P.count_only) wave_cxyz(P, T, groups[T.group], la, lb, la_min, lb_min, wp.np, W, lane);
    const int ncol = wave_geometry(T, lay, wp, W, lane);
    if (P.count_only) {
      double cnt = 0.0;
      for (int c0 = 0; c0 < ncol; c0 += 64) {
        const Column C = decode_column(T, lay, wp, W, ncol, c0 + lane);
        cnt += (double)(C.kl <= 0 ? 2 - 2 * C.kl : 0);
      }
      cnt = wave_total(cnt);
      if (lane == 63) P.task_out[task * kTaskOutStride] = cnt;
      wave_sync();
      continue;
    }
    wave_tables(T, lay, lp, wp, W.u, lane);
    double *dst = WINDOW ? lds : L.grid;
    switch (lp) {
      case 0: collocate_columns<0>(T, lay, wp, W, ncol, dst, npts, lane); break;
      case 1: collocate_columns<1>(T, lay, wp, W, ncol, dst, npts, lane); break;
      case 2: collocate_columns<2>(T, lay, wp, W, ncol, dst, npts, lane); break;
      case 3: collocate_columns<3>(T, lay, wp, W, ncol, dst, npts, lane); break;
      case 4: collocate_columns<4>(T, lay, wp, W, ncol, dst, npts, lane); break;
      case 5: if (LPB >= 5) collocate_columns<(LPB >= 5 ? 5 : 0)>(T, lay, wp, W, ncol, dst, npts, lane); break;
      case 6: if (LPB >= 6) collocate_columns<(LPB >= 6 ? 6 : 0)>(T, lay, wp, W, ncol, dst, npts, lane); break;
      default: break;  // the host routes larger lp to the per-task kernels
    }
    wave_sync();
  }
  if (WINDOW) {
    __syncthreads();
    for (int e = threadIdx.x; e < wp.window; e += blockDim.x) {
      const double v = lds[e];
      if (v != 0.0) atomicAdd(&L.grid[e], v);
    }
  }
}

// Waves per SIMD the register allocation must allow: the force/virial
// getters of FULL need more registers than the hab-only finish.
template <int LPB, bool FULL>
constexpr int integrate_occupancy() { return FULL ? (LPB <= 4 ? 2 : 1) : (LPB <= 4 ? 3 : 2); }

template <int LPB, bool FULL>
__global__ __launch_bounds__(256, (integrate_occupancy<LPB, FULL>())) void integrate_wave_kernel(
    const Task *__restrict__ tasks, const Level *__restrict__ levels,
    const Group *__restrict__ groups, PointParams P) {
  extern __shared__ double lds[];
  const WavePlan &wp = P.wp;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  const long long t = (long long)blockIdx.x * nw + wave;
  if (t >= P.ntasks) return;
  const long long task = P.task_base + t;
  const Task &T = tasks[task];
  const Level &L = levels[wp.level];
  const Layout &lay = L.lay;
  const int npts = (int)L.npts;
  const WaveLds W = wave_lds(lds, wp, wave);
  const int la = T.la_max + P.ld[0], lb = T.lb_max + P.ld[2];
  const int la_min = max(T.la_min + P.ld[1], 0), lb_min = max(T.lb_min + P.ld[3], 0);
  const int lp = la + lb;
  const int ncol = wave_geometry(T, lay, wp, W, lane);
  wave_tables(T, lay, lp, wp, W.u, lane);
  const double *grid = L.grid;
  switch (lp) {
    case 0: integrate_moments<0>(T, lay, wp, W, ncol, grid, npts, lane); break;
    case 1: integrate_moments<1>(T, lay, wp, W, ncol, grid, npts, lane); break;
    case 2: integrate_moments<2>(T, lay, wp, W, ncol, grid, npts, lane); break;
    case 3: integrate_moments<3>(T, lay, wp, W, ncol, grid, npts, lane); break;
    case 4: integrate_moments<4>(T, lay, wp, W, ncol, grid, npts, lane); break;
    case 5: if (LPB >= 5) integrate_moments<(LPB >= 5 ? 5 : 0)>(T, lay, wp, W, ncol, grid, npts, lane); break;
    case 6: if (LPB >= 6) integrate_moments<(LPB >= 6 ? 6 : 0)>(T, lay, wp, W, ncol, grid, npts, lane); break;
    default: return;  // the host routes larger lp to the per-task kernels
  }
  wave_finish<FULL>(P, T, groups[T.group], la, lb, la_min, lb_min, wp, W, task, lane);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <typename K>
static hipError_t launch_wave(K kernel, const PointParams &P, int blocks, hipStream_t s) {
  const size_t lds = (size_t)P.wp.bytes;
  if (lds > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(64 * P.wp.nwaves), lds, s, P.tasks, P.levels,
                     P.groups, P);
  return hipGetLastError();
}

// Collocation (or point counting) of P.ntasks tasks from P.task_base; lpb is
// 4 or 6; blocks is the grid size (any size: waves loop over the range).
hipError_t launch_collocate_wave(const PointParams &P, int lpb, int blocks, hipStream_t s) {
  if (P.ntasks <= 0) return hipSuccess;
  if (P.wp.nwaves < 1 || P.wp.nwaves > 8 || blocks < 1) return hipErrorInvalidValue;
  const bool win = P.wp.window > 0;
  if (lpb == 4)
    return win ? launch_wave(collocate_wave_kernel<4, true>, P, blocks, s)
               : launch_wave(collocate_wave_kernel<4, false>, P, blocks, s);
  if (lpb == 6)
    return win ? launch_wave(collocate_wave_kernel<6, true>, P, blocks, s)
               : launch_wave(collocate_wave_kernel<6, false>, P, blocks, s);
  return hipErrorInvalidValue;
}

// Integration: one wave per task, ceil(ntasks / nwaves) workgroups.
hipError_t launch_integrate_wave(const PointParams &P, int lpb, hipStream_t s) {
  if (P.ntasks <= 0) return hipSuccess;
  if (P.wp.nwaves < 1 || P.wp.nwaves > 4 || P.wp.window != 0) return hipErrorInvalidValue;
  const int blocks = (P.ntasks + P.wp.nwaves - 1) / P.wp.nwaves;
  const bool full = P.do_forces || P.compute_tau;
  if (lpb == 4)
    return full ? launch_wave(integrate_wave_kernel<4, true>, P, blocks, s)
                : launch_wave(integrate_wave_kernel<4, false>, P, blocks, s);
  if (lpb == 6)
    return full ? launch_wave(integrate_wave_kernel<6, true>, P, blocks, s)
                : launch_wave(integrate_wave_kernel<6, false>, P, blocks, s);
  return hipErrorInvalidValue;
}

}  // namespace gm
