// HIP kernels of the MI355X grid backend (gfx950, wave64).
//
// One workgroup owns one task (a product of two primitive Gaussians on one
// grid level).  Per task, in LDS and FP64:
//   prepare_pab stencil            -> cab        (src/grid/common/grid_prepare_pab.h)
//   binomial re-centring tables    -> alpha      (src/grid/ref/grid_ref_collint.h:836-861)
//   cab -> cxyz polynomial about rp              (grid_ref_collint.h:877-910)
//   per-axis Gaussian x monomial tables, or the lattice-coordinate
//   re-expansion cxyz -> cijk for the general case (grid_ref_collint.h:697-762)
// The truncation region is then walked in chunks of planes.  Plane and row
// bounds repeat the reference's floating-point operations exactly
// (orthorhombic: grid_ref_collint.h:139-167, 44-52; general: :591-664,
// :431-464), so the same set of (task, grid point) pairs is visited.
//
// Collocate flattens the points of a chunk across the lanes (consecutive
// lanes -> consecutive x -> coalesced FP64 atomics into the level grid).
// Integrate gives each lane whole rows (register accumulation along x) and
// contracts rows -> planes -> cxyz in LDS, then cxyz -> cab -> hab, forces
// and virial (src/grid/common/grid_process_vab.h).
//
// This file is compiled with -ffp-contract=off; multiply-adds that may fuse
// are written with fma() explicitly.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "grid_internal.h"

#include "grid_device.h"

namespace gm {
void count_launch(int id, bool fp32);


// Chunk bookkeeping in LDS ints.
struct Chunk {
  int *pk;    // plane index (k) per chunk plane       [plcap+1]
  int *prow;  // first row of each chunk plane (+1)     [plcap+1]
  int *pjs;   // first j of each plane                  [plcap+1]
  int *rj;    // row j                                  [rowcap+1]
  int *rp;    // row plane (chunk index)                [rowcap+1]
  int *ri;    // row first i                            [rowcap+1]
  int *roff;  // row length, then exclusive offsets     [rowcap+1]
  int *misc;  // [0] planes, [1] rows, [2..] scan scratch
  int4 *row4; // orthorhombic rows: {first i, length, grid row base, plane<<16 | j+32768}
};

__device__ __forceinline__ Chunk chunk_arrays(double *lds, int *ib, const LdsPlan &pl) {
  Chunk c;
  c.row4 = reinterpret_cast<int4 *>(lds + pl.o_row4);
  int *p = ib + pl.i_plane;
  c.pk = p;
  c.prow = p + (pl.plcap + 1);
  c.pjs = p + 2 * (pl.plcap + 1);
  int *r = ib + pl.i_row;
  c.rj = r;
  c.rp = r + (pl.rowcap + 1);
  c.ri = r + 2 * (pl.rowcap + 1);
  c.roff = r + 3 * (pl.rowcap + 1);
  c.misc = ib + pl.i_misc;
  return c;
}

__device__ __forceinline__ int find_plane(const int *prow, int npl, int r) {
  int lo = 0, hi = npl;  // prow[npl] = nrows > r
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prow[mid] <= r) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int find_row(const int *roff, int nrows, int p) {
  int lo = 0, hi = nrows;  // roff[nrows] = npts > p
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (roff[mid] <= p) lo = mid; else hi = mid;
  }
  return lo;
}

// Builds the next chunk starting at plane kbeg: planes, rows (j, plane, first
// i, length).  Returns number of planes consumed (>= 1).  Ends synchronised.
template <int NW>
__device__ int build_chunk(const Task &T, const Layout &lay, bool ortho, int kbeg,
                           const LdsPlan &pl, const Chunk &C, const int *map, int tid,
                           int nthr) {
  const int khi = ortho ? 1 - T.lo[2] : T.hi[2];
  if (tid == 0) {
    int rows = 0, npl = 0, k = kbeg;
    while (k <= khi && npl < pl.plcap) {
      int js, nr;
      if (ortho) {
        const int k1 = k <= 0 ? k : 1 - k;
        const double kr = k1 * lay.dh[2][2];
        const double krem = T.disr * T.disr - kr * kr;
        js = ortho_start(krem, lay.dh_inv[1][1]);
        nr = 2 - 2 * js;
      } else {
        js = T.lo[1];
        nr = (map[2 * pl.mapn + (k - T.lo[2])] >= 0) ? T.hi[1] - T.lo[1] + 1 : 0;
      }
      if (npl > 0 && rows + nr > pl.rowcap) break;
      C.pk[npl] = k;
      C.pjs[npl] = js;
      C.prow[npl] = rows;
      rows += nr;
      npl++;
      k++;
    }
    C.prow[npl] = rows;
    C.misc[0] = npl;
    C.misc[1] = rows;
  }
  __syncthreads();
  const int npl = C.misc[0], nrows = C.misc[1];
  for (int r = tid; r < nrows; r += nthr) {
    const int p = find_plane(C.prow, npl, r);
    const int j = C.pjs[p] + (r - C.prow[p]);
    const int k = C.pk[p];
    int i0 = 0, len = 0;
    if (ortho) {
      const int k1 = k <= 0 ? k : 1 - k;
      const double kr = k1 * lay.dh[2][2];
      const double krem = T.disr * T.disr - kr * kr;
      const int j1 = j <= 0 ? j : 1 - j;
      const double jr = j1 * lay.dh[1][1];
      const double jrem = krem - jr * jr;
      i0 = ortho_start(jrem, lay.dh_inv[0][0]);
      len = 2 - 2 * i0;
    } else if (map[1 * pl.mapn + (j - T.lo[1])] >= 0) {
      int ismin, ismax;
      if (general_row(lay.dh, T.gp, j, k, T.radius, &ismin, &ismax)) {
        // The sphere lies inside the index box; the clamp guards rounding.
        ismin = ismin < T.lo[0] ? T.lo[0] : ismin;
        ismax = ismax > T.hi[0] ? T.hi[0] : ismax;
        i0 = ismin;
        len = ismax >= ismin ? ismax - ismin + 1 : 0;
      }
    }
    C.rj[r] = j;
    C.rp[r] = p;
    C.ri[r] = i0;
    C.roff[r] = len;
    if (ortho) {
      // one 16-byte record per row for the column-owning point loops; the
      // host guarantees level grids below 2^31 points
      const int rowbase = (map[2 * pl.mapn + k + T.cmax] * lay.npts_local[1] +
                           map[pl.mapn + j + T.cmax]) * lay.npts_local[0];
      C.row4[r] = make_int4(i0, len, rowbase, (p << 16) | (j + 32768));
    }
  }
  __syncthreads();
  // exclusive scan of row lengths -> offsets (segment per thread)
  const int seg = (nrows + nthr - 1) / nthr;
  const int r0 = tid * seg, r1 = min(nrows, r0 + seg);
  int s = 0;
  for (int r = r0; r < r1; r++) s += C.roff[r];
  int total;
  int off = block_excl_scan<NW>(s, C.misc + 2, &total);
  for (int r = r0; r < r1; r++) {
    const int len = C.roff[r];
    C.roff[r] = off;
    off += len;
  }
  if (tid == 0) C.roff[nrows] = total;
  __syncthreads();
  return npl;
}

// Copies the block's task record and its level description into LDS.
__device__ __forceinline__ void stage_task(const PointParams &P, Task &sT, Level &sL, int tid,
                                           int nthr) {
  static_assert(sizeof(Task) % 4 == 0 && sizeof(Level) % 4 == 0, "word copies");
  const int *src = reinterpret_cast<const int *>(&P.tasks[launch_task(P, blockIdx.x)]);
  int *dst = reinterpret_cast<int *>(&sT);
  for (int e = tid; e < (int)(sizeof(Task) / 4); e += nthr) dst[e] = src[e];
  __syncthreads();
  const int *lsrc = reinterpret_cast<const int *>(&P.levels[sT.level]);
  int *ldst = reinterpret_cast<int *>(&sL);
  for (int e = tid; e < (int)(sizeof(Level) / 4); e += nthr) ldst[e] = lsrc[e];
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Collocate: one workgroup per task
// ---------------------------------------------------------------------------
// Lanes of a workgroup as slots of `width` lanes, one x column per lane:
// slot s handles rows s, s + nslots, ...; lane q of the slot handles the
// cube column lo[0] + pass*width + q.
struct XSlots {
  int width, slot, nslots, q, npass;
};

template <int NW>
__device__ __forceinline__ XSlots x_slots(const Task &T, int tid) {
  const int extent = 2 - 2 * T.lo[0];  // columns lo[0] .. 1 - lo[0]
  XSlots X;
  X.width = extent <= 16 ? 16 : (extent <= 32 ? 32 : 64);
  const int per_wave = 64 / X.width;
  X.slot = (tid >> 6) * per_wave + (tid & 63) / X.width;
  X.nslots = per_wave * NW;
  X.q = (tid & 63) % X.width;
  X.npass = (extent + X.width - 1) / X.width;
  return X;
}

// ORTHO: the launch covers only orthorhombic (true) or only general (false)
// tasks, so each instantiation carries one code path.
template <int NW, int LPB, bool ORTHO>
__device__ __forceinline__ void collocate_task(const PointParams &P) {
  extern __shared__ double lds[];
  __shared__ Task sT;
  __shared__ Level sL;
  const int tid = threadIdx.x, nthr = 64 * NW;
  stage_task(P, sT, sL, tid, nthr);
  const Task &T = sT;
  if (T.flags & kTaskSkip) {
    if (P.count_only && tid == 0)
      P.task_out[launch_task(P, blockIdx.x) * kTaskOutStride] = 0.0;
    return;
  }
  const LdsPlan &pl = P.plan;
  const Level &L = sL;
  const Layout &lay = L.lay;
  constexpr bool ortho = ORTHO;
  double *cab = lds + pl.o_cab;
  double *alpha = lds + pl.o_alpha;
  double *cxyz = lds + pl.o_cxyz;
  double *cijk = lds + pl.o_cijk;
  double *tab = lds + pl.o_tab;
  double *cxy = lds + pl.o_cxy;
  double *cxrow = lds + pl.o_cxrow;
  int *ib = reinterpret_cast<int *>(lds + pl.n_doubles);
  int *map = ib + pl.i_map;
  const Chunk C = chunk_arrays(lds, ib, pl);

  const int la = T.la_max + P.ld[0], lb = T.lb_max + P.ld[2];
  const int la_min = max(T.la_min + P.ld[1], 0), lb_min = max(T.lb_min + P.ld[3], 0);
  const int lp = la + lb;
  const int NP = pl.np, EXT = pl.ext;

  if (!P.count_only) {
    // ---- cab from the decontracted density block (prepare_pab) ----
    const int n1 = ncoset(la), n2 = ncoset(lb);
    for (int e = tid; e < n1 * n2; e += nthr) cab[e] = 0.0;
    for (int e = tid; e < NP * NP * NP; e += nthr) cxyz[e] = 0.0;
    __syncthreads();
    const Group &G = P.groups[T.group];
    const int a0 = ncoset(T.la_min - 1), na = ncoset(T.la_max) - a0;
    const int b0 = ncoset(T.lb_min - 1), nb = ncoset(T.lb_max) - b0;
    const CabRef cr{cab, n1};
    for (int e = tid; e < na * nb; e += nthr) {
      const int ia = a0 + e % na, jb = b0 + e / na;
      const long long q = G.cart_off + (long long)(T.o2 + jb) * G.ncoa + T.o1 + ia;
      const double p = (q < P.cart_len) ? P.cart[q] : 0.0;
      if (P.func == 100)
        cab[jb * n1 + ia] = p;
      else
        apply_func(P.func, orb(ia), orb(jb), T.zeta, T.zetb, p, cr);
    }
    build_alpha(T, la, lb, alpha, tid, nthr);
    __syncthreads();
    // ---- cab -> cxyz (pairs scatter into LDS) ----
    {
      const int np = lp + 1, nA = la + 1, nB = lb + 1;
      const double pref = T.rscale * T.pexp;
      const int a1 = ncoset(la_min - 1), ma = ncoset(la) - a1;
      const int b1 = ncoset(lb_min - 1), mb = ncoset(lb) - b1;
      for (int e = tid; e < ma * mb; e += nthr) {
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
              atomicAdd(&cxyz[(lz * NP + ly) * NP + lx], cv * (ax[lx] * ay[ly] * az[lz] * pref));
      }
    }
    // cab/alpha share LDS with the per-axis tables and chunk arrays
    __syncthreads();
    if (ortho) {
      build_ortho_tables(T, lay, lp, pl, tab, map, tid, nthr);
      __syncthreads();
    } else {
      build_general_maps(T, lay, pl, map, tid, nthr);
      __syncthreads();
      cxyz_cijk(true, lp, lay.dh, NP, cxyz, cijk, tid, nthr);
      __syncthreads();
    }
  } else {
    if (ortho) {
      build_ortho_tables(T, lay, 0, pl, tab, map, tid, nthr);
    } else {
      build_general_maps(T, lay, pl, map, tid, nthr);
    }
    __syncthreads();
  }

  const long long nx = lay.npts_local[0], ny = lay.npts_local[1];
  const int cmax = T.cmax;
  const int *mapx = map, *mapy = map + pl.mapn, *mapz = map + 2 * pl.mapn;
  double counted = 0.0;
  const int kbeg0 = T.lo[2], kend = ortho ? 1 - T.lo[2] : T.hi[2];
  for (int kbeg = kbeg0; kbeg <= kend;) {
    const int npl = build_chunk<NW>(T, lay, ortho, kbeg, pl, C, map, tid, nthr);
    const int nrows = C.misc[1];
    const int npts = C.roff[nrows];
    if (P.count_only) {
      counted += npts;
      kbeg += npl;
      __syncthreads();
      continue;
    }
    // ---- plane coefficients ----
    for (int e = tid; e < npl * NP * NP; e += nthr) {
      const int p = e / (NP * NP), ly = (e / NP) % NP, lx = e % NP;
      if (ly > lp || lx > lp - ly) continue;
      const int k = C.pk[p];
      double s = 0.0;
      if (ortho) {
        for (int lz = 0; lz <= lp - ly - lx; lz++)
          s += cxyz[(lz * NP + ly) * NP + lx] * tab[(2 * NP + lz) * EXT + k + cmax];
      } else {
        const double dk = k - T.gp[2];
        double pw = 1.0;
        for (int kl = 0; kl <= lp - ly - lx; kl++) {
          s += cijk[(kl * NP + ly) * NP + lx] * pw;
          pw *= dk;
        }
      }
      cxy[e] = s;
    }
    __syncthreads();
    // ---- row coefficients ----
    for (int e = tid; e < nrows * NP; e += nthr) {
      const int r = e / NP, lx = e % NP;
      if (lx > lp) continue;
      const int p = C.rp[r], j = C.rj[r];
      const double *cp = &cxy[p * NP * NP];
      double s = 0.0;
      if (ortho) {
        for (int ly = 0; ly <= lp - lx; ly++)
          s += cp[ly * NP + lx] * tab[(1 * NP + ly) * EXT + j + cmax];
      } else {
        const double dj = j - T.gp[1];
        double pw = 1.0;
        for (int jl = 0; jl <= lp - lx; jl++) {
          s += cp[jl * NP + lx] * pw;
          pw *= dj;
        }
      }
      cxrow[e] = s;
    }
    __syncthreads();
    if (ortho) {
      // ---- points: each lane owns one x column of the cube ----
      const XSlots X = x_slots<NW>(T, tid);
      for (int pass = 0; pass < X.npass; pass++) {
        const int i = T.lo[0] + pass * X.width + X.q;
        const bool iv = i <= 1 - T.lo[0];
        double px[LPB + 1];
#pragma unroll
        for (int l = 0; l <= LPB; l++)
          px[l] = (iv && l <= lp) ? tab[(0 * NP + l) * EXT + i + cmax] : 0.0;
        const long long ig = iv ? mapx[i + cmax] : 0;
        for (int r = X.slot; r < nrows; r += X.nslots) {
          const int4 m = C.row4[r];
          if (!iv || i < m.x || i >= m.x + m.y) continue;
          const double *cr = &cxrow[r * NP];
          double v = 0.0;
#pragma unroll
          for (int l = 0; l <= LPB; l++)
            if (l <= lp) v = fma(cr[l], px[l], v);
          const long long gidx = (long long)m.z + ig;
          if (gidx < L.npts) atomicAdd(&L.grid[gidx], v);
        }
      }
      __syncthreads();
      kbeg += npl;
      continue;
    }
    // ---- points: consecutive lanes take consecutive points of the chunk ----
    for (int base = 0; base < npts; base += nthr) {
      const int q = base + tid;
      if (q < npts) {
        const int r = find_row(C.roff, nrows, q);
        const int i = C.ri[r] + (q - C.roff[r]);
        const int j = C.rj[r];
        const int k = C.pk[C.rp[r]];
        const double *cr = &cxrow[r * NP];
        double v = 0.0;
        long long gidx;
        bool ok = true;
        if (ortho) {
          for (int lx = 0; lx <= lp; lx++)
            v = fma(cr[lx], tab[(0 * NP + lx) * EXT + i + cmax], v);
          gidx = ((long long)mapz[k + cmax] * ny + mapy[j + cmax]) * nx + mapx[i + cmax];
        } else {
          const int ig = mapx[i - T.lo[0]];
          ok = ig >= 0;
          const double di = i - T.gp[0], dj = j - T.gp[1], dk = k - T.gp[2];
          const double rx = di * lay.dh[0][0] + dj * lay.dh[1][0] + dk * lay.dh[2][0];
          const double ry = di * lay.dh[0][1] + dj * lay.dh[1][1] + dk * lay.dh[2][1];
          const double rz = di * lay.dh[0][2] + dj * lay.dh[1][2] + dk * lay.dh[2][2];
          const double arg = -T.zetp * (rx * rx + ry * ry + rz * rz);
          double dip = P.fp32 ? (double)__expf((float)arg) : exp(arg);
          for (int il = 0; il <= lp; il++) {
            v = fma(cr[il], dip, v);
            dip *= di;
          }
          gidx = ((long long)mapz[k - T.lo[2]] * ny + mapy[j - T.lo[1]]) * nx + ig;
        }
        if (ok && gidx >= 0 && gidx < L.npts) atomicAdd(&L.grid[gidx], v);
      }
    }
    __syncthreads();
    kbeg += npl;
  }
  if (P.count_only && tid == 0)
    P.task_out[launch_task(P, blockIdx.x) * kTaskOutStride] = counted;
}

template <int NW, int LPB, bool ORTHO>
__global__ __launch_bounds__(64 * NW) void collocate_kernel(PointParams P) {
  collocate_task<NW, LPB, ORTHO>(P);
}

// The same walk with P.count_only set: task points per task (its own kernel
// name keeps it apart from collocation in profiles).
template <int NW, bool ORTHO>
__global__ __launch_bounds__(64 * NW) void count_points_kernel(PointParams P) {
  collocate_task<NW, 4, ORTHO>(P);
}

// ---------------------------------------------------------------------------
// Integrate: one workgroup per task; LPB bounds lp for register accumulators
// ---------------------------------------------------------------------------


// FULL = false: hab only, no tau (forces/virial/tau code compiled out).
template <int NW, int LPB, bool FULL, bool ORTHO>
__global__ __launch_bounds__(64 * NW) void integrate_kernel(PointParams P) {
  extern __shared__ double lds[];
  __shared__ Task sT;
  __shared__ Level sL;
  const int tid = threadIdx.x, nthr = 64 * NW;
  stage_task(P, sT, sL, tid, nthr);
  const Task &T = sT;
  double *out = P.task_out + launch_task(P, blockIdx.x) * kTaskOutStride;
  if (T.flags & kTaskSkip) {
    if (P.do_forces && tid < kTaskOutStride) out[tid] = 0.0;
    return;
  }
  const LdsPlan &pl = P.plan;
  const Level &L = sL;
  const Layout &lay = L.lay;
  constexpr bool ortho = ORTHO;
  double *cab = lds + pl.o_cab;
  double *alpha = lds + pl.o_alpha;
  double *cxyz = lds + pl.o_cxyz;
  double *cijk = lds + pl.o_cijk;
  double *tab = lds + pl.o_tab;
  double *cxy = lds + pl.o_cxy;
  double *cxrow = lds + pl.o_cxrow;
  double *red = lds + pl.o_red;
  int *ib = reinterpret_cast<int *>(lds + pl.n_doubles);
  int *map = ib + pl.i_map;
  const Chunk C = chunk_arrays(lds, ib, pl);

  // l ranges of cab: process_get_ldiffs, grid_process_vab.h:222-251
  const int la = T.la_max + P.ld[0], lb = T.lb_max + P.ld[2];
  const int la_min = max(T.la_min + P.ld[1], 0), lb_min = max(T.lb_min + P.ld[3], 0);
  const int lp = la + lb;
  const int NP = pl.np, EXT = pl.ext;

  for (int e = tid; e < NP * NP * NP; e += nthr) {
    cxyz[e] = 0.0;
    if (!ortho) cijk[e] = 0.0;
  }
  if (ortho)
    build_ortho_tables(T, lay, lp, pl, tab, map, tid, nthr);
  else
    build_general_maps(T, lay, pl, map, tid, nthr);
  __syncthreads();

  const long long nx = lay.npts_local[0], ny = lay.npts_local[1];
  const int cmax = T.cmax;
  const int *mapx = map, *mapy = map + pl.mapn, *mapz = map + 2 * pl.mapn;
  const double *grid = L.grid;
  const int kbeg0 = T.lo[2], kend = ortho ? 1 - T.lo[2] : T.hi[2];
  for (int kbeg = kbeg0; kbeg <= kend;) {
    const int npl = build_chunk<NW>(T, lay, ortho, kbeg, pl, C, map, tid, nthr);
    const int nrows = C.misc[1];
    if (ortho && LPB <= 8) {
      // ---- each lane owns one x column: it accumulates the (y, z) moments
      //      Q[ly][lz] = sum_rows V(i,j,k) py_ly(j) pz_lz(k) of its column in
      //      registers; columns are then contracted with px into cxyz ----
      constexpr int NQ = (LPB + 1) * (LPB + 2) / 2;
      constexpr int SW = NQ + LPB + 1;  // staging words per column
      const XSlots X = x_slots<NW>(T, tid);
      double *stage = lds + pl.o_stage;
      for (int pass = 0; pass < X.npass; pass++) {
        for (int e = tid; e < X.width * SW; e += nthr) stage[e] = 0.0;
        __syncthreads();
        const int i = T.lo[0] + pass * X.width + X.q;
        const bool iv = i <= 1 - T.lo[0];
        double px[LPB + 1], R[LPB + 1], Q[NQ];
#pragma unroll
        for (int l = 0; l <= LPB; l++) {
          px[l] = (iv && l <= lp) ? tab[(0 * NP + l) * EXT + i + cmax] : 0.0;
          R[l] = 0.0;
        }
#pragma unroll
        for (int e = 0; e < NQ; e++) Q[e] = 0.0;
        const long long ig = iv ? mapx[i + cmax] : 0;
        int cur = -1;
        auto flush = [&](int plane) {
          const int k = C.pk[plane];
          int e = 0;
#pragma unroll
          for (int ly = 0; ly <= LPB; ly++) {
#pragma unroll
            for (int lz = 0; lz <= LPB - ly; lz++) {
              if (ly + lz <= lp) Q[e] = fma(R[ly], tab[(2 * NP + lz) * EXT + k + cmax], Q[e]);
              e++;
            }
          }
#pragma unroll
          for (int l = 0; l <= LPB; l++) R[l] = 0.0;
        };
        for (int r = X.slot; r < nrows; r += X.nslots) {
          const int4 m = C.row4[r];
          const int p = m.w >> 16;
          if (p != cur) {
            if (cur >= 0) flush(cur);
            cur = p;
          }
          if (!iv || i < m.x || i >= m.x + m.y) continue;
          const int j = (m.w & 0xffff) - 32768;
          const long long gidx = (long long)m.z + ig;
          const double g = (gidx < L.npts) ? grid[gidx] : 0.0;
#pragma unroll
          for (int ly = 0; ly <= LPB; ly++)
            if (ly <= lp) R[ly] = fma(g, tab[(1 * NP + ly) * EXT + j + cmax], R[ly]);
        }
        if (cur >= 0) flush(cur);
        if (iv) {
          double *sq = &stage[X.q * SW];
#pragma unroll
          for (int e = 0; e < NQ; e++)
            if (Q[e] != 0.0) atomicAdd(&sq[e], Q[e]);
          if (X.slot == 0) {
#pragma unroll
            for (int l = 0; l <= LPB; l++) sq[NQ + l] = px[l];
          }
        }
        __syncthreads();
        // cxyz[lz][ly][lx] += sum over columns of px[lx] * Q[ly][lz]
        for (int e = tid; e < NP * NP * NP; e += nthr) {
          const int lz = e / (NP * NP), ly = (e / NP) % NP, lx = e % NP;
          if (lz + ly + lx > lp) continue;
          const int qi = ly * (LPB + 1) - (ly * (ly - 1)) / 2 + lz;
          double s = 0.0;
          for (int c = 0; c < X.width; c++) s = fma(stage[c * SW + NQ + lx], stage[c * SW + qi], s);
          cxyz[e] += s;
        }
        __syncthreads();
      }
      kbeg += npl;
      continue;
    }
    // ---- rows: each lane integrates whole rows into registers ----
    for (int r = tid; r < nrows; r += nthr) {
      double acc[LPB + 1];
#pragma unroll
      for (int l = 0; l <= LPB; l++) acc[l] = 0.0;
      const int i0 = C.ri[r], len = C.roff[r + 1] - C.roff[r];
      const int j = C.rj[r], k = C.pk[C.rp[r]];
      if (ortho) {
        const long long rowbase = ((long long)mapz[k + cmax] * ny + mapy[j + cmax]) * nx;
        for (int q = 0; q < len; q++) {
          const int i = i0 + q;
          const long long gidx = rowbase + mapx[i + cmax];
          const double g = (gidx < L.npts) ? grid[gidx] : 0.0;
          const double *tx = &tab[i + cmax];
#pragma unroll
          for (int l = 0; l <= LPB; l++)
            if (l <= lp) acc[l] = fma(g, tx[l * EXT], acc[l]);
        }
      } else {
        const long long rowbase = ((long long)mapz[k - T.lo[2]] * ny + mapy[j - T.lo[1]]) * nx;
        const double dj = j - T.gp[1], dk = k - T.gp[2];
        for (int q = 0; q < len; q++) {
          const int i = i0 + q;
          const int ig = mapx[i - T.lo[0]];
          if (ig < 0) continue;
          const long long gidx = rowbase + ig;
          const double g = (gidx < L.npts) ? grid[gidx] : 0.0;
          const double di = i - T.gp[0];
          const double rx = di * lay.dh[0][0] + dj * lay.dh[1][0] + dk * lay.dh[2][0];
          const double ry = di * lay.dh[0][1] + dj * lay.dh[1][1] + dk * lay.dh[2][1];
          const double rz = di * lay.dh[0][2] + dj * lay.dh[1][2] + dk * lay.dh[2][2];
          const double arg = -T.zetp * (rx * rx + ry * ry + rz * rz);
          double dip = g * (P.fp32 ? (double)__expf((float)arg) : exp(arg));
#pragma unroll
          for (int l = 0; l <= LPB; l++)
            if (l <= lp) {
              acc[l] += dip;
              dip *= di;
            }
        }
      }
#pragma unroll
      for (int l = 0; l <= LPB; l++)
        if (l <= lp) cxrow[r * NP + l] = acc[l];
    }
    __syncthreads();
    // ---- rows -> planes ----
    for (int e = tid; e < npl * NP * NP; e += nthr) {
      const int p = e / (NP * NP), ly = (e / NP) % NP, lx = e % NP;
      if (ly > lp || lx > lp - ly) continue;
      double s = 0.0;
      for (int r = C.prow[p]; r < C.prow[p + 1]; r++) {
        if (ortho) {
          s += cxrow[r * NP + lx] * tab[(1 * NP + ly) * EXT + C.rj[r] + cmax];
        } else {
          s += cxrow[r * NP + lx] * ipow(C.rj[r] - T.gp[1], ly);
        }
      }
      cxy[e] = s;
    }
    __syncthreads();
    // ---- planes -> cxyz (ortho) / cijk (general) ----
    for (int e = tid; e < NP * NP * NP; e += nthr) {
      const int lz = e / (NP * NP), ly = (e / NP) % NP, lx = e % NP;
      if (lz + ly + lx > lp) continue;
      double s = 0.0;
      for (int p = 0; p < npl; p++) {
        const double w = ortho ? tab[(2 * NP + lz) * EXT + C.pk[p] + cmax]
                               : ipow(C.pk[p] - T.gp[2], lz);
        s += cxy[(p * NP + ly) * NP + lx] * w;
      }
      if (ortho)
        cxyz[e] += s;
      else
        cijk[e] += s;
    }
    __syncthreads();
    kbeg += npl;
  }
  if (!ortho) {
    cxyz_cijk(false, lp, lay.dh, NP, cxyz, cijk, tid, nthr);
  }
  // ---- cxyz -> cab (adjoint of the collocate transform, rscale = 1) ----
  const int n1 = ncoset(la), n2 = ncoset(lb);
  for (int e = tid; e < n1 * n2; e += nthr) cab[e] = 0.0;
  build_alpha(T, la, lb, alpha, tid, nthr);
  __syncthreads();
  {
    const int np = lp + 1, nA = la + 1, nB = lb + 1;
    const int a1 = ncoset(la_min - 1), ma = ncoset(la) - a1;
    const int b1 = ncoset(lb_min - 1), mb = ncoset(lb) - b1;
    for (int e = tid; e < ma * mb; e += nthr) {
      const int ia = a1 + e % ma, jb = b1 + e / ma;
      const Ang A = orb(ia), B = orb(jb);
      const double *ax = &alpha[((0 * nB + B.x) * nA + A.x) * np];
      const double *ay = &alpha[((1 * nB + B.y) * nA + A.y) * np];
      const double *az = &alpha[((2 * nB + B.z) * nA + A.z) * np];
      double s = 0.0;
      for (int lz = 0; lz <= A.z + B.z; lz++)
        for (int ly = 0; ly <= A.y + B.y; ly++)
          for (int lx = 0; lx <= A.x + B.x; lx++)
            s += cxyz[(lz * NP + ly) * NP + lx] * (ax[lx] * ay[ly] * az[lz] * T.pexp);
      cab[jb * n1 + ia] = s;
    }
  }
  __syncthreads();
  // ---- hab, forces, virial (grid_ref_integrate.c:132-171) ----
  const Group &G = P.groups[T.group];
  const CabGet cg{cab, n1};
  const bool tau = FULL && P.compute_tau != 0;
  const double za = T.zeta, zb = T.zetb;
  const int a0 = ncoset(T.la_min - 1), na = ncoset(T.la_max) - a0;
  const int b0 = ncoset(T.lb_min - 1), nb = ncoset(T.lb_max) - b0;
  // Single-task mode (P.pgf_out set, grid_ref_integrate.c:95-203): forces of
  // a and b, virials of a and b apart, and the unweighted derivatives hdab,
  // hadb, a_hdab of each hab element ([n2][n1][3] and [n2][n1][3][3] with
  // the element's Cartesian index q; tau is off for those).
  const bool pgf = P.pgf_out != nullptr;
  double f[6] = {0, 0, 0, 0, 0, 0};
  double va[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double vb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int e = tid; e < na * nb; e += nthr) {
    const int ia = a0 + e % na, jb = b0 + e / na;
    const Ang A = orb(ia), B = orb(jb);
    const long long q = G.cart_off + (long long)(T.o2 + jb) * G.ncoa + T.o1 + ia;
    if (q < P.cart_len) atomicAdd(&P.hab_cart[q], get_hab(A, B, za, zb, cg, tau));
    if (FULL && P.do_forces) {
      const double pv = (P.cart && q < P.cart_len) ? P.cart[q] : 0.0;
      for (int i = 0; i < 3; i++) {
        const double fa = get_fa(A, B, i, za, zb, cg, tau);
        const double fb = get_fb(A, B, i, za, zb, T.rab, cg, tau);
        f[i] += pv * fa;
        f[3 + i] += pv * fb;
        if (pgf && P.hdab && q < P.cart_len) P.hdab[3 * q + i] += fa;
        if (pgf && P.hadb && q < P.cart_len) P.hadb[3 * q + i] += fb;
      }
      if (P.do_virial) {
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) {
            const double a = get_va(A, B, i, j, za, zb, cg, tau);
            va[i * 3 + j] += pv * a;
            vb[i * 3 + j] += pv * get_vb(A, B, i, j, za, zb, T.rab, cg, tau);
            if (pgf && P.a_hdab && q < P.cart_len) P.a_hdab[9 * q + 3 * i + j] += a;
          }
      }
    }
  }
  if (FULL && P.do_forces) {
    const int lane = tid & 63, w = tid >> 6;
    // list mode: forces and the summed virial (15); single task: 6 + 9 + 9
    const int nout = pgf ? 24 : kTaskOutStride;
    for (int c = 0; c < nout; c++) {
      const double x = c < 6 ? f[c]
                     : pgf   ? (c < 15 ? va[c - 6] : vb[c - 15])
                             : va[c - 6] + vb[c - 6];
      const double sum = wave_sum(x);
      if (lane == 0) red[w * nout + c] = sum;
    }
    __syncthreads();
    if (tid < nout) {
      double s = 0.0;
      for (int q = 0; q < NW; q++) s += red[q * nout + tid];
      if (pgf) {
        P.pgf_out[tid] = s;
      } else {
        const double scale = (T.iatom == T.jatom) ? 1.0 : 2.0;
        out[tid] = scale * s;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Decontraction / contraction of matrix sub-blocks (grid_ref_task_list.c
// :237-271 and :462-499), one wave per (block, iset, jset) group (the
// groups are small: tens of Cartesian functions per side), four groups per
// workgroup.  Columns of the Cartesian block are processed in chunks of
// `cach`; each wave owns `wave_doubles` of LDS.
// ---------------------------------------------------------------------------
constexpr int kGroupWaves = 4;

__device__ __forceinline__ double block_elem(const Group &g, const double *blk, int sa,
                                             int sb) {
  return g.transpose ? blk[(long long)(g.sb0 + sb) * g.nsgfa + g.sa0 + sa]
                     : blk[(long long)(g.sa0 + sa) * g.nsgfb + g.sb0 + sb];
}

__global__ __launch_bounds__(64 * kGroupWaves) void decontract_kernel(
    const Group *__restrict__ groups, int ngroups, const double *__restrict__ pab,
    const double *__restrict__ sphi, double *__restrict__ cart, long long cart_len, int cach,
    int wave_doubles) {
  extern __shared__ double lds_dc[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gi = blockIdx.x * kGroupWaves + wave;
  if (gi >= ngroups) return;
  const Group &g = groups[gi];
  double *work = lds_dc + (size_t)wave * wave_doubles;  // [nsb][nc]
  const double *blk = pab + g.block_off;
  for (int c0 = 0; c0 < g.ncoa; c0 += cach) {
    const int nc = min(cach, g.ncoa - c0);
    for (int e = lane; e < g.nsb * nc; e += 64) {
      const int sb = e / nc, ca = c0 + e % nc;
      double s = 0.0;
      for (int sa = 0; sa < g.nsa; sa++)
        s = fma(block_elem(g, blk, sa, sb), sphi[g.sphi_a + (long long)sa * g.maxcoa + ca], s);
      work[e] = s;
    }
    wave_sync();
    for (int e = lane; e < g.ncob * nc; e += 64) {
      const int cb = e / nc, cc = e % nc;
      double s = 0.0;
      for (int sb = 0; sb < g.nsb; sb++)
        s = fma(sphi[g.sphi_b + (long long)sb * g.maxcob + cb], work[sb * nc + cc], s);
      const long long q = g.cart_off + (long long)cb * g.ncoa + c0 + cc;
      if (q < cart_len) cart[q] = s;
    }
    wave_sync();
  }
}

__global__ __launch_bounds__(64 * kGroupWaves) void contract_kernel(
    const Group *__restrict__ groups, int ngroups, const double *__restrict__ hab_cart,
    const double *__restrict__ sphi, double *__restrict__ hab, long long hab_len, int cach,
    int wave_doubles, int acc_doubles) {
  extern __shared__ double lds_ct[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gi = blockIdx.x * kGroupWaves + wave;
  if (gi >= ngroups) return;
  const Group &g = groups[gi];
  double *acc = lds_ct + (size_t)wave * wave_doubles;  // [nsb][nsa]
  double *work = acc + acc_doubles;                    // [nsb][nc]
  for (int e = lane; e < g.nsa * g.nsb; e += 64) acc[e] = 0.0;
  for (int c0 = 0; c0 < g.ncoa; c0 += cach) {
    const int nc = min(cach, g.ncoa - c0);
    for (int e = lane; e < g.nsb * nc; e += 64) {
      const int sb = e / nc, cc = e % nc;
      double s = 0.0;
      for (int cb = 0; cb < g.ncob; cb++)
        s = fma(sphi[g.sphi_b + (long long)sb * g.maxcob + cb],
                hab_cart[g.cart_off + (long long)cb * g.ncoa + c0 + cc], s);
      work[e] = s;
    }
    wave_sync();
    for (int e = lane; e < g.nsa * g.nsb; e += 64) {
      const int sb = e / g.nsa, sa = e % g.nsa;
      double s = 0.0;
      for (int cc = 0; cc < nc; cc++)
        s = fma(work[sb * nc + cc], sphi[g.sphi_a + (long long)sa * g.maxcoa + c0 + cc], s);
      acc[e] += s;
    }
    wave_sync();
  }
  for (int e = lane; e < g.nsa * g.nsb; e += 64) {
    const int sb = e / g.nsa, sa = e % g.nsa;
    const long long q = g.block_off + (g.transpose ? (long long)(g.sb0 + sb) * g.nsgfa + g.sa0 + sa
                                                   : (long long)(g.sa0 + sa) * g.nsgfb + g.sb0 + sb);
    if (q < hab_len) hab[q] = acc[e];
  }
}

// Staged (de)contraction: persistent waves, each taking groups
// g = wave, wave + total waves, ...; per group every operand is first copied
// into the wave's LDS with all loads in flight at once (the block's
// (sa, sb) sub-block P, the rows sb of sphi_b, a chunk of nc columns of the
// rows sa of sphi_a), then both products run from LDS:
//   decontract  W[sb][c] = sum_sa P[sa][sb] Sa[sa][c],
//               cart[cb][c] = sum_sb Sb[sb][cb] W[sb][c]
//   contract    W[sb][c] = sum_cb Sb[sb][cb] H[cb][c]   (H: the hab_cart chunk)
//               acc[sb][sa] += sum_c W[sb][c] Sa[sa][c], hab = acc
// LDS per wave (doubles): decontract nsa nsb + nsb ncob + nc (nsa + nsb);
// contract nsb ncob + nc (ncob + nsb + nsa) + nsa nsb.
// Operand staging of one group: element e of each matrix (lane + 64 u) is
// loaded into registers (all loads in flight together, while the previous
// group computes) and later written to the wave's LDS.  kStageRegs elements
// per lane and matrix; the host uses the kernels only when every matrix of
// the list fits.
// e / n for 0 <= e < 4096 and 1 <= n <= 256 (the staged kernels' operand
// indices: the host uses them only for matrices within these bounds)
// through a single-precision reciprocal instead of an integer division:
// (e + 1/2) / n lies at least 1/512 from an integer, and rounding moves
// (e + 1/2) rcp(n) by less than 4096 x 2^-22 < 1/1024.
__device__ __forceinline__ int fdiv(int e, int n) {
  return (int)(((float)e + 0.5f) * __builtin_amdgcn_rcpf((float)n));
}

constexpr int kStageRegs = 8;

struct StageRegs {
  double p[kStageRegs], b[kStageRegs], a[kStageRegs];
};

// decontract: P[sa][sb] from the block, Sb[sb][cb], Sa[sa][c] (first chunk)
__device__ __forceinline__ void dc_load(const Group &g, const double *pab, const double *sphi,
                                        int nc, int lane, StageRegs &R) {
  const double *blk = pab + g.block_off;
#pragma unroll
  for (int u = 0; u < kStageRegs; u++) {
    const int e = lane + 64 * u;
    if (e < g.nsa * g.nsb) {
      if (g.transpose) {
        const int sb = fdiv(e, g.nsa), sa = e - sb * g.nsa;
        R.p[u] = blk[(long long)(g.sb0 + sb) * g.nsgfa + g.sa0 + sa];
      } else {
        const int sa = fdiv(e, g.nsb), sb = e - sa * g.nsb;
        R.p[u] = blk[(long long)(g.sa0 + sa) * g.nsgfb + g.sb0 + sb];
      }
    }
    if (e < g.nsb * g.ncob) {
      const int sb = fdiv(e, g.ncob), cb = e - sb * g.ncob;
      R.b[u] = sphi[g.sphi_b + (long long)sb * g.maxcob + cb];
    }
    if (e < g.nsa * nc) {
      const int sa = fdiv(e, nc), c = e - sa * nc;
      R.a[u] = sphi[g.sphi_a + (long long)sa * g.maxcoa + c];
    }
  }
}

__device__ __forceinline__ void dc_store(const Group &g, int nc, int lane, const StageRegs &R,
                                         double *P, double *Sb, double *Sa) {
#pragma unroll
  for (int u = 0; u < kStageRegs; u++) {
    const int e = lane + 64 * u;
    if (e < g.nsa * g.nsb) {
      if (g.transpose) {
        const int sb = fdiv(e, g.nsa), sa = e - sb * g.nsa;
        P[sa * g.nsb + sb] = R.p[u];
      } else {
        P[e] = R.p[u];
      }
    }
    if (e < g.nsb * g.ncob) Sb[e] = R.b[u];
    if (e < g.nsa * nc) Sa[e] = R.a[u];
  }
}

// Staged (de)contraction: persistent waves, each taking groups
// g = wave, wave + total waves, ...; every operand is copied into the wave's
// LDS (the block's (sa, sb) sub-block P, the rows sb of sphi_b, chunks of nc
// columns of the rows sa of sphi_a; contract: the hab_cart chunk H), then
// both products run from LDS:
//   decontract  W[sb][c] = sum_sa P[sa][sb] Sa[sa][c],
//               cart[cb][c] = sum_sb Sb[sb][cb] W[sb][c]
//   contract    W[sb][c] = sum_cb Sb[sb][cb] H[cb][c]
//               acc[sb][sa] += sum_c W[sb][c] Sa[sa][c], hab = acc
// PF: the operands of the first chunk of the next group are loaded into
// registers before the current group computes (software pipeline).
// LDS per wave (doubles): decontract nsa nsb + nsb ncob + nc (nsa + nsb);
// contract nsb ncob + nc (ncob + nsb + nsa) + nsa nsb.
template <bool PF>
__global__ __launch_bounds__(256) void decontract_staged_kernel(
    const Group *__restrict__ groups, int ngroups, const double *__restrict__ pab,
    const double *__restrict__ sphi, double *__restrict__ cart, long long cart_len, int nc_max,
    int wave_doubles) {
  extern __shared__ double lds_ds[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  const int stride = gridDim.x * nw;
  double *base = lds_ds + (size_t)wave * wave_doubles;
  int gi = blockIdx.x * nw + wave;
  if (gi >= ngroups) return;
  Group g = groups[gi];
  StageRegs R;
  if constexpr (PF) dc_load(g, pab, sphi, min(nc_max, g.ncoa), lane, R);
  while (true) {
    const int nsa = g.nsa, nsb = g.nsb, ncoa = g.ncoa, ncob = g.ncob;
    double *P = base;                   // [nsa][nsb]
    double *Sb = P + nsa * nsb;         // [nsb][ncob]
    double *Sa = Sb + nsb * ncob;       // [nsa][nc]
    double *W = Sa + nsa * nc_max;      // [nsb][nc]
    if constexpr (PF) {
      dc_store(g, min(nc_max, ncoa), lane, R, P, Sb, Sa);
    } else {
      const double *blk = pab + g.block_off;
      for (int e = lane; e < nsa * nsb; e += 64) {
        if (g.transpose) {
          const int sb = fdiv(e, nsa), sa = e - sb * nsa;
          P[sa * nsb + sb] = blk[(long long)(g.sb0 + sb) * g.nsgfa + g.sa0 + sa];
        } else {
          const int sa = fdiv(e, nsb), sb = e - sa * nsb;
          P[e] = blk[(long long)(g.sa0 + sa) * g.nsgfb + g.sb0 + sb];
        }
      }
      for (int e = lane; e < nsb * ncob; e += 64) {
        const int sb = fdiv(e, ncob), cb = e - sb * ncob;
        Sb[e] = sphi[g.sphi_b + (long long)sb * g.maxcob + cb];
      }
    }
    const int gn = gi + stride;
    Group g2;
    if (gn < ngroups) {
      g2 = groups[gn];
      if constexpr (PF) dc_load(g2, pab, sphi, min(nc_max, g2.ncoa), lane, R);
    }
    for (int c0 = 0; c0 < ncoa; c0 += nc_max) {
      const int nc = min(nc_max, ncoa - c0);
      if (!PF || c0 > 0) {
        for (int e = lane; e < nsa * nc; e += 64) {
          const int sa = fdiv(e, nc), c = e - sa * nc;
          Sa[e] = sphi[g.sphi_a + (long long)sa * g.maxcoa + c0 + c];
        }
      }
      wave_sync();
      for (int e = lane; e < nsb * nc; e += 64) {
        const int sb = fdiv(e, nc), c = e - sb * nc;
        double acc = 0.0;
        for (int sa = 0; sa < nsa; sa++) acc = fma(P[sa * nsb + sb], Sa[sa * nc + c], acc);
        W[e] = acc;
      }
      wave_sync();
      for (int e = lane; e < ncob * nc; e += 64) {
        const int cb = fdiv(e, nc), c = e - cb * nc;
        double acc = 0.0;
        for (int sb = 0; sb < nsb; sb++) acc = fma(Sb[sb * ncob + cb], W[sb * nc + c], acc);
        const long long q = g.cart_off + (long long)cb * ncoa + c0 + c;
        if (q < cart_len) cart[q] = acc;
      }
      wave_sync();
    }
    if (gn >= ngroups) break;
    g = g2;
    gi = gn;
  }
}

// contract: H[cb][c] (first chunk), Sb[sb][cb], Sa[sa][c] (first chunk)
__device__ __forceinline__ void ct_load(const Group &g, const double *hab_cart,
                                        const double *sphi, int nc, int lane, StageRegs &R) {
#pragma unroll
  for (int u = 0; u < kStageRegs; u++) {
    const int e = lane + 64 * u;
    if (e < g.ncob * nc) {
      const int cb = fdiv(e, nc), c = e - cb * nc;
      R.p[u] = hab_cart[g.cart_off + (long long)cb * g.ncoa + c];
    }
    if (e < g.nsb * g.ncob) {
      const int sb = fdiv(e, g.ncob), cb = e - sb * g.ncob;
      R.b[u] = sphi[g.sphi_b + (long long)sb * g.maxcob + cb];
    }
    if (e < g.nsa * nc) {
      const int sa = fdiv(e, nc), c = e - sa * nc;
      R.a[u] = sphi[g.sphi_a + (long long)sa * g.maxcoa + c];
    }
  }
}

template <bool PF>
__global__ __launch_bounds__(256) void contract_staged_kernel(
    const Group *__restrict__ groups, int ngroups, const double *__restrict__ hab_cart,
    const double *__restrict__ sphi, double *__restrict__ hab, long long hab_len, int nc_max,
    int wave_doubles) {
  extern __shared__ double lds_cs[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  const int stride = gridDim.x * nw;
  double *base = lds_cs + (size_t)wave * wave_doubles;
  int gi = blockIdx.x * nw + wave;
  if (gi >= ngroups) return;
  Group g = groups[gi];
  StageRegs R;
  if constexpr (PF) ct_load(g, hab_cart, sphi, min(nc_max, g.ncoa), lane, R);
  while (true) {
    const int nsa = g.nsa, nsb = g.nsb, ncoa = g.ncoa, ncob = g.ncob;
    double *acc = base;                 // [nsb][nsa]
    double *Sb = acc + nsa * nsb;       // [nsb][ncob]
    double *H = Sb + nsb * ncob;        // [ncob][nc]
    double *W = H + ncob * nc_max;      // [nsb][nc]
    double *Sa = W + nsb * nc_max;      // [nsa][nc]
    for (int e = lane; e < nsa * nsb; e += 64) acc[e] = 0.0;
    if constexpr (PF) {
      const int nc = min(nc_max, ncoa);
#pragma unroll
      for (int u = 0; u < kStageRegs; u++) {
        const int e = lane + 64 * u;
        if (e < ncob * nc) H[e] = R.p[u];
        if (e < nsb * ncob) Sb[e] = R.b[u];
        if (e < nsa * nc) Sa[e] = R.a[u];
      }
    } else {
      for (int e = lane; e < nsb * ncob; e += 64) {
        const int sb = fdiv(e, ncob), cb = e - sb * ncob;
        Sb[e] = sphi[g.sphi_b + (long long)sb * g.maxcob + cb];
      }
    }
    const int gn = gi + stride;
    Group g2;
    if (gn < ngroups) {
      g2 = groups[gn];
      if constexpr (PF) ct_load(g2, hab_cart, sphi, min(nc_max, g2.ncoa), lane, R);
    }
    for (int c0 = 0; c0 < ncoa; c0 += nc_max) {
      const int nc = min(nc_max, ncoa - c0);
      if (!PF || c0 > 0) {
        for (int e = lane; e < ncob * nc; e += 64) {
          const int cb = fdiv(e, nc), c = e - cb * nc;
          H[e] = hab_cart[g.cart_off + (long long)cb * ncoa + c0 + c];
        }
        for (int e = lane; e < nsa * nc; e += 64) {
          const int sa = fdiv(e, nc), c = e - sa * nc;
          Sa[e] = sphi[g.sphi_a + (long long)sa * g.maxcoa + c0 + c];
        }
      }
      wave_sync();
      for (int e = lane; e < nsb * nc; e += 64) {
        const int sb = fdiv(e, nc), c = e - sb * nc;
        double s = 0.0;
        for (int cb = 0; cb < ncob; cb++) s = fma(Sb[sb * ncob + cb], H[cb * nc + c], s);
        W[e] = s;
      }
      wave_sync();
      for (int e = lane; e < nsa * nsb; e += 64) {
        const int sb = fdiv(e, nsa), sa = e - sb * nsa;
        double s = 0.0;
        for (int c = 0; c < nc; c++) s = fma(W[sb * nc + c], Sa[sa * nc + c], s);
        acc[e] += s;
      }
      wave_sync();
    }
    for (int e = lane; e < nsa * nsb; e += 64) {
      const int sb = fdiv(e, nsa), sa = e - sb * nsa;
      const long long q = g.block_off + (g.transpose ? (long long)(g.sb0 + sb) * g.nsgfa + g.sa0 + sa
                                                     : (long long)(g.sa0 + sa) * g.nsgfb + g.sb0 + sb);
      if (q < hab_len) hab[q] = acc[e];
    }
    wave_sync();
    if (gn >= ngroups) break;
    g = g2;
    gi = gn;
  }
}

// FP32-mixed integrate: the potential rounded to FP32 once per call.
__global__ void to_float_kernel(const double *src, float *dst, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

__global__ void fill_kernel(double *p, long long n, double v) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = v;
}

// forces[atom] = sum of the per-task a- or b-forces of the tasks listed for
// the atom (CSR, entry = 2*task + side), one wave per atom.
__global__ __launch_bounds__(64) void reduce_forces_kernel(const int *ptr, const int *ent,
                                                           const double *task_out,
                                                           double *forces, int natoms) {
  const int atom = blockIdx.x, lane = threadIdx.x;
  if (atom >= natoms) return;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  for (int e = ptr[atom] + lane; e < ptr[atom + 1]; e += 64) {
    const long long t = ent[e] >> 1;
    const int side = ent[e] & 1;
    const double *o = &task_out[t * kTaskOutStride + 3 * side];
    f0 += o[0];
    f1 += o[1];
    f2 += o[2];
  }
  f0 = wave_sum(f0);
  f1 = wave_sum(f1);
  f2 = wave_sum(f2);
  if (lane == 0) {
    forces[3 * atom + 0] = f0;
    forces[3 * atom + 1] = f1;
    forces[3 * atom + 2] = f2;
  }
}

__global__ __launch_bounds__(256) void reduce_virial_kernel(const double *task_out,
                                                            long long ntasks,
                                                            double *virial) {
  __shared__ double red[4][9];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  double v[9];
  for (int c = 0; c < 9; c++) v[c] = 0.0;
  for (long long t = tid; t < ntasks; t += 256)
    for (int c = 0; c < 9; c++) v[c] += task_out[t * kTaskOutStride + 6 + c];
  for (int c = 0; c < 9; c++) {
    const double x = wave_sum(v[c]);
    if (lane == 0) red[w][c] = x;
  }
  __syncthreads();
  if (tid < 9) virial[tid] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
}

// ---------------------------------------------------------------------------
// Host-side launchers (called from grid_api.cpp)
// ---------------------------------------------------------------------------

// Dynamic LDS above 64 KiB has to be requested per kernel.
template <typename K>
static hipError_t allow_lds(K kernel, size_t lds) {
  if (lds <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

template <typename K>
static hipError_t launch_task_kernel(K kernel, const PointParams &P, int nthreads, size_t lds,
                                     hipStream_t s) {
  const hipError_t e = allow_lds(kernel, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kernel, dim3(P.ntasks), dim3(nthreads), lds, s, P);
  return hipGetLastError();
}

template <bool ORTHO>
static hipError_t launch_collocate_kind(const PointParams &P, int lp, size_t lds, hipStream_t s) {
  if (P.count_only) return launch_task_kernel(count_points_kernel<2, ORTHO>, P, 128, lds, s);
  switch (lpb_bucket(lp)) {
    case 4: return launch_task_kernel(collocate_kernel<2, 4, ORTHO>, P, 128, lds, s);
    case 8: return launch_task_kernel(collocate_kernel<2, 8, ORTHO>, P, 128, lds, s);
    case 12: return launch_task_kernel(collocate_kernel<2, 12, ORTHO>, P, 128, lds, s);
    default: return launch_task_kernel(collocate_kernel<2, 16, ORTHO>, P, 128, lds, s);
  }
}

// P.task_base / P.ntasks select the range of the task array; ortho says
// which kind of tasks the range holds.
hipError_t launch_collocate(const PointParams &P, bool ortho, hipStream_t s) {
  const size_t lds = P.plan.bytes;
  const int lp = P.plan.np - 1;
  if (lp > 16) return hipErrorInvalidValue;
  if (P.ntasks <= 0) return hipSuccess;
  if (!P.count_only) count_launch(11, false);  // FP64 points (FP32 exp only, general cells)
  return ortho ? launch_collocate_kind<true>(P, lp, lds, s)
               : launch_collocate_kind<false>(P, lp, lds, s);
}

template <bool FULL, bool ORTHO>
static hipError_t launch_integrate_kind(const PointParams &P, int lp, size_t lds, hipStream_t s) {
  switch (lpb_bucket(lp)) {
    case 4: return launch_task_kernel(integrate_kernel<2, 4, FULL, ORTHO>, P, 128, lds, s);
    case 8: return launch_task_kernel(integrate_kernel<2, 8, FULL, ORTHO>, P, 128, lds, s);
    case 12: return launch_task_kernel(integrate_kernel<2, 12, FULL, ORTHO>, P, 128, lds, s);
    default: return launch_task_kernel(integrate_kernel<2, 16, FULL, ORTHO>, P, 128, lds, s);
  }
}

hipError_t launch_integrate(const PointParams &P, bool ortho, hipStream_t s) {
  const size_t lds = P.plan.bytes;
  const int lp = P.plan.np - 1;
  if (lp > 16) return hipErrorInvalidValue;
  if (P.ntasks <= 0) return hipSuccess;
  count_launch(11, false);  // FP64 points (FP32 exp only, general cells)
  const bool full = P.do_forces || P.compute_tau;
  if (ortho)
    return full ? launch_integrate_kind<true, true>(P, lp, lds, s)
                : launch_integrate_kind<false, true>(P, lp, lds, s);
  return full ? launch_integrate_kind<true, false>(P, lp, lds, s)
              : launch_integrate_kind<false, false>(P, lp, lds, s);
}

// Staged kernels: nwaves waves per workgroup of wave_doubles each, at most
// `blocks` workgroups (persistent: the waves loop over the groups).
// pf: the list's matrices fit kStageRegs elements per lane (prefetching
// kernels).
hipError_t launch_decontract_staged(const Group *groups, int ngroups, const double *pab,
                                    const double *sphi, double *cart, long long cart_len, int nc,
                                    int wave_doubles, int nwaves, int blocks, bool pf,
                                    hipStream_t s) {
  if (ngroups == 0) return hipSuccess;
  const size_t lds = sizeof(double) * (size_t)wave_doubles * nwaves;
  blocks = std::max(1, std::min(blocks, (ngroups + nwaves - 1) / nwaves));
  auto go = [&](auto kernel) {
    const hipError_t e = allow_lds(kernel, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(64 * nwaves), lds, s, groups, ngroups, pab,
                       sphi, cart, cart_len, nc, wave_doubles);
    return hipGetLastError();
  };
  return pf ? go(decontract_staged_kernel<true>) : go(decontract_staged_kernel<false>);
}

hipError_t launch_contract_staged(const Group *groups, int ngroups, const double *hab_cart,
                                  const double *sphi, double *hab, long long hab_len, int nc,
                                  int wave_doubles, int nwaves, int blocks, bool pf,
                                  hipStream_t s) {
  if (ngroups == 0) return hipSuccess;
  const size_t lds = sizeof(double) * (size_t)wave_doubles * nwaves;
  blocks = std::max(1, std::min(blocks, (ngroups + nwaves - 1) / nwaves));
  auto go = [&](auto kernel) {
    const hipError_t e = allow_lds(kernel, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(64 * nwaves), lds, s, groups, ngroups,
                       hab_cart, sphi, hab, hab_len, nc, wave_doubles);
    return hipGetLastError();
  };
  return pf ? go(contract_staged_kernel<true>) : go(contract_staged_kernel<false>);
}

int stage_regs() { return kStageRegs; }

hipError_t launch_decontract(const Group *groups, int ngroups, const double *pab,
                             long long pab_len, const double *sphi, double *cart,
                             long long cart_len, int cach, int max_nsb, hipStream_t s) {
  if (ngroups == 0) return hipSuccess;
  (void)pab_len;
  const int wave_doubles = (cach * max_nsb + 1) & ~1;  // 16-byte aligned wave regions
  const size_t lds = sizeof(double) * (size_t)wave_doubles * kGroupWaves;
  const int blocks = (ngroups + kGroupWaves - 1) / kGroupWaves;
  const hipError_t e = allow_lds(decontract_kernel, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(decontract_kernel, dim3(blocks), dim3(64 * kGroupWaves), lds, s, groups,
                     ngroups, pab, sphi, cart, cart_len, cach, wave_doubles);
  return hipGetLastError();
}

hipError_t launch_contract(const Group *groups, int ngroups, const double *hab_cart,
                           const double *sphi, double *hab, long long hab_len, int cach,
                           int max_nsb, int max_nsa_nsb, hipStream_t s) {
  if (ngroups == 0) return hipSuccess;
  const int acc_doubles = (max_nsa_nsb + 1) & ~1;
  const int wave_doubles = acc_doubles + ((cach * max_nsb + 1) & ~1);
  const size_t lds = sizeof(double) * (size_t)wave_doubles * kGroupWaves;
  const int blocks = (ngroups + kGroupWaves - 1) / kGroupWaves;
  const hipError_t e = allow_lds(contract_kernel, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(contract_kernel, dim3(blocks), dim3(64 * kGroupWaves), lds, s, groups,
                     ngroups, hab_cart, sphi, hab, hab_len, cach, wave_doubles, acc_doubles);
  return hipGetLastError();
}

hipError_t launch_to_float(const double *src, float *dst, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(to_float_kernel, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

hipError_t launch_fill(double *p, long long n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_reduce_forces(const int *ptr, const int *ent, const double *task_out,
                                double *forces, int natoms, hipStream_t s) {
  if (natoms <= 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_forces_kernel, dim3(natoms), dim3(64), 0, s, ptr, ent, task_out,
                     forces, natoms);
  return hipGetLastError();
}

// Virial per list of a batched task list: one workgroup per list.
__global__ __launch_bounds__(64) void reduce_virial_lists_kernel(const int *ptr, const int *ent,
                                                                 const double *task_out,
                                                                 double *virial, int nlists) {
  const int l = blockIdx.x, lane = threadIdx.x;
  if (l >= nlists) return;
  double v[9];
  for (int c = 0; c < 9; c++) v[c] = 0.0;
  for (int e = ptr[l] + lane; e < ptr[l + 1]; e += 64)
    for (int c = 0; c < 9; c++) v[c] += task_out[(long long)ent[e] * kTaskOutStride + 6 + c];
  for (int c = 0; c < 9; c++) {
    const double x = wave_sum(v[c]);
    if (lane == 0) virial[9 * l + c] = x;
  }
}

hipError_t launch_reduce_virial_lists(const int *ptr, const int *ent, const double *task_out,
                                      double *virial, int nlists, hipStream_t s) {
  if (nlists <= 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_virial_lists_kernel, dim3(nlists), dim3(64), 0, s, ptr, ent, task_out,
                     virial, nlists);
  return hipGetLastError();
}

hipError_t launch_reduce_virial(const double *task_out, long long ntasks, double *virial,
                                hipStream_t s) {
  hipLaunchKernelGGL(reduce_virial_kernel, dim3(1), dim3(256), 0, s, task_out, ntasks, virial);
  return hipGetLastError();
}

}  // namespace gm
