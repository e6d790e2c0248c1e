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

#include "grid_internal.h"

namespace gm {

// ---------------------------------------------------------------------------
// Cartesian orbital tables (ordering of src/grid/common/grid_common.h:95-105)
// ---------------------------------------------------------------------------
struct OrbTable {
  signed char v[kNumOrb][3];
};

constexpr OrbTable make_orb_table() {
  OrbTable t{};
  int c = 0;
  for (int l = 0; l <= kMaxL; l++)
    for (int lx = l; lx >= 0; lx--)
      for (int lz = 0; lz <= l - lx; lz++) {
        t.v[c][0] = (signed char)lx;
        t.v[c][1] = (signed char)(l - lx - lz);
        t.v[c][2] = (signed char)lz;
        c++;
      }
  return t;
}

__constant__ OrbTable c_orb = make_orb_table();

__device__ __forceinline__ int ncoset(int l) {
  return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6;
}

__device__ __forceinline__ int coset(int lx, int ly, int lz) {
  const int l = lx + ly + lz;
  if (l == 0) return 0;
  return ncoset(l - 1) + ((l - lx) * (l - lx + 1)) / 2 + lz;
}

__device__ __forceinline__ int pmod(int a, int m) { return ((a % m) + m) % m; }

struct Ang {
  int x, y, z;
};

__device__ __forceinline__ Ang orb(int idx) {
  return Ang{c_orb.v[idx][0], c_orb.v[idx][1], c_orb.v[idx][2]};
}
__device__ __forceinline__ int comp(const Ang &a, int i) {
  return i == 0 ? a.x : (i == 1 ? a.y : a.z);
}
__device__ __forceinline__ Ang up(Ang a, int i) {
  a.x += (i == 0);
  a.y += (i == 1);
  a.z += (i == 2);
  return a;
}
// Lowering clamps at zero like down() in grid_common.h:143-147.
__device__ __forceinline__ Ang dn(Ang a, int i) {
  if (i == 0) a.x = a.x > 0 ? a.x - 1 : 0;
  if (i == 1) a.y = a.y > 0 ? a.y - 1 : 0;
  if (i == 2) a.z = a.z > 0 ? a.z - 1 : 0;
  return a;
}
__device__ __forceinline__ int aidx(const Ang &a) { return coset(a.x, a.y, a.z); }

// ---------------------------------------------------------------------------
// Small block-level helpers
// ---------------------------------------------------------------------------
template <int NW>
__device__ __forceinline__ int block_excl_scan(int x, int *scratch, int *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int incl = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) scratch[w] = incl;
  __syncthreads();
  int pre = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < NW; q++) {
    const int v = scratch[q];
    pre += (q < w) ? v : 0;
    tot += v;
  }
  *total = tot;
  __syncthreads();
  return pre + incl - x;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---------------------------------------------------------------------------
// prepare_pab stencils, src/grid/common/grid_prepare_pab.h:33-430
// (accumulated into LDS; different density elements can hit the same cab
// element, hence LDS atomics for every func but AB)
// ---------------------------------------------------------------------------
struct CabRef {
  double *m;
  int n1;
  __device__ __forceinline__ void add(Ang a, Ang b, double v) const {
    atomicAdd(&m[aidx(b) * n1 + aidx(a)], v);
  }
};

__device__ void mixed_b(int d, double fa, Ang a, Ang b, double zb, const CabRef &c) {
  const int i1 = d, i2 = (d + 1) % 3;
  const int b1 = comp(b, i1), b2 = comp(b, i2);
  c.add(a, dn(dn(b, i2), i1), b1 * b2 * fa);
  c.add(a, up(dn(b, i2), i1), -2.0 * zb * b2 * fa);
  c.add(a, dn(up(b, i2), i1), -2.0 * zb * b1 * fa);
  c.add(a, up(up(b, i2), i1), +4.0 * zb * zb * fa);
}

__device__ void pure_b(int d, double fa, Ang a, Ang b, double zb, const CabRef &c) {
  const int bd = comp(b, d);
  c.add(a, dn(dn(b, d), d), bd * (bd - 1) * fa);
  c.add(a, b, -2.0 * zb * (2 * bd + 1) * fa);
  c.add(a, up(up(b, d), d), +4.0 * zb * zb * fa);
}

__device__ void apply_func(int func, Ang a, Ang b, double za, double zb, double p,
                           const CabRef &c) {
  if (func == 200) {  // DADB
    for (int i = 0; i < 3; i++) {
      const int ai = comp(a, i), bi = comp(b, i);
      c.add(dn(a, i), dn(b, i), 0.5 * ai * bi * p);
      c.add(dn(a, i), up(b, i), -1.0 * ai * zb * p);
      c.add(up(a, i), dn(b, i), -1.0 * za * bi * p);
      c.add(up(a, i), up(b, i), 2.0 * za * zb * p);
    }
  } else if (func >= 301 && func <= 303) {  // ADBmDAB
    const int d = func - 301;
    c.add(a, dn(b, d), +comp(b, d) * p);
    c.add(a, up(b, d), -2.0 * zb * p);
    c.add(dn(a, d), b, -comp(a, d) * p);
    c.add(up(a, d), b, +2.0 * za * p);
  } else if (func >= 411 && func <= 433) {  // ARDBmDARB
    const int d = (func / 10) % 10 - 1, r = func % 10 - 1;
    const Ang br = up(b, r);
    c.add(a, dn(br, d), comp(b, d) * p);
    c.add(a, up(br, d), -2.0 * zb * p);
    c.add(dn(a, d), br, -comp(a, d) * p);
    c.add(up(a, d), br, +2.0 * za * p);
  } else if (func >= 501 && func <= 503) {  // DABpADB
    const int d = func - 501;
    c.add(a, dn(b, d), comp(b, d) * p);
    c.add(a, up(b, d), -2.0 * zb * p);
    c.add(dn(a, d), b, comp(a, d) * p);
    c.add(up(a, d), b, -2.0 * za * p);
  } else if (func >= 601 && func <= 603) {  // DX, DY, DZ
    const int d = func - 601;
    const int ad = comp(a, d), bd = comp(b, d);
    c.add(dn(a, d), dn(b, d), ad * bd * p);
    c.add(dn(a, d), up(b, d), -2.0 * ad * zb * p);
    c.add(up(a, d), dn(b, d), -2.0 * za * bd * p);
    c.add(up(a, d), up(b, d), +4.0 * za * zb * p);
  } else if (func >= 701 && func <= 703) {  // DXDY, DYDZ, DZDX
    const int d1 = func - 701, d2 = (d1 + 1) % 3;
    const int a1 = comp(a, d1), a2 = comp(a, d2);
    mixed_b(d1, a1 * a2 * p, dn(dn(a, d2), d1), b, zb, c);
    mixed_b(d1, -2.0 * za * a2 * p, up(dn(a, d2), d1), b, zb, c);
    mixed_b(d1, -2.0 * za * a1 * p, dn(up(a, d2), d1), b, zb, c);
    mixed_b(d1, 4.0 * za * za * p, up(up(a, d2), d1), b, zb, c);
  } else if (func >= 801 && func <= 803) {  // DXDX, DYDY, DZDZ
    const int d = func - 801;
    const int ad = comp(a, d);
    pure_b(d, ad * (ad - 1) * p, dn(dn(a, d), d), b, zb, c);
    pure_b(d, -2.0 * za * (2 * ad + 1) * p, a, b, zb, c);
    pure_b(d, 4.0 * za * za * p, up(up(a, d), d), b, zb, c);
  } else if (func >= 901 && func <= 903) {  // DAB
    const int d = func - 901;
    c.add(dn(a, d), b, -comp(a, d) * p);
    c.add(up(a, d), b, +2.0 * za * p);
  } else if (func >= 904 && func <= 906) {  // ADB
    const int d = func - 904;
    c.add(a, dn(b, d), -comp(b, d) * p);
    c.add(a, up(b, d), +2.0 * zb * p);
  } else if (func >= 1001 && func <= 1003) {  // CORE
    const int d = func - 1001;
    c.add(up(a, d), b, +2.0 * za * p);
  }
}

// ---------------------------------------------------------------------------
// Per-task setup shared by collocate and integrate
// ---------------------------------------------------------------------------

// alpha[d][ib][ia][l]: (x-a)^ia (x-b)^ib = sum_l alpha (x-p)^l
__device__ void build_alpha(const Task &T, int la, int lb, double *alpha, int tid,
                            int nthr) {
  const int np = la + lb + 1, na = la + 1, nb = lb + 1;
  for (int e = tid; e < 3 * na * nb; e += nthr) {
    const int d = e / (na * nb), rem = e % (na * nb);
    const int ib = rem / na, ia = rem % na;
    double *row = &alpha[e * np];
    for (int l = 0; l < np; l++) row[l] = 0.0;
    const double pa = T.rp[d] - T.ra[d];
    const double pb = T.rp[d] - T.rb[d];
    double bin_a = 1.0, pow_a = 1.0;
    for (int k = 0; k <= ia; k++) {
      double bin_b = 1.0, pow_b = 1.0;
      for (int l = 0; l <= ib; l++) {
        row[ia - k + ib - l] += bin_a * bin_b * pow_a * pow_b;
        bin_b *= ((double)(ib - l)) / ((double)(l + 1));
        pow_b *= pb;
      }
      bin_a *= ((double)(ia - k)) / ((double)(k + 1));
      pow_a *= pa;
    }
  }
}

// Orthorhombic per-axis tables pol[d][l][c] = (c h - roff)^l exp(-zetp (c h - roff)^2)
// and periodic index maps map[d][c] = modulo(center + c - shift, npts_global).
__device__ void build_ortho_tables(const Task &T, const Layout &lay, int lp,
                                   const LdsPlan &pl, double *tab, int *map, int tid,
                                   int nthr) {
  const int NP = pl.np, EXT = pl.ext, cmax = T.cmax;
  for (int d = 0; d < 3; d++) {
    const int n = 2 - 2 * T.lo[d];
    const double h = lay.dh[d][d];
    for (int e = tid; e < n; e += nthr) {
      const int c = T.lo[d] + e;
      const double t = c * h - T.roff[d];
      double v = exp(-T.zetp * t * t);
      for (int l = 0; l <= lp; l++) {
        tab[(d * NP + l) * EXT + c + cmax] = v;
        v *= t;
      }
    }
    for (int e = tid; e < 2 * cmax + 1; e += nthr)
      map[d * pl.mapn + e] = pmod(T.center[d] + (e - cmax) - lay.shift_local[d],
                                  lay.npts_global[d]);
  }
}

// General-case index maps: -1 marks indices outside the (border-masked) local
// grid, grid_ref_collint.h:514-529.
__device__ void build_general_maps(const Task &T, const Layout &lay,
                                   const LdsPlan &pl, int *map, int tid, int nthr) {
  for (int d = 0; d < 3; d++) {
    const int n = T.hi[d] - T.lo[d] + 1;
    for (int e = tid; e < n; e += nthr) {
      const int g = pmod(T.lo[d] + e - lay.shift_local[d], lay.npts_global[d]);
      map[d * pl.mapn + e] = (T.blo[d] <= g && g <= T.bhi[d]) ? g : -1;
    }
  }
}

// Multinomial weight of (x-xp)^lx (y-yp)^ly (z-zp)^lz -> lattice offsets,
// grid_ref_collint.h:706-761.  fac table up to kMaxL.
__device__ __forceinline__ double dfac(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; i++) f *= i;
  return f;
}

__device__ __forceinline__ double ipow(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; i++) r *= x;
  return r;
}

// forward: cijk[kl][jl][il] = sum M cxyz ; backward: cxyz += M^T cijk.
// One output element per thread, looping over its decompositions.
__device__ void cxyz_cijk(bool forward, int lp, const double dh[3][3], int NP,
                          double *cxyz, double *cijk, int tid, int nthr) {
  const int np = lp + 1;
  for (int e = tid; e < np * np * np; e += nthr) {
    const int o2 = e / (np * np), o1 = (e / np) % np, o0 = e % np;
    if (o0 + o1 + o2 > lp) continue;
    double s = 0.0;
    if (forward) {
      // output (kl, jl, il) = (o2, o1, o0): split il = ilx+ily+ilz etc.
      const int il = o0, jl = o1, kl = o2;
      for (int ilx = 0; ilx <= il; ilx++)
        for (int ily = 0; ily <= il - ilx; ily++) {
          const int ilz = il - ilx - ily;
          for (int jlx = 0; jlx <= jl; jlx++)
            for (int jly = 0; jly <= jl - jlx; jly++) {
              const int jlz = jl - jlx - jly;
              for (int klx = 0; klx <= kl; klx++)
                for (int kly = 0; kly <= kl - klx; kly++) {
                  const int klz = kl - klx - kly;
                  const int lx = ilx + jlx + klx, ly = ily + jly + kly,
                            lz = ilz + jlz + klz;
                  const double m =
                      ipow(dh[0][0], ilx) * ipow(dh[1][0], jlx) * ipow(dh[2][0], klx) *
                      ipow(dh[0][1], ily) * ipow(dh[1][1], jly) * ipow(dh[2][1], kly) *
                      ipow(dh[0][2], ilz) * ipow(dh[1][2], jlz) * ipow(dh[2][2], klz) *
                      dfac(lx) * dfac(ly) * dfac(lz) /
                      (dfac(ilx) * dfac(ily) * dfac(ilz) * dfac(jlx) * dfac(jly) *
                       dfac(jlz) * dfac(klx) * dfac(kly) * dfac(klz));
                  s += cxyz[(lz * NP + ly) * NP + lx] * m;
                }
            }
        }
      cijk[(kl * NP + jl) * NP + il] = s;
    } else {
      // output (lz, ly, lx) = (o2, o1, o0): split lx = ilx+jlx+klx etc.
      const int lx = o0, ly = o1, lz = o2;
      for (int ilx = 0; ilx <= lx; ilx++)
        for (int jlx = 0; jlx <= lx - ilx; jlx++) {
          const int klx = lx - ilx - jlx;
          for (int ily = 0; ily <= ly; ily++)
            for (int jly = 0; jly <= ly - ily; jly++) {
              const int kly = ly - ily - jly;
              for (int ilz = 0; ilz <= lz; ilz++)
                for (int jlz = 0; jlz <= lz - ilz; jlz++) {
                  const int klz = lz - ilz - jlz;
                  const int il = ilx + ily + ilz, jl = jlx + jly + jlz,
                            kl = klx + kly + klz;
                  const double m =
                      ipow(dh[0][0], ilx) * ipow(dh[1][0], jlx) * ipow(dh[2][0], klx) *
                      ipow(dh[0][1], ily) * ipow(dh[1][1], jly) * ipow(dh[2][1], kly) *
                      ipow(dh[0][2], ilz) * ipow(dh[1][2], jlz) * ipow(dh[2][2], klz) *
                      dfac(lx) * dfac(ly) * dfac(lz) /
                      (dfac(ilx) * dfac(ily) * dfac(ilz) * dfac(jlx) * dfac(jly) *
                       dfac(jlz) * dfac(klx) * dfac(kly) * dfac(klz));
                  s += cijk[(kl * NP + jl) * NP + il] * m;
                }
            }
        }
      cxyz[(lz * NP + ly) * NP + lx] += s;
    }
  }
}

// Orthorhombic row start, grid_ref_collint.h:147 / :49.
__device__ __forceinline__ int ortho_start(double remain, double hinv) {
  return (int)ceil(-1e-8 - sqrt(fmax(0.0, remain)) * hinv);
}

// General row i-range, grid_ref_collint.h:450-464.  Returns false if the row
// misses the sphere.
__device__ __forceinline__ bool general_row(const double dh[3][3], const double gp[3],
                                            int j, int k, double radius, int *ismin,
                                            int *ismax) {
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = 0; i < 3; i++) {
    const double v = (0 - gp[0]) * dh[0][i] + (j - gp[1]) * dh[1][i] + (k - gp[2]) * dh[2][i];
    a += dh[0][i] * dh[0][i];
    b += 2.0 * v * dh[0][i];
    c += v * v;
  }
  const double d = b * b - 4.0 * a * (c - radius * radius);
  if (!(0.0 < d)) return false;
  const double sd = sqrt(d);
  const double inv_2a = 1.0 / (2.0 * a);
  *ismin = (int)ceil((-b - sd) * inv_2a);
  *ismax = (int)floor((-b + sd) * inv_2a);
  return true;
}

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
  const int *src = reinterpret_cast<const int *>(&P.tasks[P.task_base + blockIdx.x]);
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
__global__ __launch_bounds__(64 * NW) void collocate_kernel(PointParams P) {
  extern __shared__ double lds[];
  __shared__ Task sT;
  __shared__ Level sL;
  const int tid = threadIdx.x, nthr = 64 * NW;
  stage_task(P, sT, sL, tid, nthr);
  const Task &T = sT;
  if (T.flags & kTaskSkip) {
    if (P.count_only && tid == 0)
      P.task_out[(long long)(P.task_base + blockIdx.x) * kTaskOutStride] = 0.0;
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
          double dip = exp(-T.zetp * (rx * rx + ry * ry + rz * rz));
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
    P.task_out[(long long)(P.task_base + blockIdx.x) * kTaskOutStride] = counted;
}

// ---------------------------------------------------------------------------
// Integrate: one workgroup per task; LPB bounds lp for register accumulators
// ---------------------------------------------------------------------------

struct CabGet {
  const double *m;
  int n1;
  __device__ __forceinline__ double operator()(Ang a, Ang b) const {
    return m[aidx(b) * n1 + aidx(a)];
  }
};

// grid_process_vab.h:29-205
__device__ __forceinline__ double fa_plain(Ang a, Ang b, int i, double za, const CabGet &c) {
  return 2.0 * za * c(up(a, i), b) - comp(a, i) * c(dn(a, i), b);
}
__device__ __forceinline__ double fb_plain(Ang a, Ang b, int i, double zb,
                                           const double rab[3], const CabGet &c) {
  return 2.0 * zb * (c(up(a, i), b) - rab[i] * c(a, b)) - comp(b, i) * c(a, dn(b, i));
}
__device__ __forceinline__ double va_plain(Ang a, Ang b, int i, int j, double za,
                                           const CabGet &c) {
  return 2.0 * za * c(up(up(a, j), i), b) - comp(a, j) * c(up(dn(a, j), i), b);
}
__device__ __forceinline__ double vb_plain(Ang a, Ang b, int i, int j, double zb,
                                           const double rab[3], const CabGet &c) {
  return 2.0 * zb *
             (c(up(up(a, j), i), b) - c(up(a, i), b) * rab[j] - c(up(a, j), b) * rab[i] +
              c(a, b) * rab[j] * rab[i]) -
         comp(b, j) * c(a, up(dn(b, j), i));
}

#define TAU_SUM(F)                                                    \
  do {                                                                \
    for (int k = 0; k < 3; k++) {                                     \
      const int ak = comp(a, k), bk = comp(b, k);                     \
      const Ang ad = dn(a, k), au = up(a, k), bd = dn(b, k), bu = up(b, k); \
      s += 0.5 * ak * bk * F(ad, bd);                                 \
      s -= za * bk * F(au, bd);                                       \
      s -= ak * zb * F(ad, bu);                                       \
      s += 2.0 * za * zb * F(au, bu);                                 \
    }                                                                 \
  } while (0)

__device__ double get_hab(Ang a, Ang b, double za, double zb, const CabGet &c, bool tau) {
  if (!tau) return c(a, b);
  double s = 0.0;
#define F(x, y) c(x, y)
  TAU_SUM(F);
#undef F
  return s;
}
__device__ double get_fa(Ang a, Ang b, int i, double za, double zb, const CabGet &c,
                         bool tau) {
  if (!tau) return fa_plain(a, b, i, za, c);
  double s = 0.0;
#define F(x, y) fa_plain(x, y, i, za, c)
  TAU_SUM(F);
#undef F
  return s;
}
__device__ double get_fb(Ang a, Ang b, int i, double za, double zb, const double rab[3],
                         const CabGet &c, bool tau) {
  if (!tau) return fb_plain(a, b, i, zb, rab, c);
  double s = 0.0;
#define F(x, y) fb_plain(x, y, i, zb, rab, c)
  TAU_SUM(F);
#undef F
  return s;
}
__device__ double get_va(Ang a, Ang b, int i, int j, double za, double zb, const CabGet &c,
                         bool tau) {
  if (!tau) return va_plain(a, b, i, j, za, c);
  double s = 0.0;
#define F(x, y) va_plain(x, y, i, j, za, c)
  TAU_SUM(F);
#undef F
  return s;
}
__device__ double get_vb(Ang a, Ang b, int i, int j, double za, double zb,
                         const double rab[3], const CabGet &c, bool tau) {
  if (!tau) return vb_plain(a, b, i, j, zb, rab, c);
  double s = 0.0;
#define F(x, y) vb_plain(x, y, i, j, zb, rab, c)
  TAU_SUM(F);
#undef F
  return s;
}
#undef TAU_SUM

// FULL = false: hab only, no tau (forces/virial/tau code compiled out).
template <int NW, int LPB, bool FULL, bool ORTHO>
__global__ __launch_bounds__(64 * NW) void integrate_kernel(PointParams P) {
  extern __shared__ double lds[];
  __shared__ Task sT;
  __shared__ Level sL;
  const int tid = threadIdx.x, nthr = 64 * NW;
  stage_task(P, sT, sL, tid, nthr);
  const Task &T = sT;
  double *out = P.task_out + (long long)(P.task_base + blockIdx.x) * kTaskOutStride;
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
          double dip = g * exp(-T.zetp * (rx * rx + ry * ry + rz * rz));
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
  double f[6] = {0, 0, 0, 0, 0, 0};
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int e = tid; e < na * nb; e += nthr) {
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
    const int lane = tid & 63, w = tid >> 6;
    for (int c = 0; c < kTaskOutStride; c++) {
      const double x = wave_sum(c < 6 ? f[c] : v[c - 6]);
      if (lane == 0) red[w * kTaskOutStride + c] = x;
    }
    __syncthreads();
    if (tid < kTaskOutStride) {
      double s = 0.0;
      for (int q = 0; q < NW; q++) s += red[q * kTaskOutStride + tid];
      const double scale = (T.iatom == T.jatom) ? 1.0 : 2.0;
      out[tid] = scale * s;
    }
  }
}

// ---------------------------------------------------------------------------
// Decontraction / contraction of matrix sub-blocks (grid_ref_task_list.c
// :237-271 and :462-499), one workgroup per (block, iset, jset) group.
// Columns of the Cartesian block are processed in chunks of `cach`.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_elem(const Group &g, const double *blk, int sa,
                                             int sb) {
  return g.transpose ? blk[(long long)(g.sb0 + sb) * g.nsgfa + g.sa0 + sa]
                     : blk[(long long)(g.sa0 + sa) * g.nsgfb + g.sb0 + sb];
}

__global__ __launch_bounds__(256) void decontract_kernel(const Group *groups,
                                                         const double *pab, long long pab_len,
                                                         const double *sphi, double *cart,
                                                         long long cart_len, int cach) {
  extern __shared__ double work[];
  const Group g = groups[blockIdx.x];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const double *blk = pab + g.block_off;
  for (int c0 = 0; c0 < g.ncoa; c0 += cach) {
    const int nc = min(cach, g.ncoa - c0);
    for (int e = tid; e < g.nsb * nc; e += nthr) {
      const int sb = e / nc, ca = c0 + e % nc;
      double s = 0.0;
      for (int sa = 0; sa < g.nsa; sa++)
        s += block_elem(g, blk, sa, sb) * sphi[g.sphi_a + (long long)sa * g.maxcoa + ca];
      work[e] = s;
    }
    __syncthreads();
    for (int e = tid; e < g.ncob * nc; e += nthr) {
      const int cb = e / nc, cc = e % nc;
      double s = 0.0;
      for (int sb = 0; sb < g.nsb; sb++)
        s += sphi[g.sphi_b + (long long)sb * g.maxcob + cb] * work[sb * nc + cc];
      const long long q = g.cart_off + (long long)cb * g.ncoa + c0 + cc;
      if (q < cart_len) cart[q] = s;
    }
    __syncthreads();
  }
  (void)pab_len;
}

__global__ __launch_bounds__(256) void contract_kernel(const Group *groups,
                                                       const double *hab_cart,
                                                       const double *sphi, double *hab,
                                                       long long hab_len, int cach) {
  extern __shared__ double lds2[];
  const Group g = groups[blockIdx.x];
  const int tid = threadIdx.x, nthr = blockDim.x;
  double *acc = lds2;                    // [nsb][nsa]
  double *work = lds2 + g.nsa * g.nsb;   // [nsb][cach]
  for (int e = tid; e < g.nsa * g.nsb; e += nthr) acc[e] = 0.0;
  __syncthreads();
  for (int c0 = 0; c0 < g.ncoa; c0 += cach) {
    const int nc = min(cach, g.ncoa - c0);
    for (int e = tid; e < g.nsb * nc; e += nthr) {
      const int sb = e / nc, cc = e % nc;
      double s = 0.0;
      for (int cb = 0; cb < g.ncob; cb++)
        s += sphi[g.sphi_b + (long long)sb * g.maxcob + cb] *
             hab_cart[g.cart_off + (long long)cb * g.ncoa + c0 + cc];
      work[e] = s;
    }
    __syncthreads();
    for (int e = tid; e < g.nsa * g.nsb; e += nthr) {
      const int sb = e / g.nsa, sa = e % g.nsa;
      double s = 0.0;
      for (int cc = 0; cc < nc; cc++)
        s += work[sb * nc + cc] * sphi[g.sphi_a + (long long)sa * g.maxcoa + c0 + cc];
      acc[e] += s;
    }
    __syncthreads();
  }
  for (int e = tid; e < g.nsa * g.nsb; e += nthr) {
    const int sb = e / g.nsa, sa = e % g.nsa;
    const long long q = g.block_off + (g.transpose ? (long long)(g.sb0 + sb) * g.nsgfa + g.sa0 + sa
                                                   : (long long)(g.sa0 + sa) * g.nsgfb + g.sb0 + sb);
    if (q < hab_len) hab[q] = acc[e];
  }
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
  const bool full = P.do_forces || P.compute_tau;
  if (ortho)
    return full ? launch_integrate_kind<true, true>(P, lp, lds, s)
                : launch_integrate_kind<false, true>(P, lp, lds, s);
  return full ? launch_integrate_kind<true, false>(P, lp, lds, s)
              : launch_integrate_kind<false, false>(P, lp, lds, s);
}

hipError_t launch_decontract(const Group *groups, int ngroups, const double *pab,
                             long long pab_len, const double *sphi, double *cart,
                             long long cart_len, int cach, int max_nsb, hipStream_t s) {
  if (ngroups == 0) return hipSuccess;
  const size_t lds = sizeof(double) * (size_t)cach * (size_t)max_nsb;
  hipLaunchKernelGGL(decontract_kernel, dim3(ngroups), dim3(256), lds, s, groups, pab,
                     pab_len, sphi, cart, cart_len, cach);
  return hipGetLastError();
}

hipError_t launch_contract(const Group *groups, int ngroups, const double *hab_cart,
                           const double *sphi, double *hab, long long hab_len, int cach,
                           int max_nsb, int max_nsa_nsb, hipStream_t s) {
  if (ngroups == 0) return hipSuccess;
  const size_t lds = sizeof(double) * ((size_t)max_nsa_nsb + (size_t)cach * (size_t)max_nsb);
  hipLaunchKernelGGL(contract_kernel, dim3(ngroups), dim3(256), lds, s, groups, hab_cart,
                     sphi, hab, hab_len, cach);
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

hipError_t launch_reduce_virial(const double *task_out, long long ntasks, double *virial,
                                hipStream_t s) {
  hipLaunchKernelGGL(reduce_virial_kernel, dim3(1), dim3(256), 0, s, task_out, ntasks, virial);
  return hipGetLastError();
}

}  // namespace gm
