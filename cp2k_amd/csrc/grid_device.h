// Device helpers shared by the per-task kernels (grid_kernels.hip) and the
// wave-per-task kernels (grid_wave.hip): Cartesian orbital tables, the
// prepare_pab stencils, binomial re-centring, per-axis tables, truncation
// bounds and the process_vab getters.  Compiled with -ffp-contract=off.
#pragma once

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

static __constant__ OrbTable c_orb = make_orb_table();

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

static __device__ void mixed_b(int d, double fa, Ang a, Ang b, double zb, const CabRef &c) {
  const int i1 = d, i2 = (d + 1) % 3;
  const int b1 = comp(b, i1), b2 = comp(b, i2);
  c.add(a, dn(dn(b, i2), i1), b1 * b2 * fa);
  c.add(a, up(dn(b, i2), i1), -2.0 * zb * b2 * fa);
  c.add(a, dn(up(b, i2), i1), -2.0 * zb * b1 * fa);
  c.add(a, up(up(b, i2), i1), +4.0 * zb * zb * fa);
}

static __device__ void pure_b(int d, double fa, Ang a, Ang b, double zb, const CabRef &c) {
  const int bd = comp(b, d);
  c.add(a, dn(dn(b, d), d), bd * (bd - 1) * fa);
  c.add(a, b, -2.0 * zb * (2 * bd + 1) * fa);
  c.add(a, up(up(b, d), d), +4.0 * zb * zb * fa);
}

static __device__ void apply_func(int func, Ang a, Ang b, double za, double zb, double p,
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
static __device__ void build_alpha(const Task &T, int la, int lb, double *alpha, int tid,
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
static __device__ void build_ortho_tables(const Task &T, const Layout &lay, int lp,
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
static __device__ void build_general_maps(const Task &T, const Layout &lay,
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
static __device__ void cxyz_cijk(bool forward, int lp, const double dh[3][3], int NP,
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

// Forces (and virial) of one (a, b) pair without tau: the plain getters
// above with each cab element looked up once (c(a, b), c(a + 1_i, b) and
// c(a + 1_i + 1_j, b) serve several of them).
__device__ __forceinline__ void pair_forces_plain(Ang a, Ang b, double za, double zb,
                                                  const double rab[3], const CabGet &c,
                                                  double pv, bool virial, double *f, double *v) {
  const double *row = c.m + aidx(b) * c.n1;  // c(x, b) = row[aidx(x)]
  const double c0 = row[aidx(a)];
  double cu[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    cu[i] = row[aidx(up(a, i))];
    f[i] += pv * (2.0 * za * cu[i] - comp(a, i) * row[aidx(dn(a, i))]);
    f[3 + i] += pv * (2.0 * zb * (cu[i] - rab[i] * c0) - comp(b, i) * c(a, dn(b, i)));
  }
  if (!virial) return;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const double cuu = row[aidx(up(up(a, j), i))];
      v[i * 3 + j] += pv * (2.0 * za * cuu - comp(a, j) * row[aidx(up(dn(a, j), i))]);
      v[i * 3 + j] += pv * (2.0 * zb * (cuu - cu[i] * rab[j] - cu[j] * rab[i] + c0 * rab[j] * rab[i]) -
                            comp(b, j) * c(a, up(dn(b, j), i)));
    }
}

static __device__ double get_hab(Ang a, Ang b, double za, double zb, const CabGet &c, bool tau) {
  if (!tau) return c(a, b);
  double s = 0.0;
#define F(x, y) c(x, y)
  TAU_SUM(F);
#undef F
  return s;
}
static __device__ double get_fa(Ang a, Ang b, int i, double za, double zb, const CabGet &c,
                         bool tau) {
  if (!tau) return fa_plain(a, b, i, za, c);
  double s = 0.0;
#define F(x, y) fa_plain(x, y, i, za, c)
  TAU_SUM(F);
#undef F
  return s;
}
static __device__ double get_fb(Ang a, Ang b, int i, double za, double zb, const double rab[3],
                         const CabGet &c, bool tau) {
  if (!tau) return fb_plain(a, b, i, zb, rab, c);
  double s = 0.0;
#define F(x, y) fb_plain(x, y, i, zb, rab, c)
  TAU_SUM(F);
#undef F
  return s;
}
static __device__ double get_va(Ang a, Ang b, int i, int j, double za, double zb, const CabGet &c,
                         bool tau) {
  if (!tau) return va_plain(a, b, i, j, za, c);
  double s = 0.0;
#define F(x, y) va_plain(x, y, i, j, za, c)
  TAU_SUM(F);
#undef F
  return s;
}
static __device__ double get_vb(Ang a, Ang b, int i, int j, double za, double zb,
                         const double rab[3], const CabGet &c, bool tau) {
  if (!tau) return vb_plain(a, b, i, j, zb, rab, c);
  double s = 0.0;
#define F(x, y) vb_plain(x, y, i, j, zb, rab, c)
  TAU_SUM(F);
#undef F
  return s;
}
#undef TAU_SUM

// ---------------------------------------------------------------------------
// Wave-level helpers of the wave-per-task kernels
// ---------------------------------------------------------------------------
// LDS written by one lane and read by another lane of the same wave: LDS
// requests of a wave complete in order, so only compiler reordering must be
// prevented.
static __device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int CTRL, int ROWS>
static __device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, true);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes (all active); the total is valid in lane 63.
static __device__ __forceinline__ double wave_total(double v) {
  v += dpp_mov<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += dpp_mov<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0x141, 0xf>(v);  // row_half_mirror
  v += dpp_mov<0x140, 0xf>(v);  // row_mirror
  v += dpp_mov<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_mov<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

static __device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}

static __device__ __forceinline__ int pm(int a, int m) {
  const int r = a % m;
  return r < 0 ? r + m : r;
}

static __device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// a mod n (in [0, n)).  One period either side of [0, n) takes two selects;
// anything farther (cells smaller than the task) the remainder.
static __device__ __forceinline__ int wrap(int a, int n) {
  a = a < 0 ? a + n : a;
  a = a >= n ? a - n : a;
  if ((unsigned)a >= (unsigned)n) a = pm(a, n);
  return a;
}

// Packed index of (lx, ly, lz), lx + ly + lz <= lp: lz slowest, lx fastest.
static __device__ __forceinline__ int packed_index(int lx, int ly, int lz, int lp) {
  int off = 0;
  for (int z = 0; z < lz; z++) off += (lp - z + 1) * (lp - z + 2) / 2;
  for (int y = 0; y < ly; y++) off += lp - lz - y + 1;
  return off + lx;
}

// Multiply-add in the evaluation type: FP64, or FP32 in FP32-mixed mode.
static __device__ __forceinline__ double fmar(double a, double b, double c) { return fma(a, b, c); }
static __device__ __forceinline__ float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }



// Sums over the 64 lanes of N per-lane values v[0..N-1], stored to out[e]
// (each entry by one lane).  Recursive halving: at each of the six steps a
// lane keeps half of its entries (the lower or upper half, by one bit of its
// lane id), adds its partner's copy of them and sends the other half: about
// N exchanges and N additions in all, against 6 N of each for N separate
// wave_total calls.  Partners: row_mirror (lane bit 3), row_half_mirror
// (bit 2), quad_perm xor 2, xor 1 (all DPP), then xor 16 and xor 32.  The
// mirrors also flip the lower bits, which is why they come first: a step's
// partners must hold the same entries, i.e. agree in the bits used so far.
template <int S>
static __device__ __forceinline__ double halving_partner(double v) {
  if constexpr (S == 0) return dpp_mov<0x140, 0xf>(v);
  else if constexpr (S == 1) return dpp_mov<0x141, 0xf>(v);
  else if constexpr (S == 2) return dpp_mov<0x4e, 0xf>(v);
  else if constexpr (S == 3) return dpp_mov<0xb1, 0xf>(v);
  else if constexpr (S == 4) return __shfl_xor(v, 16, 64);
  else return __shfl_xor(v, 32, 64);
}

template <int N, int S>
struct HalvingSum {
  static constexpr int M = (N + 1) / 2;  // entries kept at this step
  static constexpr int kBit = S == 0 ? 8 : S == 1 ? 4 : S == 2 ? 2 : S == 3 ? 1 : S == 4 ? 16 : 32;
  // v holds N entries of which the first cnt are real (the rest padding);
  // they are the sums' entries base .. base + cnt - 1.
  static __device__ __forceinline__ void run(const double *v, int lane, int base, int cnt,
                                             double *out) {
    const bool hi = (lane & kBit) != 0;
    double w[M];
#pragma unroll
    for (int i = 0; i < M; i++) {
      const double lo_v = v[i];
      const double hi_v = (M + i < N) ? v[(M + i < N) ? M + i : 0] : 0.0;
      const double keep = hi ? hi_v : lo_v;
      const double send = hi ? lo_v : hi_v;
      w[i] = keep + halving_partner<S>(send);
    }
    const int b = base + (hi ? M : 0);
    const int c = hi ? max(cnt - M, 0) : min(cnt, M);
    if constexpr (S < 5) {
      HalvingSum<M, S + 1>::run(w, lane, b, c, out);
    } else {
#pragma unroll
      for (int i = 0; i < M; i++)
        if (i < c) out[b + i] = w[i];
    }
  }
};

template <int N>
__device__ __forceinline__ void wave_sum_store(const double (&v)[N], double *out, int lane) {
  HalvingSum<N, 0>::run(v, lane, 0, N, out);
}

}  // namespace gm
