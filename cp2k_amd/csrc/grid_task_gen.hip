// Task-list generation on the device for replicated real-space grids: from
// the orbital neighbour list to the argument set of grid_create_task_list.
//
//   neighbour-list loop    src/task_list_methods.F:246-294
//   task_list_inner_loop   :652-719  per pair, every (iset, jset, ipgf, jpgf)
//                          whose set / kind / pgf radii reach |rab| is a task
//   compute_pgf_properties :755-800  radius = exp_radius_very_extended
//                          (src/aobasis/ao_util.F:207-290) with prefactor
//                          exp(-zeta zetb / zetp |rab|^2) and cutoff 1
//   exp_radius             ao_util.F:95-190  bracket growth r -> 2r + 1, then
//                          interval contraction at 0.38 / 0.62 (single-
//                          precision literals), exponentials in single
//                          precision, r**l by binary powering
//   gaussian_gridlevel     src/pw_env/gaussian_gridlevels.F:146-166
//   get_atom_pair          :2070-2136  blocks = distinct (min atom, max atom,
//                          image) of the tasks, numbered in that order
//   tasks_sort             :2931-2960  stable order by (level, image, iatom,
//                          jatom, iset, jset, ipgf, jpgf)
//   rs_calc_offsets        :2439-2484  block offsets in the pab buffer
//
// Pipeline (library stream): a wave per neighbour-list entry counts its
// tasks (screening only: a lane per candidate (iset, jset, ipgf, jpgf) in
// loop order, ballot counts); exclusive scan; a wave per entry writes its
// tasks at its offset in loop order (ballot prefix), with the radius
// bisection per lane and the task's sort keys; two stable radix sorts (the
// (iset, jset, ipgf, jpgf) key, then the (level, image, iatom, jatom) key)
// give the reference's order, in which ties keep neighbour-list order as
// the merge sort does; a radix sort of the entries' block keys numbers the
// blocks; a gather writes the lists.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/grid_task_gen.h"

namespace gm {
hipStream_t library_stream();
}

namespace {

[[noreturn]] void tg_fatal(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "grid_mi355x error: ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  va_end(ap);
  fflush(stderr);
  abort();
}

#define TG_HIP(call)                                                                        \
  do {                                                                                      \
    const hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                                   \
      tg_fatal("%s failed at %s:%d: %s", #call, __FILE__, __LINE__, hipGetErrorString(e_)); \
  } while (0)

constexpr int kMaxLevels = 16;
constexpr int kMaxSetL = 10;  // lmax of a set: la_max + lb_max <= 20
constexpr int kWavesPerBlock = 4;

// Basis data of the kinds, flattened.  ints: per kind nset, nsgf, largest
// npgf (3 * nkinds), then lmax [nkinds][maxset] and npgf [nkinds][maxset].
// doubles: zet, pgf_radius [nkinds][maxset][maxpgf], set_radius
// [nkinds][maxset], kind_radius [nkinds].
struct Kinds {
  const int *kint;
  const double *kdbl;
  int nkinds, maxset, maxpgf;
  __device__ int nset(int k) const { return kint[3 * k]; }
  __device__ int nsgf(int k) const { return kint[3 * k + 1]; }
  __device__ int mpgf(int k) const { return kint[3 * k + 2]; }
  __device__ int lmax(int k, int s) const { return kint[3 * nkinds + k * maxset + s]; }
  __device__ int npgf(int k, int s) const {
    return kint[3 * nkinds + nkinds * maxset + k * maxset + s];
  }
  __device__ double zet(int k, int s, int p) const {
    return kdbl[((size_t)k * maxset + s) * maxpgf + p];
  }
  __device__ double pgf_radius(int k, int s, int p) const {
    return kdbl[(size_t)nkinds * maxset * maxpgf + ((size_t)k * maxset + s) * maxpgf + p];
  }
  __device__ double set_radius(int k, int s) const {
    return kdbl[2 * (size_t)nkinds * maxset * maxpgf + k * maxset + s];
  }
  __device__ double kind_radius(int k) const {
    return kdbl[2 * (size_t)nkinds * maxset * maxpgf + (size_t)nkinds * maxset + k];
  }
};

// Bit fields of the sort keys (widths from the input sizes).  Task key hi:
// level | image | iatom | jatom (zero based); task key lo: iset | jset | ipgf
// | jpgf; block key: row | col | image, and bit blk_bits set for an entry
// without tasks (sorts last).
struct KeyBits {
  int jat, iat, img, lev, hi_bits;
  int pgf, js, is, lo_bits;
  int blk_col, blk_row, blk_bits;
};

struct GenParams {
  int npairs;
  const int *pair_atoms;    // [npairs][2], 1-based
  const double *pair_rab;   // [npairs][3]
  const int *pair_image;    // [npairs], 1-based, or null (all 1)
  const int *atom_kinds;    // [natoms], 1-based
  const double *atom_pos;   // [natoms][3]
  Kinds K;
  KeyBits B;
  int nlevels;
  double cutoff[kMaxLevels];
  double rel_cutoff, eps;
  long long *counts;        // [npairs]
  const long long *offsets; // [npairs]
  unsigned long long *pair_key;  // [npairs] block key
  int *pair_val;            // [npairs] entry index
  // tasks in neighbour-list order
  unsigned long long *t_hi;
  unsigned int *t_lo;
  int *t_pair;
  double *t_radius;
  int *t_iota;
};

__device__ __forceinline__ unsigned long long lanes_below(int lane) {
  return (1ull << lane) - 1ull;
}

// EXP of a REAL(sp) argument: the argument rounded to float, the exponential
// rounded to float.
__device__ __forceinline__ double exp_sp(double x) {
  return (double)(float)exp((double)(float)x);
}

// x**n for integer n >= 0 by right-to-left binary powering (the expansion
// of an integer power of a real with a variable exponent).
__device__ __forceinline__ double powi(double x, int n) {
  double y = (n & 1) ? x : 1.0;
  while (n >>= 1) {
    x = x * x;
    if (n & 1) y = y * x;
  }
  return y;
}

__device__ double exp_radius(int l, double alpha, double threshold, double prefactor,
                             double epsabs, double rlow) {
  const double c0 = (double)0.38f, c1 = (double)0.62f;
  const double a = fabs(alpha), t = fabs(threshold);
  double radius = rlow;
  if (prefactor == 0.0) return radius;
  double r = fmax(sqrt(0.5 * (double)l / a), radius);
  const double d = fabs(prefactor);
  double g = d;
  if (l != 0) g = g * exp_sp(-a * r * r) * powi(r, l);
  if (g < t) return radius;
  radius = r * 2.0 + 1.0;
  int i = 1;
  for (; i <= 5000; i++) {
    g = d * exp_sp(-a * radius * radius) * powi(radius, l);
    if (g < t) break;
    r = radius;
    radius = r * 2.0 + 1.0;
  }
  double dr = 0.0;
  for (i = i + 1; i <= 5000; i++) {
    const double rd = radius - r;
    if (rd < epsabs || rd == dr) return radius;
    const double s0 = r + rd * c0, s1 = r + rd * c1;
    const double g0 = d * exp_sp(-a * s0 * s0) * powi(s0, l);
    const double g1 = d * exp_sp(-a * s1 * s1) * powi(s1, l);
    if (g0 < t) {
      radius = s0;
    } else {
      r = s0;
      if (g1 < t)
        radius = s1;
      else
        r = s1;
    }
    dr = rd;
  }
  return radius;
}

// exp_radius_very_extended without pab (prefactor_local = prefactor).
__device__ double product_radius(int la_max, int lb_max, double rad_a, double rad_b,
                                 double zetp, double eps, double prefactor) {
  double poly[2 * kMaxSetL + 1];
  const int lp = la_max + lb_max;
  for (int i = 0; i <= lp; i++) poly[i] = 0.0;
  for (int lxa = 0; lxa <= la_max; lxa++)
    for (int lxb = 0; lxb <= lb_max; lxb++) {
      double coef[2 * kMaxSetL + 1];
      for (int i = 0; i <= lxa + lxb; i++) coef[i] = 0.0;
      double bini = 1.0, s1 = 1.0;
      for (int i = 0; i <= lxa; i++) {
        double binj = 1.0, s2 = 1.0;
        for (int j = 0; j <= lxb; j++) {
          coef[lxa + lxb - i - j] += bini * binj * s1 * s2;
          binj = (binj * (double)(lxb - j)) / (double)(j + 1);
          s2 = s2 * rad_b;
        }
        bini = (bini * (double)(lxa - i)) / (double)(i + 1);
        s1 = s1 * rad_a;
      }
      for (int i = 0; i <= lxa + lxb; i++) poly[i] = fmax(poly[i], coef[i]);
    }
  double radius = 0.0;
  for (int i = 0; i <= lp; i++)
    radius = fmax(radius, exp_radius(i, zetp, eps, poly[i] * prefactor, 1.0e-2, radius));
  return radius;
}

__device__ int grid_level(const GenParams &P, double zetp) {
  int level = 1;
  const double need = fabs(zetp) * P.rel_cutoff;
  for (int i = 0; i < P.nlevels; i++)
    if (P.cutoff[i] + 1e-6 >= need) level = i + 1;
  return level;
}

// One neighbour-list entry and its candidate space: candidate c in
// [0, ncand) is (is, js, ip, jp) in loop order (jp fastest).
struct Entry {
  int ia, ja, ka, kb, image;
  int nsb, mpa, mpb, ncand;
  double rab[3], rab2, dab, kind_radius_b;
};

__device__ Entry load_entry(const GenParams &P, int p) {
  Entry e;
  e.ia = P.pair_atoms[2 * p] - 1;
  e.ja = P.pair_atoms[2 * p + 1] - 1;
  e.ka = P.atom_kinds[e.ia] - 1;
  e.kb = P.atom_kinds[e.ja] - 1;
  e.image = P.pair_image ? P.pair_image[p] : 1;
  for (int d = 0; d < 3; d++) e.rab[d] = P.pair_rab[3 * p + d];
  e.rab2 = e.rab[0] * e.rab[0] + e.rab[1] * e.rab[1] + e.rab[2] * e.rab[2];
  e.dab = sqrt(e.rab2);
  e.kind_radius_b = P.K.kind_radius(e.kb);
  e.nsb = P.K.nset(e.kb);
  e.mpa = P.K.mpgf(e.ka);
  e.mpb = P.K.mpgf(e.kb);
  e.ncand = P.K.nset(e.ka) * e.nsb * e.mpa * e.mpb;
  return e;
}

// The screening of task_list_inner_loop for candidate c (the CYCLEs when a
// radius sum is below |rab|).
__device__ bool screen(const GenParams &P, const Entry &e, int c, int &is, int &js, int &ip,
                       int &jp) {
  jp = c % e.mpb;
  int q = c / e.mpb;
  ip = q % e.mpa;
  q /= e.mpa;
  js = q % e.nsb;
  is = q / e.nsb;
  const Kinds &K = P.K;
  if (ip >= K.npgf(e.ka, is) || jp >= K.npgf(e.kb, js)) return false;
  const double sra = K.set_radius(e.ka, is), srb = K.set_radius(e.kb, js);
  if (sra + e.kind_radius_b < e.dab) return false;
  if (sra + srb < e.dab) return false;
  const double rpa = K.pgf_radius(e.ka, is, ip);
  if (rpa + srb < e.dab) return false;
  if (rpa + K.pgf_radius(e.kb, js, jp) < e.dab) return false;
  return true;
}

__global__ __launch_bounds__(64 * kWavesPerBlock) void count_kernel(GenParams P) {
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (p >= P.npairs) return;  // uniform per wave
  const Entry e = load_entry(P, p);
  long long n = 0;
  for (int base = 0; base < e.ncand; base += 64) {
    int is, js, ip, jp;
    const bool ok = base + lane < e.ncand && screen(P, e, base + lane, is, js, ip, jp);
    n += __popcll(__ballot(ok));
  }
  if (lane == 0) {
    P.counts[p] = n;
    const int row = min(e.ia, e.ja), col = max(e.ia, e.ja);
    P.pair_key[p] = n ? ((unsigned long long)row << P.B.blk_row |
                         (unsigned long long)col << P.B.blk_col | (unsigned long long)(e.image - 1))
                      : 1ull << P.B.blk_bits;
    P.pair_val[p] = p;
  }
}

__global__ __launch_bounds__(64 * kWavesPerBlock) void fill_kernel(GenParams P) {
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (p >= P.npairs) return;
  const Entry e = load_entry(P, p);
  const Kinds &K = P.K;
  const double *ra = P.atom_pos + 3 * e.ia;
  long long pos = P.offsets[p];
  for (int base = 0; base < e.ncand; base += 64) {
    int is = 0, js = 0, ip = 0, jp = 0;
    const bool ok = base + lane < e.ncand && screen(P, e, base + lane, is, js, ip, jp);
    const unsigned long long m = __ballot(ok);
    if (ok) {
      const long long q = pos + __popcll(m & lanes_below(lane));
      const double zeta = K.zet(e.ka, is, ip), zetb = K.zet(e.kb, js, jp);
      const double zetp = zeta + zetb;
      const double f = zetb / zetp;
      const double prefactor = exp(-zeta * f * e.rab2);
      // rp = ra + zetb / zetp rab, rb = ra + rab; rad_a = |ra - rp|, rad_b = |rb - rp|
      double sa = 0.0, sb = 0.0;
      for (int d = 0; d < 3; d++) {
        const double rp = ra[d] + f * e.rab[d];
        const double rb = ra[d] + e.rab[d];
        sa = sa + (ra[d] - rp) * (ra[d] - rp);
        sb = sb + (rb - rp) * (rb - rp);
      }
      const double radius = product_radius(K.lmax(e.ka, is), K.lmax(e.kb, js), sqrt(sa),
                                           sqrt(sb), zetp, P.eps, prefactor);
      const int level = grid_level(P, zetp);
      P.t_hi[q] = (unsigned long long)(level - 1) << P.B.lev |
                  (unsigned long long)(e.image - 1) << P.B.img |
                  (unsigned long long)e.ia << P.B.iat | (unsigned long long)e.ja << P.B.jat;
      P.t_lo[q] = (unsigned)is << P.B.is | (unsigned)js << P.B.js | (unsigned)ip << P.B.pgf |
                  (unsigned)jp;
      P.t_pair[q] = p;
      P.t_radius[q] = radius;
      P.t_iota[q] = (int)q;
    }
    pos += __popcll(m);
  }
}

__global__ void gather_hi_kernel(const unsigned long long *hi, const int *idx,
                                 unsigned long long *out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = hi[idx[i]];
}

__global__ void total_kernel(const long long *counts, const long long *offsets, int n,
                             long long *total) {
  if (threadIdx.x == 0 && blockIdx.x == 0) total[0] = n ? offsets[n - 1] + counts[n - 1] : 0;
}

__global__ void zero_kernel(long long *x, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0;
}

// Sorted block keys -> head flags (first entry of each distinct key with tasks).
__global__ void block_heads_kernel(const unsigned long long *key, int n, int blk_bits,
                                   int *head) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool valid = !(key[i] >> blk_bits);
  head[i] = valid && (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// bid = inclusive scan of the heads: entry val[i] belongs to block bid[i]
// (0: no tasks); each head records its block's size nsgf(row) * nsgf(col).
__global__ void block_number_kernel(GenParams P, const unsigned long long *key, const int *val,
                                    const int *head, const int *bid, int n, int *pair_block,
                                    long long *block_size) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool valid = !(key[i] >> P.B.blk_bits);
  pair_block[val[i]] = valid ? bid[i] : 0;
  if (head[i]) {
    const int p = val[i];
    const int ka = P.atom_kinds[P.pair_atoms[2 * p] - 1] - 1;
    const int kb = P.atom_kinds[P.pair_atoms[2 * p + 1] - 1] - 1;
    block_size[bid[i] - 1] = (long long)P.K.nsgf(ka) * P.K.nsgf(kb);
  }
}

struct OutLists {
  int *level, *iatom, *jatom, *iset, *jset, *ipgf, *jpgf, *block;
  double *radius, *rab;
};

__global__ void gather_tasks_kernel(GenParams P, const unsigned long long *hi_sorted,
                                    const int *perm, const int *pair_block, long long n,
                                    OutLists O) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int g = perm[i];
  const unsigned long long hi = hi_sorted[i];
  const unsigned lo = P.t_lo[g];
  const int p = P.t_pair[g];
  const unsigned pm = (1u << P.B.pgf) - 1u;
  const unsigned sm = (1u << (P.B.is - P.B.js)) - 1u;
  O.level[i] = (int)(hi >> P.B.lev) + 1;
  O.iatom[i] = P.pair_atoms[2 * p];
  O.jatom[i] = P.pair_atoms[2 * p + 1];
  O.iset[i] = (int)(lo >> P.B.is) + 1;
  O.jset[i] = (int)((lo >> P.B.js) & sm) + 1;
  O.ipgf[i] = (int)((lo >> P.B.pgf) & pm) + 1;
  O.jpgf[i] = (int)(lo & pm) + 1;
  O.block[i] = pair_block[p];
  O.radius[i] = P.t_radius[g];
  for (int d = 0; d < 3; d++) O.rab[3 * i + d] = P.pair_rab[3 * p + d];
}

// Growable device arrays kept between calls.
struct Arena {
  std::vector<std::pair<void *, size_t>> slots;
  void *get(int slot, size_t bytes) {
    if ((int)slots.size() <= slot) slots.resize(slot + 1, {nullptr, 0});
    auto &s = slots[slot];
    if (bytes > s.second) {
      if (s.first) TG_HIP(hipFree(s.first));
      s.first = nullptr;
      TG_HIP(hipMalloc(&s.first, bytes));
      s.second = bytes;
    }
    return s.first;
  }
};

struct GenState {
  std::mutex mu;
  Arena arena;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.f;
};

GenState &gen_state() {
  static GenState *st = new GenState();  // kept until the process ends
  return *st;
}

int bits_for(long long n) {  // bits of the values 0 .. n - 1
  int b = 0;
  while ((1LL << b) < n) b++;
  return b;
}

template <typename T>
T *arena_array(Arena &A, int slot, size_t n) {
  return static_cast<T *>(A.get(slot, sizeof(T) * std::max<size_t>(n, 1)));
}

template <typename T>
const T *upload(Arena &A, int slot, const T *src, size_t n, hipStream_t s) {
  T *p = arena_array<T>(A, slot, n);
  if (n) TG_HIP(hipMemcpyAsync(p, src, sizeof(T) * n, hipMemcpyHostToDevice, s));
  return p;
}

unsigned grid_for(long long n, int threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace

extern "C" {

long long grid_mi355x_generate_task_list(
    const int npairs, const int *pair_atoms, const double *pair_rab, const int *pair_image,
    const int natoms, const int *atom_kinds, const double *atom_positions, const int nkinds,
    const grid_basis_set *const *basis_sets, const int maxset, const int maxpgf,
    const double *set_radius, const double *pgf_radius, const double *kind_radius,
    const int nlevels, const double *cutoff, const double rel_cutoff, const double eps_rho,
    const long long capacity, int *level_list, int *iatom_list, int *jatom_list,
    int *iset_list, int *jset_list, int *ipgf_list, int *jpgf_list, int *border_mask_list,
    int *block_num_list, double *radius_list, double *rab_list, int *nblocks,
    int *block_offsets, long long *buffer_size) {
  if (npairs < 0 || natoms < 1 || nkinds < 1) tg_fatal("generate_task_list: bad sizes");
  if (nlevels < 1 || nlevels > kMaxLevels)
    tg_fatal("generate_task_list: %d grid levels (1..%d supported)", nlevels, kMaxLevels);
  if (maxset < 1 || maxset > 256 || maxpgf < 1 || maxpgf > 256)
    tg_fatal("generate_task_list: maxset / maxpgf must be in 1..256");
  if (natoms > (1 << 21)) tg_fatal("generate_task_list: more than 2^21 atoms");
  if (!(eps_rho > 0.0)) tg_fatal("generate_task_list: eps_rho must be positive");
  // kinds
  std::vector<int> kint(3 * nkinds + 2 * nkinds * maxset, 0);
  const size_t npk = (size_t)nkinds * maxset * maxpgf;
  std::vector<double> kdbl(2 * npk + (size_t)nkinds * maxset + nkinds, 0.0);
  for (int k = 0; k < nkinds; k++) {
    const grid_basis_set *b = basis_sets[k];
    if (!b) tg_fatal("generate_task_list: basis set of kind %d is NULL", k + 1);
    if (b->nset > maxset) tg_fatal("generate_task_list: kind %d has more than maxset sets", k + 1);
    int mp = 0;
    double kr = 0.0;
    for (int s = 0; s < b->nset; s++) {
      if (b->lmax[s] < 0 || b->lmax[s] > kMaxSetL)
        tg_fatal("generate_task_list: kind %d set %d: l above %d", k + 1, s + 1, kMaxSetL);
      if (b->npgf[s] < 0 || b->npgf[s] > maxpgf || b->npgf[s] > b->maxpgf)
        tg_fatal("generate_task_list: kind %d set %d: npgf above maxpgf", k + 1, s + 1);
      mp = std::max(mp, b->npgf[s]);
      kint[3 * nkinds + k * maxset + s] = b->lmax[s];
      kint[3 * nkinds + nkinds * maxset + k * maxset + s] = b->npgf[s];
      for (int p = 0; p < b->npgf[s]; p++) {
        const size_t at = ((size_t)k * maxset + s) * maxpgf + p;
        kdbl[at] = b->zet[s * b->maxpgf + p];
        kdbl[npk + at] = pgf_radius[at];
      }
      kdbl[2 * npk + k * maxset + s] = set_radius[k * maxset + s];
      kr = std::max(kr, set_radius[k * maxset + s]);
    }
    kint[3 * k] = b->nset;
    kint[3 * k + 1] = b->nsgf;
    kint[3 * k + 2] = mp;
    kdbl[2 * npk + (size_t)nkinds * maxset + k] = kind_radius ? kind_radius[k] : kr;
  }
  int max_image = 1;
  for (int p = 0; p < npairs; p++) {
    for (int q = 0; q < 2; q++)
      if (pair_atoms[2 * p + q] < 1 || pair_atoms[2 * p + q] > natoms)
        tg_fatal("generate_task_list: entry %d names atom %d of %d", p + 1, pair_atoms[2 * p + q],
                 natoms);
    if (pair_image) {
      if (pair_image[p] < 1) tg_fatal("generate_task_list: entry %d has image < 1", p + 1);
      max_image = std::max(max_image, pair_image[p]);
    }
  }
  for (int a = 0; a < natoms; a++)
    if (atom_kinds[a] < 1 || atom_kinds[a] > nkinds)
      tg_fatal("generate_task_list: atom %d has kind %d of %d", a + 1, atom_kinds[a], nkinds);
  KeyBits B{};
  const int ba = bits_for(natoms), bimg = bits_for(max_image), blev = bits_for(nlevels);
  B.jat = 0;
  B.iat = ba;
  B.img = 2 * ba;
  B.lev = 2 * ba + bimg;
  B.hi_bits = B.lev + blev;
  const int bs = bits_for(maxset), bp = bits_for(maxpgf);
  B.pgf = bp;
  B.js = 2 * bp;
  B.is = 2 * bp + bs;
  B.lo_bits = B.is + bs;
  B.blk_col = bimg;
  B.blk_row = bimg + ba;
  B.blk_bits = bimg + 2 * ba;
  if (B.hi_bits > 64 || B.blk_bits > 63)
    tg_fatal("generate_task_list: %d atoms x %d images x %d levels exceed the sort key", natoms,
             max_image, nlevels);

  GenState &st = gen_state();
  std::lock_guard<std::mutex> guard(st.mu);
  Arena &A = st.arena;
  const hipStream_t s = gm::library_stream();
  if (!st.ev0) {
    TG_HIP(hipEventCreate(&st.ev0));
    TG_HIP(hipEventCreate(&st.ev1));
  }
  TG_HIP(hipEventRecord(st.ev0, s));
  GenParams P{};
  P.npairs = npairs;
  P.pair_atoms = upload(A, 0, pair_atoms, 2 * (size_t)npairs, s);
  P.pair_rab = upload(A, 1, pair_rab, 3 * (size_t)npairs, s);
  P.pair_image = pair_image ? upload(A, 2, pair_image, (size_t)npairs, s) : nullptr;
  P.atom_kinds = upload(A, 3, atom_kinds, (size_t)natoms, s);
  P.atom_pos = upload(A, 4, atom_positions, 3 * (size_t)natoms, s);
  P.K.kint = upload(A, 5, kint.data(), kint.size(), s);
  P.K.kdbl = upload(A, 6, kdbl.data(), kdbl.size(), s);
  P.K.nkinds = nkinds;
  P.K.maxset = maxset;
  P.K.maxpgf = maxpgf;
  P.B = B;
  P.nlevels = nlevels;
  for (int l = 0; l < nlevels; l++) P.cutoff[l] = cutoff[l];
  P.rel_cutoff = rel_cutoff;
  P.eps = eps_rho;
  P.counts = arena_array<long long>(A, 7, npairs);
  long long *offsets = arena_array<long long>(A, 8, npairs);
  P.offsets = offsets;
  P.pair_key = arena_array<unsigned long long>(A, 9, npairs);
  P.pair_val = arena_array<int>(A, 10, npairs);
  long long *total = arena_array<long long>(A, 11, 1);

  const unsigned wave_blocks = grid_for(npairs, kWavesPerBlock);
  long long ntasks = 0;
  if (npairs > 0) {
    hipLaunchKernelGGL(count_kernel, dim3(wave_blocks), dim3(64 * kWavesPerBlock), 0, s, P);
    TG_HIP(hipGetLastError());
    size_t tb = 0;
    TG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, P.counts, offsets, npairs, s));
    void *tmp = A.get(12, tb);
    TG_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, P.counts, offsets, npairs, s));
    hipLaunchKernelGGL(total_kernel, dim3(1), dim3(64), 0, s, (const long long *)P.counts,
                       (const long long *)offsets, npairs, total);
    TG_HIP(hipGetLastError());
    TG_HIP(hipMemcpyAsync(&ntasks, total, sizeof(long long), hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
  }
  if (ntasks > capacity || ntasks == 0) {
    if (ntasks == 0) {
      *nblocks = 0;
      *buffer_size = 0;
    }
    TG_HIP(hipEventRecord(st.ev1, s));
    TG_HIP(hipEventSynchronize(st.ev1));
    TG_HIP(hipEventElapsedTime(&st.last_ms, st.ev0, st.ev1));
    return ntasks;
  }
  if (ntasks >= (1LL << 31)) tg_fatal("generate_task_list: %lld tasks exceed 2^31", ntasks);
  const int n = (int)ntasks;
  P.t_hi = arena_array<unsigned long long>(A, 13, n);
  P.t_lo = arena_array<unsigned>(A, 14, n);
  P.t_pair = arena_array<int>(A, 15, n);
  P.t_radius = arena_array<double>(A, 16, n);
  P.t_iota = arena_array<int>(A, 17, n);
  hipLaunchKernelGGL(fill_kernel, dim3(wave_blocks), dim3(64 * kWavesPerBlock), 0, s, P);
  TG_HIP(hipGetLastError());

  // scratch of the sorts and scans
  unsigned *lo_sorted = arena_array<unsigned>(A, 18, n);
  int *perm1 = arena_array<int>(A, 19, n);
  unsigned long long *hi1 = arena_array<unsigned long long>(A, 20, n);
  unsigned long long *hi2 = arena_array<unsigned long long>(A, 21, n);
  int *perm2 = arena_array<int>(A, 22, n);
  unsigned long long *bkey = arena_array<unsigned long long>(A, 23, npairs);
  int *bval = arena_array<int>(A, 24, npairs);
  int *head = arena_array<int>(A, 25, npairs);
  int *bid = arena_array<int>(A, 26, npairs);
  long long *bsize = arena_array<long long>(A, 27, (size_t)npairs + 1);
  long long *boff = arena_array<long long>(A, 28, (size_t)npairs + 1);
  int *pair_block = arena_array<int>(A, 29, npairs);
  const int lo_bits = std::max(B.lo_bits, 1), hi_bits = std::max(B.hi_bits, 1);
  size_t tb[5] = {0, 0, 0, 0, 0};
  TG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb[0], P.t_lo, lo_sorted, P.t_iota, perm1,
                                            n, 0, lo_bits, s));
  TG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb[1], hi1, hi2, perm1, perm2, n, 0,
                                            hi_bits, s));
  TG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb[2], P.pair_key, bkey, P.pair_val, bval,
                                            npairs, 0, B.blk_bits + 1, s));
  TG_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb[3], head, bid, npairs, s));
  TG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb[4], bsize, boff, npairs + 1, s));
  const size_t tmax = *std::max_element(tb, tb + 5);
  void *tmp = A.get(12, tmax);
  size_t tbx;

  // stable sort: (iset, jset, ipgf, jpgf), then (level, image, iatom, jatom)
  tbx = tmax;
  TG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tbx, P.t_lo, lo_sorted, P.t_iota, perm1, n, 0,
                                            lo_bits, s));
  hipLaunchKernelGGL(gather_hi_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s,
                     (const unsigned long long *)P.t_hi, (const int *)perm1, hi1, (long long)n);
  TG_HIP(hipGetLastError());
  tbx = tmax;
  TG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tbx, hi1, hi2, perm1, perm2, n, 0, hi_bits, s));
  // blocks
  tbx = tmax;
  TG_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tbx, P.pair_key, bkey, P.pair_val, bval, npairs,
                                            0, B.blk_bits + 1, s));
  hipLaunchKernelGGL(block_heads_kernel, dim3(grid_for(npairs, 256)), dim3(256), 0, s,
                     (const unsigned long long *)bkey, npairs, B.blk_bits, head);
  TG_HIP(hipGetLastError());
  tbx = tmax;
  TG_HIP(hipcub::DeviceScan::InclusiveSum(tmp, tbx, head, bid, npairs, s));
  hipLaunchKernelGGL(zero_kernel, dim3(grid_for(npairs + 1, 256)), dim3(256), 0, s, bsize,
                     (long long)npairs + 1);
  TG_HIP(hipGetLastError());
  hipLaunchKernelGGL(block_number_kernel, dim3(grid_for(npairs, 256)), dim3(256), 0, s, P,
                     (const unsigned long long *)bkey, (const int *)bval, (const int *)head,
                     (const int *)bid, npairs, pair_block, bsize);
  TG_HIP(hipGetLastError());
  tbx = tmax;
  TG_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tbx, bsize, boff, npairs + 1, s));

  OutLists O;
  O.level = arena_array<int>(A, 30, n);
  O.iatom = arena_array<int>(A, 31, n);
  O.jatom = arena_array<int>(A, 32, n);
  O.iset = arena_array<int>(A, 33, n);
  O.jset = arena_array<int>(A, 34, n);
  O.ipgf = arena_array<int>(A, 35, n);
  O.jpgf = arena_array<int>(A, 36, n);
  O.block = arena_array<int>(A, 37, n);
  O.radius = arena_array<double>(A, 38, n);
  O.rab = arena_array<double>(A, 39, 3 * (size_t)n);
  hipLaunchKernelGGL(gather_tasks_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, P,
                     (const unsigned long long *)hi2, (const int *)perm2,
                     (const int *)pair_block, (long long)n, O);
  TG_HIP(hipGetLastError());
  TG_HIP(hipEventRecord(st.ev1, s));

  auto down = [&](void *dst, const void *src, size_t bytes) {
    TG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  };
  int nb = 0;
  down(&nb, bid + (npairs - 1), sizeof(int));
  down(level_list, O.level, sizeof(int) * n);
  down(iatom_list, O.iatom, sizeof(int) * n);
  down(jatom_list, O.jatom, sizeof(int) * n);
  down(iset_list, O.iset, sizeof(int) * n);
  down(jset_list, O.jset, sizeof(int) * n);
  down(ipgf_list, O.ipgf, sizeof(int) * n);
  down(jpgf_list, O.jpgf, sizeof(int) * n);
  down(block_num_list, O.block, sizeof(int) * n);
  down(radius_list, O.radius, sizeof(double) * n);
  down(rab_list, O.rab, sizeof(double) * 3 * (size_t)n);
  std::vector<long long> offs((size_t)npairs + 1);
  down(offs.data(), boff, sizeof(long long) * ((size_t)npairs + 1));
  TG_HIP(hipStreamSynchronize(s));
  TG_HIP(hipEventElapsedTime(&st.last_ms, st.ev0, st.ev1));
  if (offs[nb] >= (1LL << 31))
    tg_fatal("generate_task_list: the blocks hold %lld doubles (int offsets)", offs[nb]);
  for (int b = 0; b < nb; b++) block_offsets[b] = (int)offs[b];
  *nblocks = nb;
  *buffer_size = offs[nb];
  for (int i = 0; i < n; i++) border_mask_list[i] = 0;
  return ntasks;
}

double grid_mi355x_generate_task_list_ms(void) { return gen_state().last_ms; }

}  // extern "C"
