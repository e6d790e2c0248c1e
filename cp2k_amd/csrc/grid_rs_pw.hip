// Real-space multigrids <-> plane-wave grids on the device: the step after
// collocation and the step before integration, for one process (one GPU)
// and the Fourier interpolator (pw_interp), the default of CP2K.
//
//   density_rs2pw     src/pw_env/rs_pw_interface.F:64-128: each level's grid
//                     (transfer_rs2pw) is Fourier transformed (forward, 1/N),
//                     its coefficients are added into the finest level's
//                     G space by Miller index (pw_axpy through the grids'
//                     reference mapping), and the sum is transformed back
//                     (unnormalised) to the finest real-space grid.
//   potential_pw2rs   :139-226: the potential on the finest grid is
//                     transformed forward (1/N), its coefficients are copied
//                     into each coarser level's G space by Miller index,
//                     transformed back and scaled by the ratio of the grid
//                     volume elements; the finest level is a copy.
//   transfer_rs2pw /  src/pw/realspace_grid_types.F:651-769 (the serial case):
//   transfer_pw2rs    the real region of a grid with a halo of border points
//                     per side, and its periodic fill.
//
// Miller indices of a level with n points along an axis are
// -floor(n/2) .. -floor(n/2) + n - 1 (pw_grid_bounds_from_n,
// src/pw/pw_grid_info.F:247-254); FFT index k holds h = k for
// k <= n - 1 - floor(n/2), else h = k - n.  Real data go through complex
// transforms (full G space); the real-space result is the real part of the
// backward transform (CP2K's copy from complex to real).  For odd grid
// sizes, CP2K's default multigrids, this is the reference's result; for even
// sizes the coarse levels' Nyquist planes follow the same full-space rule.
//
// Layout: every grid [nz][ny][nx], x fastest (as the grid calls); G space
// the same box of complex values (re, im interleaved) in FFT index order.
// The transforms are hipFFT double-complex plans, cached per grid shape; the
// scatter / gather between levels and the real <-> complex copies are the
// HIP kernels below (HBM-bound, one element per lane).

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/grid_rs_pw.h"

namespace gm {
hipStream_t library_stream();
}

namespace {

[[noreturn]] void rs_fatal(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "grid_mi355x error: ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  va_end(ap);
  fflush(stderr);
  abort();
}

#define RS_HIP(call)                                                                       \
  do {                                                                                     \
    const hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                                  \
      rs_fatal("%s failed at %s:%d: %s", #call, __FILE__, __LINE__, hipGetErrorString(e_)); \
  } while (0)

#define RS_FFT(call)                                                                \
  do {                                                                              \
    const hipfftResult r_ = (call);                                                 \
    if (r_ != HIPFFT_SUCCESS) rs_fatal("%s failed at %s:%d: hipfft status %d", #call, \
                                       __FILE__, __LINE__, (int)r_);                \
  } while (0)

struct Dims {
  int x, y, z;
  __host__ __device__ long long n() const { return (long long)x * y * z; }
};

// Miller index of FFT index k on an axis of n points.
__device__ __forceinline__ int miller(int k, int n) { return k <= n - 1 - n / 2 ? k : k - n; }

// FFT index of Miller index h on an axis of n points (|h| <= n / 2).
__device__ __forceinline__ int fft_index(int h, int n) { return h >= 0 ? h : h + n; }

__global__ void real_to_complex_kernel(const double *__restrict__ src, double2 *__restrict__ dst,
                                       long long n, double scale) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = make_double2(scale * src[i], 0.0);
}

__global__ void real_part_kernel(const double2 *__restrict__ src, double *__restrict__ dst,
                                 long long n, double scale) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = scale * src[i].x;
}

__global__ void zero_complex_kernel(double2 *__restrict__ p, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = make_double2(0.0, 0.0);
}

// fine[map(k)] += scale * coarse[k] over the coarse box (the map is
// one-to-one: a coarse axis has at most as many points as the fine one).
__global__ void embed_add_kernel(const double2 *__restrict__ coarse, Dims c,
                                 double2 *__restrict__ fine, Dims f, double scale) {
  const long long n = c.n();
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int kx = (int)(i % c.x), ky = (int)((i / c.x) % c.y), kz = (int)(i / ((long long)c.x * c.y));
    const int fx = fft_index(miller(kx, c.x), f.x);
    const int fy = fft_index(miller(ky, c.y), f.y);
    const int fz = fft_index(miller(kz, c.z), f.z);
    double2 &d = fine[((long long)fz * f.y + fy) * f.x + fx];
    const double2 v = coarse[i];
    d.x += scale * v.x;
    d.y += scale * v.y;
  }
}

// coarse[k] = scale * fine[map(k)] over the coarse box.
__global__ void gather_kernel(const double2 *__restrict__ fine, Dims f,
                              double2 *__restrict__ coarse, Dims c, double scale) {
  const long long n = c.n();
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int kx = (int)(i % c.x), ky = (int)((i / c.x) % c.y), kz = (int)(i / ((long long)c.x * c.y));
    const int fx = fft_index(miller(kx, c.x), f.x);
    const int fy = fft_index(miller(ky, c.y), f.y);
    const int fz = fft_index(miller(kz, c.z), f.z);
    const double2 v = fine[((long long)fz * f.y + fy) * f.x + fx];
    coarse[i] = make_double2(scale * v.x, scale * v.y);
  }
}

// pw[k][j][i] = rs[k + b][j + b][i + b] (the real region of the halo grid).
__global__ void real_region_kernel(const double *__restrict__ rs, Dims r, double *__restrict__ pw,
                                   Dims p, Dims b) {
  const long long n = p.n();
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
    const int i = (int)(e % p.x), j = (int)((e / p.x) % p.y), k = (int)(e / ((long long)p.x * p.y));
    pw[e] = rs[((long long)(k + b.z) * r.y + (j + b.y)) * r.x + (i + b.x)];
  }
}

// rs[k][j][i] = pw[(k - b) mod n][(j - b) mod n][(i - b) mod n] (halo filled
// periodically, src/pw/realspace_grid_types.F:703-737).
__global__ void periodic_fill_kernel(const double *__restrict__ pw, Dims p, double *__restrict__ rs,
                                     Dims r, Dims b) {
  const long long n = r.n();
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
    const int i = (int)(e % r.x), j = (int)((e / r.x) % r.y), k = (int)(e / ((long long)r.x * r.y));
    int im = i - b.x, jm = j - b.y, km = k - b.z;
    im += im < 0 ? p.x : (im >= p.x ? -p.x : 0);
    jm += jm < 0 ? p.y : (jm >= p.y ? -p.y : 0);
    km += km < 0 ? p.z : (km >= p.z ? -p.z : 0);
    rs[e] = pw[((long long)km * p.y + jm) * p.x + im];
  }
}

unsigned grid_blocks(long long n) {
  const long long b = (n + 255) / 256;
  return (unsigned)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

// Device buffer that grows on demand.
struct Buf {
  void *p = nullptr;
  size_t bytes = 0;
  void *get(size_t n) {
    if (n > bytes) {
      if (p) RS_HIP(hipFree(p));
      p = nullptr;
      RS_HIP(hipMalloc(&p, n));
      bytes = n;
    }
    return p;
  }
};

struct RsPwState {
  std::mutex mu;
  std::map<std::tuple<int, int, int>, hipfftHandle> plans;
  Buf fine, level;
};

RsPwState &state() {
  static RsPwState *st = new RsPwState();  // lives until the process ends
  return *st;
}

hipfftHandle plan_of(RsPwState &st, const Dims &d, hipStream_t s) {
  const auto key = std::make_tuple(d.x, d.y, d.z);
  auto it = st.plans.find(key);
  if (it == st.plans.end()) {
    hipfftHandle h;
    // hipfftPlan3d takes the slowest dimension first: [z][y][x] row-major
    RS_FFT(hipfftPlan3d(&h, d.z, d.y, d.x, HIPFFT_Z2Z));
    it = st.plans.emplace(key, h).first;
  }
  RS_FFT(hipfftSetStream(it->second, s));
  return it->second;
}

Dims dims_of(const int n[3], const char *what, int level) {
  if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0) rs_fatal("%s: level %d has a non-positive size", what, level + 1);
  return Dims{n[0], n[1], n[2]};
}

void check_levels(int nlevels, const int (*npts)[3], const char *what) {
  if (nlevels < 1) rs_fatal("%s: needs at least one level", what);
  const Dims f = dims_of(npts[0], what, 0);
  for (int l = 1; l < nlevels; l++) {
    const Dims c = dims_of(npts[l], what, l);
    if (c.x > f.x || c.y > f.y || c.z > f.z)
      rs_fatal("%s: level %d (%d x %d x %d) is larger than the finest level 1", what, l + 1, c.x,
               c.y, c.z);
  }
}

}  // namespace

extern "C" {

void grid_mi355x_density_rs2pw(const int nlevels, const int (*npts)[3],
                               const double *const *rs_grids, double *rho_r, double *rho_g) {
  check_levels(nlevels, npts, "density_rs2pw");
  RsPwState &st = state();
  std::lock_guard<std::mutex> g(st.mu);
  const hipStream_t s = gm::library_stream();
  const Dims f = dims_of(npts[0], "density_rs2pw", 0);
  double2 *G = (double2 *)st.fine.get(sizeof(double2) * (size_t)f.n());
  if (nlevels == 1) {
    // one level: rho is the grid itself, rho_g its transform (:100-105)
    RS_HIP(hipMemcpyAsync(rho_r, rs_grids[0], sizeof(double) * (size_t)f.n(),
                          hipMemcpyDeviceToDevice, s));
    if (rho_g) {
      hipLaunchKernelGGL(real_to_complex_kernel, dim3(grid_blocks(f.n())), dim3(256), 0, s,
                         rs_grids[0], G, f.n(), 1.0 / (double)f.n());
      RS_HIP(hipGetLastError());
      RS_FFT(hipfftExecZ2Z(plan_of(st, f, s), (hipfftDoubleComplex *)G, (hipfftDoubleComplex *)G,
                           HIPFFT_FORWARD));
      RS_HIP(hipMemcpyAsync(rho_g, G, sizeof(double2) * (size_t)f.n(), hipMemcpyDeviceToDevice, s));
    }
    RS_HIP(hipStreamSynchronize(s));
    return;
  }
  hipLaunchKernelGGL(zero_complex_kernel, dim3(grid_blocks(f.n())), dim3(256), 0, s, G, f.n());
  RS_HIP(hipGetLastError());
  for (int l = 0; l < nlevels; l++) {
    const Dims c = dims_of(npts[l], "density_rs2pw", l);
    double2 *W = (double2 *)st.level.get(sizeof(double2) * (size_t)c.n());
    hipLaunchKernelGGL(real_to_complex_kernel, dim3(grid_blocks(c.n())), dim3(256), 0, s,
                       rs_grids[l], W, c.n(), 1.0);
    RS_HIP(hipGetLastError());
    RS_FFT(hipfftExecZ2Z(plan_of(st, c, s), (hipfftDoubleComplex *)W, (hipfftDoubleComplex *)W,
                         HIPFFT_FORWARD));
    hipLaunchKernelGGL(embed_add_kernel, dim3(grid_blocks(c.n())), dim3(256), 0, s, W, c, G, f,
                       1.0 / (double)c.n());
    RS_HIP(hipGetLastError());
  }
  if (rho_g)
    RS_HIP(hipMemcpyAsync(rho_g, G, sizeof(double2) * (size_t)f.n(), hipMemcpyDeviceToDevice, s));
  RS_FFT(hipfftExecZ2Z(plan_of(st, f, s), (hipfftDoubleComplex *)G, (hipfftDoubleComplex *)G,
                       HIPFFT_BACKWARD));
  hipLaunchKernelGGL(real_part_kernel, dim3(grid_blocks(f.n())), dim3(256), 0, s, G, rho_r, f.n(),
                     1.0);
  RS_HIP(hipGetLastError());
  RS_HIP(hipStreamSynchronize(s));
}

void grid_mi355x_potential_pw2rs(const int nlevels, const int (*npts)[3], const double *v_r,
                                 double *const *rs_grids) {
  check_levels(nlevels, npts, "potential_pw2rs");
  RsPwState &st = state();
  std::lock_guard<std::mutex> g(st.mu);
  const hipStream_t s = gm::library_stream();
  const Dims f = dims_of(npts[0], "potential_pw2rs", 0);
  RS_HIP(hipMemcpyAsync(rs_grids[0], v_r, sizeof(double) * (size_t)f.n(), hipMemcpyDeviceToDevice, s));
  if (nlevels > 1) {
    double2 *G = (double2 *)st.fine.get(sizeof(double2) * (size_t)f.n());
    hipLaunchKernelGGL(real_to_complex_kernel, dim3(grid_blocks(f.n())), dim3(256), 0, s, v_r, G,
                       f.n(), 1.0);
    RS_HIP(hipGetLastError());
    RS_FFT(hipfftExecZ2Z(plan_of(st, f, s), (hipfftDoubleComplex *)G, (hipfftDoubleComplex *)G,
                         HIPFFT_FORWARD));
    for (int l = 1; l < nlevels; l++) {
      const Dims c = dims_of(npts[l], "potential_pw2rs", l);
      double2 *W = (double2 *)st.level.get(sizeof(double2) * (size_t)c.n());
      hipLaunchKernelGGL(gather_kernel, dim3(grid_blocks(c.n())), dim3(256), 0, s, G, f, W, c,
                         1.0 / (double)f.n());
      RS_HIP(hipGetLastError());
      RS_FFT(hipfftExecZ2Z(plan_of(st, c, s), (hipfftDoubleComplex *)W, (hipfftDoubleComplex *)W,
                           HIPFFT_BACKWARD));
      // the volume element ratio dvol_l / dvol_1 = N_1 / N_l (:188-194)
      hipLaunchKernelGGL(real_part_kernel, dim3(grid_blocks(c.n())), dim3(256), 0, s, W,
                         rs_grids[l], c.n(), (double)f.n() / (double)c.n());
      RS_HIP(hipGetLastError());
    }
  }
  RS_HIP(hipStreamSynchronize(s));
}

void grid_mi355x_transfer_rs2pw(const int npts[3], const int border[3], const double *rs,
                                double *pw) {
  const Dims p = dims_of(npts, "transfer_rs2pw", 0);
  if (border[0] < 0 || border[1] < 0 || border[2] < 0) rs_fatal("transfer_rs2pw: negative border");
  const Dims b{border[0], border[1], border[2]};
  const Dims r{p.x + 2 * b.x, p.y + 2 * b.y, p.z + 2 * b.z};
  const hipStream_t s = gm::library_stream();
  hipLaunchKernelGGL(real_region_kernel, dim3(grid_blocks(p.n())), dim3(256), 0, s, rs, r, pw, p, b);
  RS_HIP(hipGetLastError());
  RS_HIP(hipStreamSynchronize(s));
}

void grid_mi355x_transfer_pw2rs(const int npts[3], const int border[3], const double *pw,
                                double *rs) {
  const Dims p = dims_of(npts, "transfer_pw2rs", 0);
  if (border[0] < 0 || border[1] < 0 || border[2] < 0) rs_fatal("transfer_pw2rs: negative border");
  if (border[0] > p.x || border[1] > p.y || border[2] > p.z)
    rs_fatal("transfer_pw2rs: a border wider than the grid");
  const Dims b{border[0], border[1], border[2]};
  const Dims r{p.x + 2 * b.x, p.y + 2 * b.y, p.z + 2 * b.z};
  const hipStream_t s = gm::library_stream();
  hipLaunchKernelGGL(periodic_fill_kernel, dim3(grid_blocks(r.n())), dim3(256), 0, s, pw, p, rs, r, b);
  RS_HIP(hipGetLastError());
  RS_HIP(hipStreamSynchronize(s));
}

}  // extern "C"
