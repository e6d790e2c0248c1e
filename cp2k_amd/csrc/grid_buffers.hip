// Packing of matrix blocks into the grid calls' pab buffer and unpacking of
// the hab buffer into matrix blocks, on the device: the work of
// rs_pack_buffer / rs_unpack_buffer (src/task_list_methods.F:2597-2674) for
// HBM-resident blocks.  Each atom pair's block is a contiguous column-major
// run of doubles at pair_offsets[i] in the buffer.
//
// HBM-bound copies.  One wave per atom pair: its 64 lanes copy the pair's
// doubles in coalesced runs (blocks hold tens to thousands of doubles).

#include <hip/hip_runtime.h>

namespace gm {
namespace {

constexpr int kWaves = 4;  // waves (pairs) per workgroup

__global__ __launch_bounds__(256) void pack_kernel(const double *const *__restrict__ blocks,
                                                   const int *__restrict__ sizes,
                                                   const long long *__restrict__ offsets,
                                                   int npairs, double *__restrict__ dest) {
  const int p = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (p >= npairs) return;
  const double *src = blocks[p];
  double *dst = dest + offsets[p];
  const int n = sizes[p];
  for (int e = threadIdx.x & 63; e < n; e += 64) dst[e] = src[e];
}

// Blocks of distinct atom pairs are distinct memory, so the adds need no
// atomics (the reference locks per matrix row only for DBCSR's bookkeeping).
__global__ __launch_bounds__(256) void unpack_kernel(double *const *__restrict__ blocks,
                                                     const int *__restrict__ sizes,
                                                     const long long *__restrict__ offsets,
                                                     int npairs, const double *__restrict__ src) {
  const int p = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (p >= npairs) return;
  double *dst = blocks[p];
  const double *s = src + offsets[p];
  const int n = sizes[p];
  for (int e = threadIdx.x & 63; e < n; e += 64) dst[e] += s[e];
}

}  // namespace

hipError_t launch_pack(const double *const *blocks, const int *sizes, const long long *offsets,
                       int npairs, double *dest, hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  const int groups = (npairs + kWaves - 1) / kWaves;
  hipLaunchKernelGGL(pack_kernel, dim3(groups), dim3(256), 0, s, blocks, sizes, offsets, npairs,
                     dest);
  return hipGetLastError();
}

hipError_t launch_unpack(double *const *blocks, const int *sizes, const long long *offsets,
                         int npairs, const double *src, hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  const int groups = (npairs + kWaves - 1) / kWaves;
  hipLaunchKernelGGL(unpack_kernel, dim3(groups), dim3(256), 0, s, blocks, sizes, offsets,
                     npairs, src);
  return hipGetLastError();
}

}  // namespace gm
