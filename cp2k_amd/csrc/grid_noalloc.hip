// Scratch memory of the allocation-free collocate entry point
// (grid_mi355x_collocate_task_list_noalloc): a file-scope device array that
// holds, per call, the level table, the decontracted density blocks and the
// per-task cxyz coefficients.  Each device holds its own copy for the life
// of the process; its address on a device comes from hipGetSymbolAddress,
// taken once per device (the array's name is not a device address in host
// code).  Calls that overlap on one device share the array and race.

#include <hip/hip_runtime.h>

#include <mutex>

#include "grid_internal.h"

namespace gm {

constexpr long long kNoAllocScratchDoubles = 16LL << 20;  // 128 MiB
constexpr int kMaxDevices = 64;
constexpr int kLevelPack = 8;  // levels written per launch of write_levels_kernel

static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
std::mutex scratch_mu;
double *scratch_ptr[kMaxDevices] = {};

struct LevelPack {
  Level lv[kLevelPack];
  int n;
};

// The level table is written by a kernel from its arguments: no copy
// from host memory and nothing for the runtime to stage.
__global__ void write_levels_kernel(LevelPack p, Level *dst) {
  const int l = threadIdx.x;
  if (l < p.n) dst[l] = p.lv[l];
}
}  // namespace

// Device address of the scratch array on the current device and its size
// in doubles; nullptr on failure (hipGetSymbolAddress's error in *err).
double *noalloc_scratch(long long *doubles, hipError_t *err) {
  int dev = 0;
  *err = hipGetDevice(&dev);
  if (*err != hipSuccess) return nullptr;
  if (dev < 0 || dev >= kMaxDevices) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> g(scratch_mu);
  if (!scratch_ptr[dev]) {
    void *p = nullptr;
    *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (*err != hipSuccess) return nullptr;
    scratch_ptr[dev] = static_cast<double *>(p);
  }
  *doubles = kNoAllocScratchDoubles;
  *err = hipSuccess;
  return scratch_ptr[dev];
}

// Writes n levels from host memory h to device memory dst (in the scratch).
hipError_t launch_write_levels(const Level *h, int n, Level *dst, hipStream_t s) {
  for (int l0 = 0; l0 < n; l0 += kLevelPack) {
    LevelPack p{};
    p.n = n - l0 < kLevelPack ? n - l0 : kLevelPack;
    for (int l = 0; l < p.n; l++) p.lv[l] = h[l0 + l];
    hipLaunchKernelGGL(write_levels_kernel, dim3(1), dim3(kLevelPack), 0, s, p, dst + l0);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace gm
