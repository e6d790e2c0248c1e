// Host side of the MI355X grid backend: the C-ABI of src/grid/grid_task_list.h,
// src/grid/common/grid_library.h, src/grid/common/grid_basis_set.h and
// src/offload/offload_buffer.h, plus the backend-level grid_gpu_* symbols and
// the extensions in include/grid_mi355x.h, and the single-task entry points of
// include/grid_pgf_product.h.
//
// Compiled with -ffp-contract=off: the task geometry computed here (cube
// centres, discretised radii, index boxes) must round exactly like the
// reference's, because it decides which grid points belong to a task.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <array>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/grid_buffers.h"
#include "../../include/grid_library.h"
#include "../../include/grid_mi355x.h"
#include "../../include/grid_pgf_product.h"
#include "../../include/grid_task_list.h"
#include "grid_internal.h"

namespace gm {
hipError_t launch_collocate(const PointParams &P, bool ortho, hipStream_t s);
hipError_t launch_integrate(const PointParams &P, bool ortho, hipStream_t s);
hipError_t launch_decontract(const Group *groups, int ngroups, const double *pab,
                             long long pab_len, const double *sphi, double *cart,
                             long long cart_len, int cach, int max_nsb, hipStream_t s);
hipError_t launch_decontract_staged(const Group *groups, int ngroups, const double *pab,
                                    const double *sphi, double *cart, long long cart_len, int nc,
                                    int wave_doubles, int nwaves, int blocks, bool pf,
                                    hipStream_t s);
hipError_t launch_contract_staged(const Group *groups, int ngroups, const double *hab_cart,
                                  const double *sphi, double *hab, long long hab_len, int nc,
                                  int wave_doubles, int nwaves, int blocks, bool pf,
                                  hipStream_t s);
int stage_regs();
hipError_t launch_contract(const Group *groups, int ngroups, const double *hab_cart,
                           const double *sphi, double *hab, long long hab_len, int cach,
                           int max_nsb, int max_nsa_nsb, hipStream_t s);
hipError_t launch_fill(double *p, long long n, double v, hipStream_t s);
hipError_t launch_to_float(const double *src, float *dst, long long n, hipStream_t s);
hipError_t launch_reduce_forces(const int *ptr, const int *ent, const double *task_out,
                                double *forces, int natoms, hipStream_t s);
hipError_t launch_reduce_virial_lists(const int *ptr, const int *ent, const double *task_out,
                                     double *virial, int nlists, hipStream_t s);
hipError_t launch_reduce_virial(const double *task_out, long long ntasks, double *virial,
                                hipStream_t s);
hipError_t launch_cxyz_wave(const PointParams &P, hipStream_t s);
hipError_t launch_cxyz_lane(const PointParams &P, hipStream_t s);
hipError_t launch_collocate_wave(const PointParams &P, int lpb, hipStream_t s);
hipError_t launch_collocate_tiles(const PointParams &P, int lpb, int nunits, hipStream_t s);
hipError_t launch_integrate_wave(const PointParams &P, int lpb, hipStream_t s);
hipError_t launch_finish_wave(const PointParams &P, hipStream_t s);
bool finish_lane_ok(const PointParams &P, int lp);
hipError_t launch_finish_lane(const PointParams &P, hipStream_t s);
hipError_t launch_integrate_mfma(const PointParams &P, hipStream_t s);
hipError_t launch_finish_block(const PointParams &P, hipStream_t s);
hipError_t launch_general_collocate(const PointParams &P, int lpb, hipStream_t s);
hipError_t launch_general_cijk(const PointParams &P, hipStream_t s);
hipError_t launch_general_cijk_lane(const PointParams &P, hipStream_t s);
hipError_t launch_general_integrate(const PointParams &P, int lo, int hi, hipStream_t s);
hipError_t launch_general_tiles(const PointParams &P, int lpb, int nunits, hipStream_t s);
double *noalloc_scratch(long long *doubles, hipError_t *err);
hipError_t launch_write_levels(const Level *h, int n, Level *dst, hipStream_t s);
hipError_t launch_pack(const double *const *blocks, const int *sizes, const long long *offsets,
                       int npairs, double *dest, hipStream_t s);
hipError_t launch_unpack(double *const *blocks, const int *sizes, const long long *offsets,
                         int npairs, const double *src, hipStream_t s);
}  // namespace gm

using namespace gm;

// ---------------------------------------------------------------------------
// Errors: the reference has no return codes and aborts (src/grid/grid_task_list.c
// :124-133, src/offload/offload_runtime.h:66-79); so does this library.
// ---------------------------------------------------------------------------
[[noreturn]] static void fatal(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "grid_mi355x error: ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  va_end(ap);
  fflush(stderr);
  abort();
}

#define HIP_CHECK(call)                                                        \
  do {                                                                         \
    const hipError_t err_ = (call);                                            \
    if (err_ != hipSuccess)                                                    \
      fatal("%s failed at %s:%d: %s", #call, __FILE__, __LINE__,               \
            hipGetErrorString(err_));                                          \
  } while (0)

static inline int ncoset(int l) { return l < 0 ? 0 : (l + 1) * (l + 2) * (l + 3) / 6; }
static inline int pmod(int a, int m) { return ((a % m) + m) % m; }

// ---------------------------------------------------------------------------
// Library state (src/grid/common/grid_library.c)
// ---------------------------------------------------------------------------
namespace {
constexpr int kNumKernelIds = 15;
constexpr int kCounterLp = 20;

struct PendingTiming {
  int id;
  hipEvent_t a, b;
};

struct LibState {
  std::mutex mu;
  bool initialized = false;
  grid_library_config config = {GRID_BACKEND_AUTO, false, false};
  hipStream_t own_stream = nullptr;
  hipStream_t user_stream = nullptr;
  bool stream_set = false;
  int precision = -1;
  bool capturing = false;  // a batch graph is being captured: no timing events
  bool noalloc = false;    // inside the allocation-free entry point: no timing events
  int wave_kernels = -1;  // wave-per-task kernels for orthorhombic tasks
  bool timing = false;
  double kernel_ms[kNumKernelIds] = {};
  long kernel_n[kNumKernelIds] = {};
  std::vector<PendingTiming> pending;
  std::vector<hipEvent_t> event_pool;
  long counters[4][kCounterLp] = {};
  long launches[kNumKernelIds][2] = {};  // eager point-kernel launches: [id][FP64, FP32]
} S;

}  // namespace

namespace gm {
// Called by the launchers of the point kernels: one launch of kernel id
// (the timing ids) with FP32 (true) or FP64 point evaluation.
void count_launch(int id, bool fp32) {
  if (id >= 0 && id < kNumKernelIds) S.launches[id][fp32 ? 1 : 0]++;
}
}  // namespace gm

namespace {
hipStream_t stream() {
  if (S.stream_set) return S.user_stream;
  if (!S.own_stream) HIP_CHECK(hipStreamCreateWithFlags(&S.own_stream, hipStreamNonBlocking));
  return S.own_stream;
}

}  // namespace

namespace gm {
// The stream all library work is launched on (the other sources launch there).
hipStream_t library_stream() { return stream(); }
}  // namespace gm

namespace {
int precision() {
  if (S.precision < 0) {
    const char *e = getenv("GRID_MI355X_PRECISION");
    S.precision = (e && atoi(e) == 1) ? GRID_MI355X_FP32_MIXED : GRID_MI355X_FP64;
  }
  return S.precision;
}

// GRID_MI355X_WAVE=0 routes every task to the per-task kernels (comparison
// runs); the default uses the wave-per-task kernels wherever they apply.
bool wave_kernels() {
  if (S.wave_kernels < 0) {
    const char *e = getenv("GRID_MI355X_WAVE");
    S.wave_kernels = (e && atoi(e) == 0) ? 0 : 1;
  }
  return S.wave_kernels == 1;
}

hipEvent_t get_event() {
  if (!S.event_pool.empty()) {
    hipEvent_t e = S.event_pool.back();
    S.event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  HIP_CHECK(hipEventCreate(&e));
  return e;
}

// Brackets one launch with events when timing is enabled.
struct TimedLaunch {
  int id;
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t s;
  TimedLaunch(int id_, hipStream_t s_) : id(id_), s(s_) {
    if (S.timing && !S.capturing && !S.noalloc && id >= 0) {
      a = get_event();
      b = get_event();
      HIP_CHECK(hipEventRecord(a, s));
    }
  }
  ~TimedLaunch() {
    if (a) {
      (void)hipEventRecord(b, s);
      S.pending.push_back({id, a, b});
    }
  }
};

void resolve_timings() {
  for (auto &p : S.pending) {
    HIP_CHECK(hipEventSynchronize(p.b));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
    S.kernel_ms[p.id] += ms;
    S.kernel_n[p.id] += 1;
    S.event_pool.push_back(p.a);
    S.event_pool.push_back(p.b);
  }
  S.pending.clear();
}

void sync_stream(hipStream_t s) {
  HIP_CHECK(hipStreamSynchronize(s));
  resolve_timings();
}

void ensure_device() {
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    fatal("no HIP device is visible (%s); the MI355X grid backend has no CPU path",
          e != hipSuccess ? hipGetErrorString(e) : "device count 0");
}

// Device allocation that grows on demand.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  void *get(size_t n) {
    if (n > bytes) {
      if (p) HIP_CHECK(hipFree(p));
      p = nullptr;
      bytes = 0;
      if (n) HIP_CHECK(hipMalloc(&p, n));
      bytes = n;
    }
    return p;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// l-range shifts of cab for a collocate func, src/grid/common/grid_prepare_pab.h:447-517.
bool func_ldiffs(int func, int d[4]) {
  int a, b;
  if (func == 100) {
    a = b = 0;
  } else if (func == 200 || (func >= 301 && func <= 303) || (func >= 501 && func <= 503) ||
             (func >= 601 && func <= 603) || (func >= 901 && func <= 906) ||
             (func >= 1001 && func <= 1003)) {
    a = b = 1;
  } else if (func >= 411 && func <= 433 && func % 10 >= 1 && func % 10 <= 3 &&
             (func / 10) % 10 >= 1 && (func / 10) % 10 <= 3) {
    d[0] = 1; d[1] = -1; d[2] = 2; d[3] = -1;
    return true;
  } else if ((func >= 701 && func <= 703) || (func >= 801 && func <= 803)) {
    a = b = 2;
  } else {
    return false;
  }
  d[0] = a; d[1] = -a; d[2] = b; d[3] = -b;
  return true;
}

// process_get_ldiffs, src/grid/common/grid_process_vab.h:222-251.
// Packed index of (a, b, c), a + b + c <= lp: a fastest, c slowest (the
// layout of the device's packed cxyz and cijk).
int packed_idx(int a, int b, int c, int lp) {
  int off = 0;
  for (int z = 0; z < c; z++) off += (lp - z + 1) * (lp - z + 2) / 2;
  for (int y = 0; y < b; y++) off += lp - c - y + 1;
  return off + a;
}

double factorial(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; i++) f *= i;
  return f;
}

// Appends M[o][i] and then M^T for the change of variables of a general
// cell: cijk[o] = sum_i M[o][i] cxyz[i] with (x - xp)^lx (y - yp)^ly
// (z - zp)^lz expanded in the lattice offsets (di, dj, dk) through
// r = di dh[0] + dj dh[1] + dk dh[2] (the multinomial expansion of
// src/grid/ref/grid_ref_collint.h:706-761).
void append_general_matrix(const double dh[3][3], int lp, std::vector<double> &out) {
  const int nc = (lp + 1) * (lp + 2) * (lp + 3) / 6;
  std::vector<double> M((size_t)nc * nc, 0.0);
  // (x - xp)^lx = (di dh00 + dj dh10 + dk dh20)^lx = sum lx! / (a! b! c!) ...
  for (int lz = 0; lz <= lp; lz++)
    for (int ly = 0; ly <= lp - lz; ly++)
      for (int lx = 0; lx <= lp - lz - ly; lx++) {
        const int in = packed_idx(lx, ly, lz, lp);
        const int L[3] = {lx, ly, lz};
        // split each Cartesian power over the three lattice directions
        std::vector<std::array<int, 3>> parts[3];
        for (int d = 0; d < 3; d++)
          for (int a = 0; a <= L[d]; a++)
            for (int b = 0; b <= L[d] - a; b++) parts[d].push_back({a, b, L[d] - a - b});
        for (const auto &px : parts[0])
          for (const auto &py : parts[1])
            for (const auto &pz : parts[2]) {
              const std::array<int, 3> *P3[3] = {&px, &py, &pz};
              double m = 1.0;
              for (int d = 0; d < 3; d++) {
                const auto &pd = *P3[d];
                m *= factorial(L[d]) / (factorial(pd[0]) * factorial(pd[1]) * factorial(pd[2]));
                for (int e = 0; e < 3; e++) m *= std::pow(dh[e][d], pd[e]);
              }
              const int il = px[0] + py[0] + pz[0], jl = px[1] + py[1] + pz[1],
                        kl = px[2] + py[2] + pz[2];
              M[(size_t)packed_idx(il, jl, kl, lp) * nc + in] += m;
            }
      }
  out.insert(out.end(), M.begin(), M.end());
  for (int i = 0; i < nc; i++)
    for (int o = 0; o < nc; o++) out.push_back(M[(size_t)o * nc + i]);
}

void integrate_ldiffs(bool forces, bool virial, bool tau, int d[4]) {
  d[0] = d[1] = d[2] = d[3] = 0;
  if (forces || virial) { d[0] += 1; d[1] -= 1; d[3] -= 1; }
  if (virial) { d[0] += 1; d[2] += 1; }
  if (tau) { d[0] += 1; d[2] += 1; d[1] -= 1; d[3] -= 1; }
}
}  // namespace

// ---------------------------------------------------------------------------
// Task list
// ---------------------------------------------------------------------------
namespace {
// Orthorhombic tasks of one level: [begin, end) in the sorted task array,
// sorted by lp0 = la_max + lb_max; lp_start[v] = first task with lp0 >= v.
constexpr int kLp0Slots = 12;
struct OrthoLevel {
  int begin = 0, end = 0;
  int max_ext = 0;      // largest cube extent 2 - 2 lo over the three axes
  bool classed = true;  // every task has a geometry class
  long long points = 0;  // task points of the level (classed levels)
  int lp_start[kLp0Slots] = {};
  // collocation tiles (classed levels): nx * by * bz points
  bool tiled = false;
  int by = 0, bz = 0;
  // launch shape of the tile kernels (grid_mi355x_autotune): workgroup
  // rounds per CU over the level's items, largest waves per workgroup
  int rounds = 4, max_waves = 12;
  // general tasks of the level: largest index box per axis and rows (j, k)
  int gni = 0, gnj = 0, gnk = 0;
  long long grows = 0;
};

// Items of one tile, sorted by lp0: start[v] = first item with lp0 >= v.
struct TileInfo {
  int level, gy0, gz0;
  int start[kLp0Slots + 1];
};

// Workgroups of the tile kernels for one lp shift: per level and lp bucket
// (lp <= 4, lp 5..6) a range of TileUnit records in `d`.
struct TilePlan {
  DevBuf d;
  std::vector<std::array<std::pair<int, int>, 3>> range;  // [level][bucket] = (first, count)
  std::pair<int, int> mrange[3];                          // [bucket]: all levels (merged)
};

// Launch class of a task in merged mode.
enum TaskClass : signed char { kClsOrthoClassed = 0, kClsOrthoOther = 1, kClsGeneral = 2, kClsSkip = 3 };

// Merged launches for one lp shift: task index lists of the launch kinds.
enum MergedList {
  kMSetup, kMOrthoA, kMOrthoB, kMGenA, kMGenB, kMOrthoRest, kMGenRest,
  kMDirectA, kMDirectB,  // classed orthorhombic tasks of levels without collocation tiles
  kMGenDirA, kMGenDirB,  // general tasks of levels without collocation tiles
  kMSetupHi, kMOrthoC, kMDirectC,  // collocation of classed orthorhombic tasks with lp 7..8
  kMOrthoRestC,          // orthorhombic tasks the per-task kernels collocate
  kMOrthoHl,             // integration of classed orthorhombic tasks with lp 7..11
  kMGenHl,               // integration of general tasks with lp 7..11
  kMGenRestI,            // general tasks the per-task kernels integrate
  kMNum
};
struct MergedPlan {
  DevBuf d;
  int off[kMNum + 1] = {};  // list m = d[off[m] .. off[m + 1])
  int lp[kMNum] = {};       // largest lp (with the shift) in list m
};

// Arguments of grid_create_task_list, kept for building batched lists.
struct ListInputs {
  bool ortho = false;
  int ntasks = 0, nlevels = 0, natoms = 0, nkinds = 0, nblocks = 0;
  std::vector<int> block_offsets, atom_kinds, level, iatom, jatom, iset, jset, ipgf, jpgf,
      border_mask, block_num, npts_global, npts_local, shift_local, border_width;
  std::vector<double> atom_positions, radius, rab, dh, dh_inv;
  std::vector<const grid_basis_set *> basis_sets;
  std::vector<char> level_ortho;  // per level (batched lists of mixed cells)
};

// Largest LDS tile in doubles; the rest of LDS holds the wave regions.
constexpr long long kTileDoubles = 14336;
// waves of a general tile workgroup (lp <= 4; 8 for lp 5..6)
constexpr int kGenTileWaves = 16;

// Orthorhombic row start, src/grid/ref/grid_ref_collint.h:147 / :49 (the
// device copy is ortho_start in grid_device.h).
inline int ortho_start_host(double remain, double hinv) {
  return (int)ceil(-1e-8 - sqrt(fmax(0.0, remain)) * hinv);
}

struct TaskListImpl {
  static constexpr unsigned kMagic = 0x6d693335u;
  unsigned magic = kMagic;
  bool empty = true;
  bool orthorhombic = false;
  int ntasks = 0, nlevels = 0, natoms = 0, nkinds = 0, nblocks = 0, ngroups = 0;
  std::vector<Layout> layouts;
  std::vector<int> npts_local;  // [nlevels][3]
  // device-resident description
  DevBuf d_tasks, d_groups, d_sphi, d_fptr, d_fent, d_cart, d_hab_cart, d_task_out;
  DevBuf d_geo, d_geo_cols, d_geo_kmin, d_cxyz;  // geometry classes, per-task cxyz
  DevBuf d_geo_rows, d_items;                      // row starts, collocation tile items
  DevBuf d_gridf;                                  // FP32-mixed: potential rounded to FP32
  DevBuf d_gmat, d_gmat_off;                       // general cells: cxyz -> cijk matrices
  std::vector<TileInfo> tiles;
  std::map<int, TilePlan> tile_plans;              // by lp shift
  std::vector<TileInfo> gtiles;                    // tiles of the general tasks
  std::map<int, TilePlan> gtile_plans;             // by lp shift
  DevBuf d_gitems;                                 // their items
  DevBuf d_queue;                                  // work-queue cursors (merged launches)
  // merged launches: the tile collocation and the point integration as
  // persistent launches over device work queues (set per batch, by the
  // batch autotuner or grid_mi355x_batch_set_work_queue)
  bool work_queue = false;
  // Per-call level tables.  A call returns once its work is enqueued, so the
  // table of a call still in flight must not be rewritten: each call takes
  // the next slot of a ring (pinned staging + device copy), and waits for
  // the copy of that slot's previous call before writing it.
  static constexpr int kLevelSlots = 8;
  struct LevelSlot {
    Level *h = nullptr;
    DevBuf d;
    hipEvent_t copied = nullptr;
  };
  LevelSlot level_slots[kLevelSlots];
  int next_level_slot = 0, last_level_slot = 0;
  long long cart_len = 0;
  // maxima for the LDS plan
  int max_la = 0, max_lb = 0, max_lp = 0, max_cmax = 0, max_rows_ortho = 0;
  int max_range = 0, max_rows_general = 0, max_planes = 0;
  bool has_ortho = false, has_general = false;
  int max_ncoa = 0, max_nsb = 0, max_nsa_nsb = 0, max_nsa = 0, max_ncob = 0;
  long long block_end = 0;  // max over groups of block offset + block size
  long n_ortho = 0, n_general = 0, n_skip = 0;
  long hist[2][kCounterLp] = {};  // [ortho?][lp of the basis sets]
  std::vector<OrthoLevel> olev;   // per level
  std::vector<OrthoLevel> glev;   // per level: general tasks (begin, end, lp_start)
  std::vector<signed char> cls;   // per sorted task: a TaskClass
  std::vector<int> task_level;    // per sorted task: grid level (zero based)
  int max_ext_all = 0;            // largest cube extent of the orthorhombic tasks
  // merged launches (batched lists): one launch per kind over all levels
  bool merged = false;
  std::map<int, MergedPlan> merged_plans;  // by lp shift
  const Level *static_levels = nullptr;    // fixed level table (graph capture)
  ListInputs inputs;                       // the create arguments (for batching)
  std::vector<signed char> lp0s;  // lp0 per sorted task
  std::vector<int> highl_before;  // sorted tasks before s outside lane_cxyz_task
  int ngeo = 0;
  long long class_points = 0;     // task points of the tasks with a geometry class
  std::vector<long long> class_npts;  // per sorted task: points of its class (or 0)
  long long points = -1;
  std::vector<long long> points_by_lp;  // task points per la_max + lb_max
  // per-call scratch when the caller's buffers have no device mirror
  DevBuf pab_dev, hab_dev, forces_dev;
  std::vector<DevBuf> grid_dev;

  ~TaskListImpl() {
    DevBuf *bufs[] = {&d_tasks, &d_groups, &d_sphi, &d_fptr, &d_fent,
                      &d_cart, &d_hab_cart, &d_task_out, &pab_dev, &hab_dev, &forces_dev,
                      &d_geo, &d_geo_cols, &d_geo_kmin, &d_cxyz, &d_geo_rows, &d_items,
                      &d_gridf, &d_gmat, &d_gmat_off};
    for (auto &kv : tile_plans) kv.second.d.release();
    for (auto &kv : merged_plans) kv.second.d.release();
    for (auto &kv : gtile_plans) kv.second.d.release();
    d_gitems.release();
    d_queue.release();
    for (DevBuf *b : bufs) b->release();
    for (auto &b : grid_dev) b.release();
    for (LevelSlot &ls : level_slots) {
      if (ls.copied) (void)hipEventSynchronize(ls.copied);
      if (ls.copied) (void)hipEventDestroy(ls.copied);
      if (ls.h) (void)hipHostFree(ls.h);
      ls.d.release();
    }
    magic = 0;
  }
};

TaskListImpl *checked(const void *p) {
  TaskListImpl *t = (TaskListImpl *)p;
  if (!t || t->magic != TaskListImpl::kMagic) fatal("invalid or freed task list handle");
  return t;
}

template <typename T>
void upload(DevBuf &b, const std::vector<T> &v) {
  const size_t n = v.size() * sizeof(T);
  void *p = b.get(n ? n : sizeof(T));
  if (n) HIP_CHECK(hipMemcpy(p, v.data(), n, hipMemcpyHostToDevice));
}

int lds_limit();
TilePlan &tile_plan(TaskListImpl *tl, int sh, bool general = false);

// Geometry of a task from ra, rab, the exponents and the radius (FP64, in the
// reference's operation order): kTaskSkip if the task maps to no point
// (2 radius < the largest |dh| entry), kTaskOrtho for the orthorhombic path
// (ortho_cxyz_to_grid) or 0 for the general path (general_cijk_to_grid); the
// record's flags are set accordingly.  `num` names the task in messages.
int task_geometry(Task &T, const Layout &y, bool ortho, int bm, int num) {
  // cab_to_grid, src/grid/ref/grid_ref_collint.h:928-947
  double hmax = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) hmax = fmax(hmax, fabs(y.dh[i][j]));
  T.zetp = T.zeta + T.zetb;
  const double f = T.zetb / T.zetp;
  const double rab2 = T.rab[0] * T.rab[0] + T.rab[1] * T.rab[1] + T.rab[2] * T.rab[2];
  T.pexp = exp(-T.zeta * f * rab2);
  for (int d = 0; d < 3; d++) {
    T.rp[d] = T.ra[d] + f * T.rab[d];
    T.rb[d] = T.ra[d] + T.rab[d];
  }
  if (2.0 * T.radius < hmax) {
    T.flags = kTaskSkip;
    return kTaskSkip;
  }
  if (ortho) {
    // ortho_cxyz_to_grid, grid_ref_collint.h:222-257
    T.flags = kTaskOrtho;
    for (int d = 0; d < 3; d++) {
      double s2 = 0.0;
      for (int e = 0; e < 3; e++) s2 += y.dh_inv[e][d] * T.rp[e];
      T.center[d] = (int)floor(s2);
    }
    for (int d = 0; d < 3; d++) T.roff[d] = T.rp[d] - ((double)T.center[d]) * y.dh[d][d];
    const double drmin = fmin(y.dh[0][0], fmin(y.dh[1][1], y.dh[2][2]));
    T.disr = drmin * fmax(1.0, ceil(T.radius / drmin));
    T.cmax = 0;
    for (int d = 0; d < 3; d++) {
      T.lo[d] = (int)ceil(-1e-8 - T.disr * y.dh_inv[d][d]);
      T.hi[d] = 1 - T.lo[d];
      if (y.npts_global[d] != y.npts_local[d]) {
        // non-periodic direction: the cube must fit without wrapping (:247-253)
        const int off = pmod(T.center[d] + T.lo[d] - y.shift_local[d], y.npts_global[d]) - T.lo[d];
        if (!(off + T.hi[d] < y.npts_local[d]) || !(off + T.lo[d] >= 0))
          fatal("task %d: cube does not fit into the non-periodic local grid", num);
      }
      T.cmax = std::max(T.cmax, T.hi[d]);
    }
    return kTaskOrtho;
  }
  // general_cijk_to_grid, grid_ref_collint.h:591-654
  T.flags = 0;
  for (int d = 0; d < 3; d++) {
    T.blo[d] = 0;
    T.bhi[d] = y.npts_local[d] - 1;
    if (bm & (1 << (2 * d))) T.blo[d] += y.border_width[d];
    if (bm & (1 << (2 * d + 1))) T.bhi[d] -= y.border_width[d];
    T.gp[d] = 0.0;
    for (int e = 0; e < 3; e++) T.gp[d] += y.dh_inv[e][d] * T.rp[e];
    T.lo[d] = 2147483647;
    T.hi[d] = -2147483647 - 1;
  }
  for (int si = -1; si <= 1; si++)
    for (int sj = -1; sj <= 1; sj++)
      for (int sk = -1; sk <= 1; sk++) {
        const double x = T.rp[0] + si * T.radius;
        const double yy = T.rp[1] + sj * T.radius;
        const double z = T.rp[2] + sk * T.radius;
        for (int d = 0; d < 3; d++) {
          const double r = y.dh_inv[0][d] * x + y.dh_inv[1][d] * yy + y.dh_inv[2][d] * z;
          T.lo[d] = std::min(T.lo[d], (int)floor(r));
          T.hi[d] = std::max(T.hi[d], (int)ceil(r));
        }
      }
  return 0;
}

void build_task_list(TaskListImpl *tl, bool orthorhombic, int ntasks, int nlevels,
                     int natoms, int nkinds, int nblocks, const int *block_offsets,
                     const double *atom_positions, const int *atom_kinds,
                     const grid_basis_set *const *basis_sets, const int *level_list,
                     const int *iatom_list, const int *jatom_list, const int *iset_list,
                     const int *jset_list, const int *ipgf_list, const int *jpgf_list,
                     const int *border_mask_list, const int *block_num_list,
                     const double *radius_list, const double *rab_list,
                     const int *npts_global, const int *npts_local, const int *shift_local,
                     const int *border_width, const double *dh, const double *dh_inv,
                     const char *level_ortho = nullptr) {
  const grid_library_config cfg = grid_library_get_config();
  if (cfg.backend != GRID_BACKEND_AUTO && cfg.backend != GRID_BACKEND_GPU)
    fatal("grid backend %d is not provided by this library (only AUTO/GPU, the "
          "MI355X backend)", (int)cfg.backend);
  if (cfg.validate)
    fatal("VALIDATE needs CP2K's ref backend, which this library does not contain");

  tl->orthorhombic = orthorhombic;
  tl->ntasks = ntasks;
  tl->nlevels = nlevels;
  tl->natoms = natoms;
  tl->nkinds = nkinds;
  tl->nblocks = nblocks;
  tl->npts_local.assign(npts_local, npts_local + 3 * (size_t)std::max(nlevels, 0));
  {
    ListInputs &in = tl->inputs;
    auto cp = [](auto &v, const auto *p, size_t n) { v.assign(p, p + (p ? n : 0)); };
    in.ortho = orthorhombic;
    in.ntasks = ntasks;
    in.nlevels = nlevels;
    in.natoms = natoms;
    in.nkinds = nkinds;
    in.nblocks = nblocks;
    const size_t nt = (size_t)std::max(ntasks, 0), nl = (size_t)std::max(nlevels, 0);
    cp(in.block_offsets, block_offsets, (size_t)std::max(nblocks, 0));
    cp(in.atom_positions, atom_positions, 3 * (size_t)std::max(natoms, 0));
    cp(in.atom_kinds, atom_kinds, (size_t)std::max(natoms, 0));
    cp(in.basis_sets, basis_sets, (size_t)std::max(nkinds, 0));
    cp(in.level, level_list, nt);
    cp(in.iatom, iatom_list, nt);
    cp(in.jatom, jatom_list, nt);
    cp(in.iset, iset_list, nt);
    cp(in.jset, jset_list, nt);
    cp(in.ipgf, ipgf_list, nt);
    cp(in.jpgf, jpgf_list, nt);
    cp(in.border_mask, border_mask_list, nt);
    cp(in.block_num, block_num_list, nt);
    cp(in.radius, radius_list, nt);
    cp(in.rab, rab_list, 3 * nt);
    cp(in.npts_global, npts_global, 3 * nl);
    cp(in.npts_local, npts_local, 3 * nl);
    cp(in.shift_local, shift_local, 3 * nl);
    cp(in.border_width, border_width, 3 * nl);
    cp(in.dh, dh, 9 * nl);
    cp(in.dh_inv, dh_inv, 9 * nl);
    if (level_ortho) in.level_ortho.assign(level_ortho, level_ortho + nl);
    else in.level_ortho.assign(nl, orthorhombic ? 1 : 0);
  }
  // src/grid/grid_task_list.c:65-70
  tl->empty = (nblocks == 0 || ntasks == 0 || nlevels == 0);
  if (tl->empty) return;
  ensure_device();

  tl->layouts.resize(nlevels);
  for (int l = 0; l < nlevels; l++) {
    Layout &y = tl->layouts[l];
    for (int i = 0; i < 3; i++) {
      y.npts_global[i] = npts_global[3 * l + i];
      y.npts_local[i] = npts_local[3 * l + i];
      y.shift_local[i] = shift_local[3 * l + i];
      y.border_width[i] = border_width[3 * l + i];
      if (y.npts_local[i] <= 0 || y.npts_global[i] <= 0)
        fatal("level %d: non-positive grid dimension", l + 1);
      for (int j = 0; j < 3; j++) {
        y.dh[i][j] = dh[9 * l + 3 * i + j];
        y.dh_inv[i][j] = dh_inv[9 * l + 3 * i + j];
      }
    }
    if ((long long)y.npts_local[0] * y.npts_local[1] * y.npts_local[2] >= (1LL << 31))
      fatal("level %d: local grids of 2^31 points or more are not supported", l + 1);
    // the general-cell integrate reads a z plane through a buffer resource
    // whose range check needs planes below 2^31 bytes
    if ((long long)y.npts_local[0] * y.npts_local[1] >= (1LL << 28))
      fatal("level %d: local grid planes of 2^28 points or more are not supported", l + 1);
  }

  // --- validate indices (the reference trusts its caller; a wrong index here
  //     would become an out-of-bounds access on the device) ---
  int max_set_l = 0;
  for (int k = 0; k < nkinds; k++) {
    const grid_basis_set *b = basis_sets[k];
    if (!b) fatal("basis set of kind %d is NULL", k + 1);
    for (int s = 0; s < b->nset; s++) {
      if (b->lmax[s] > 5 || b->lmin[s] < 0 || b->lmin[s] > b->lmax[s])
        fatal("kind %d set %d: unsupported l range %d..%d (max 5)", k + 1, s + 1,
              b->lmin[s], b->lmax[s]);
      if (b->npgf[s] * ncoset(b->lmax[s]) > b->maxco)
        fatal("kind %d set %d: npgf*ncoset(lmax) exceeds maxco", k + 1, s + 1);
      if (b->first_sgf[s] < 1 || b->first_sgf[s] - 1 + b->nsgf_set[s] > b->nsgf)
        fatal("kind %d set %d: spherical functions outside 1..nsgf", k + 1, s + 1);
      max_set_l = std::max(max_set_l, b->lmax[s]);
    }
  }
  for (int a = 0; a < natoms; a++)
    if (atom_kinds[a] < 1 || atom_kinds[a] > nkinds) fatal("atom %d: bad kind", a + 1);
  for (int t = 0; t < ntasks; t++) {
    const int lv = level_list[t], ia = iatom_list[t], ja = jatom_list[t];
    if (lv < 1 || lv > nlevels) fatal("task %d: level %d out of range", t + 1, lv);
    if (ia < 1 || ia > natoms || ja < 1 || ja > natoms) fatal("task %d: atom out of range", t + 1);
    if (block_num_list[t] < 1 || block_num_list[t] > nblocks)
      fatal("task %d: block %d out of range", t + 1, block_num_list[t]);
    const grid_basis_set *ib = basis_sets[atom_kinds[ia - 1] - 1];
    const grid_basis_set *jb = basis_sets[atom_kinds[ja - 1] - 1];
    if (iset_list[t] < 1 || iset_list[t] > ib->nset || jset_list[t] < 1 || jset_list[t] > jb->nset)
      fatal("task %d: set out of range", t + 1);
    if (ipgf_list[t] < 1 || ipgf_list[t] > ib->npgf[iset_list[t] - 1] || jpgf_list[t] < 1 ||
        jpgf_list[t] > jb->npgf[jset_list[t] - 1])
      fatal("task %d: pgf out of range", t + 1);
  }

  // --- order: kernel class (orthorhombic, general, too small to map), then
  //     level, la_max + lb_max (one polynomial degree bound per launch),
  //     la_max (uniform waves in the lane-per-task setup), block, iset, jset
  //     (src/grid/ref/grid_ref_task_list.c:26-37) ---
  std::vector<double> level_hmax(nlevels, 0.0);
  for (int l = 0; l < nlevels; l++)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        level_hmax[l] = fmax(level_hmax[l], fabs(tl->layouts[l].dh[i][j]));
  std::vector<char> cls(ntasks), lpk(ntasks), lak(ntasks);
  for (int t = 0; t < ntasks; t++) {
    lak[t] = (char)basis_sets[atom_kinds[iatom_list[t] - 1] - 1]->lmax[iset_list[t] - 1];
    lpk[t] = (char)(lak[t] +
                    basis_sets[atom_kinds[jatom_list[t] - 1] - 1]->lmax[jset_list[t] - 1]);
    if (2.0 * radius_list[t] < level_hmax[level_list[t] - 1])
      cls[t] = 2;
    else
      cls[t] = ((level_ortho ? level_ortho[level_list[t] - 1] != 0 : orthorhombic) &&
                border_mask_list[t] == 0) ? 0 : 1;  // the task's kernel class, as below
  }
  std::vector<int> order(ntasks);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (cls[a] != cls[b]) return cls[a] < cls[b];
    if (level_list[a] != level_list[b]) return level_list[a] < level_list[b];
    if (lpk[a] != lpk[b]) return lpk[a] < lpk[b];
    if (lak[a] != lak[b]) return lak[a] < lak[b];
    if (block_num_list[a] != block_num_list[b]) return block_num_list[a] < block_num_list[b];
    if (iset_list[a] != iset_list[b]) return iset_list[a] < iset_list[b];
    return jset_list[a] < jset_list[b];
  });

  // --- groups: one per (block, iset, jset) over all levels ---
  std::vector<long long> keys(ntasks);
  long long smax = 1;
  for (int k = 0; k < nkinds; k++) smax = std::max<long long>(smax, basis_sets[k]->nset);
  for (int t = 0; t < ntasks; t++)
    keys[t] = ((long long)(block_num_list[t] - 1) * smax + (iset_list[t] - 1)) * smax +
              (jset_list[t] - 1);
  std::vector<long long> ukeys(keys);
  std::sort(ukeys.begin(), ukeys.end());
  ukeys.erase(std::unique(ukeys.begin(), ukeys.end()), ukeys.end());
  tl->ngroups = (int)ukeys.size();
  std::vector<Group> groups(tl->ngroups);
  std::vector<int> group_of(ntasks);
  std::vector<char> group_init(tl->ngroups, 0);
  // sphi of all kinds, concatenated
  std::vector<long long> sphi_off(nkinds);
  std::vector<double> sphi_all;
  for (int k = 0; k < nkinds; k++) {
    const grid_basis_set *b = basis_sets[k];
    sphi_off[k] = (long long)sphi_all.size();
    sphi_all.insert(sphi_all.end(), b->sphi, b->sphi + (size_t)b->nsgf * b->maxco);
  }
  tl->max_ncoa = tl->max_nsb = tl->max_nsa_nsb = tl->max_nsa = tl->max_ncob = 0;
  tl->block_end = 0;
  for (int t = 0; t < ntasks; t++) {
    const int g = (int)(std::lower_bound(ukeys.begin(), ukeys.end(), keys[t]) - ukeys.begin());
    group_of[t] = g;
    const int ia = iatom_list[t] - 1, ja = jatom_list[t] - 1;
    if (group_init[g]) {
      if (groups[g].iatom != ia || groups[g].jatom != ja)
        fatal("block %d is used with two different atom pairs", block_num_list[t]);
      continue;
    }
    group_init[g] = 1;
    const int ik = atom_kinds[ia] - 1, jk = atom_kinds[ja] - 1;
    const grid_basis_set *ib = basis_sets[ik], *jb = basis_sets[jk];
    const int is = iset_list[t] - 1, js = jset_list[t] - 1;
    Group &G = groups[g];
    G.iatom = ia;
    G.jatom = ja;
    G.block_off = block_offsets[block_num_list[t] - 1];
    G.transpose = ia <= ja;
    G.ncoa = ib->npgf[is] * ncoset(ib->lmax[is]);
    G.ncob = jb->npgf[js] * ncoset(jb->lmax[js]);
    G.nsa = ib->nsgf_set[is];
    G.nsb = jb->nsgf_set[js];
    G.sa0 = ib->first_sgf[is] - 1;
    G.sb0 = jb->first_sgf[js] - 1;
    G.nsgfa = ib->nsgf;
    G.nsgfb = jb->nsgf;
    G.maxcoa = ib->maxco;
    G.maxcob = jb->maxco;
    G.sphi_a = sphi_off[ik] + (long long)G.sa0 * ib->maxco;
    G.sphi_b = sphi_off[jk] + (long long)G.sb0 * jb->maxco;
    if (G.block_off < 0) fatal("block %d has a negative offset", block_num_list[t]);
    tl->block_end = std::max(tl->block_end, (long long)G.block_off + (long long)G.nsgfa * G.nsgfb);
    tl->max_ncoa = std::max(tl->max_ncoa, G.ncoa);
    tl->max_nsb = std::max(tl->max_nsb, G.nsb);
    tl->max_nsa_nsb = std::max(tl->max_nsa_nsb, G.nsa * G.nsb);
    tl->max_nsa = std::max(tl->max_nsa, G.nsa);
    tl->max_ncob = std::max(tl->max_ncob, G.ncob);
  }
  long long cart = 0;
  for (auto &G : groups) {
    G.cart_off = cart;
    cart += (long long)G.ncoa * G.ncob;
  }
  tl->cart_len = cart;

  // --- per-task geometry (FP64, operation order of the reference) ---
  std::vector<Task> tasks(ntasks);
  tl->max_la = tl->max_lb = tl->max_lp = tl->max_cmax = tl->max_rows_ortho = 0;
  tl->max_range = tl->max_rows_general = tl->max_planes = 0;
  tl->has_ortho = tl->has_general = false;
  tl->n_ortho = tl->n_general = tl->n_skip = 0;
  memset(tl->hist, 0, sizeof(tl->hist));
  for (int s = 0; s < ntasks; s++) {
    const int t = order[s];
    Task &T = tasks[s];
    memset(&T, 0, sizeof(T));
    const int ia = iatom_list[t] - 1, ja = jatom_list[t] - 1;
    const grid_basis_set *ib = basis_sets[atom_kinds[ia] - 1];
    const grid_basis_set *jb = basis_sets[atom_kinds[ja] - 1];
    const int is = iset_list[t] - 1, js = jset_list[t] - 1;
    const int ip = ipgf_list[t] - 1, jp = jpgf_list[t] - 1;
    const int level = level_list[t] - 1;
    const Layout &y = tl->layouts[level];
    const int bm = border_mask_list[t];
    T.level = level;
    T.group = group_of[t];
    T.iatom = ia;
    T.jatom = ja;
    T.la_max = ib->lmax[is];
    T.la_min = ib->lmin[is];
    T.lb_max = jb->lmax[js];
    T.lb_min = jb->lmin[js];
    T.o1 = ip * ncoset(T.la_max);
    T.o2 = jp * ncoset(T.lb_max);
    T.zeta = ib->zet[is * ib->maxpgf + ip];
    T.zetb = jb->zet[js * jb->maxpgf + jp];
    T.rscale = (ia == ja) ? 1.0 : 2.0;
    T.radius = radius_list[t];
    for (int d = 0; d < 3; d++) {
      T.ra[d] = atom_positions[3 * ia + d];
      T.rab[d] = rab_list[3 * t + d];
    }
    const int lp0 = T.la_max + T.lb_max;
    tl->max_la = std::max(tl->max_la, T.la_max);
    tl->max_lb = std::max(tl->max_lb, T.lb_max);
    tl->max_lp = std::max(tl->max_lp, lp0);
    const int kind =
        task_geometry(T, y, (level_ortho ? level_ortho[level] != 0 : orthorhombic) && bm == 0, bm,
                      t + 1);
    if (kind == kTaskSkip) {
      tl->n_skip++;
      continue;
    }
    if (kind == kTaskOrtho) {
      tl->has_ortho = true;
      tl->max_cmax = std::max(tl->max_cmax, T.cmax);
      tl->max_rows_ortho = std::max(tl->max_rows_ortho, 2 - 2 * T.lo[1]);
      tl->max_planes = std::max(tl->max_planes, 2 - 2 * T.lo[2]);
      tl->n_ortho++;
      tl->hist[1][std::min(lp0, kCounterLp - 1)]++;
    } else {
      tl->has_general = true;
      for (int d = 0; d < 3; d++) tl->max_range = std::max(tl->max_range, T.hi[d] - T.lo[d] + 1);
      tl->max_rows_general = std::max(tl->max_rows_general, T.hi[1] - T.lo[1] + 1);
      tl->max_planes = std::max(tl->max_planes, T.hi[2] - T.lo[2] + 1);
      tl->n_general++;
      tl->hist[0][std::min(lp0, kCounterLp - 1)]++;
    }
  }
  (void)max_set_l;

  // --- geometry classes: orthorhombic tasks of one level with the same
  //     discretised radius share their truncation region (it does not depend
  //     on the centre).  Per class: the columns (i, j) of the central plane,
  //     j-major, with the first plane kl of each column (planes kl .. 1 - kl),
  //     from the bounds of src/grid/ref/grid_ref_collint.h:139-167 ---
  std::vector<GeoClass> geo;
  std::vector<int> geo_cols, geo_kmin, geo_rows;
  std::map<std::pair<int, unsigned long long>, int> class_of;
  tl->class_points = 0;
  const int kNoRow = std::numeric_limits<int>::max();
  for (int s = 0; s < ntasks; s++) {
    Task &T = tasks[s];
    T.geo = -1;
    if (!(T.flags & kTaskOrtho)) continue;
    if (2 - 2 * std::min({T.lo[0], T.lo[1], T.lo[2]}) > kWaveMaxExt) continue;
    unsigned long long bits;
    memcpy(&bits, &T.disr, sizeof bits);
    const auto key = std::make_pair(T.level, bits);
    auto it = class_of.find(key);
    if (it == class_of.end()) {
      const Layout &y = tl->layouts[T.level];
      GeoClass G{};
      G.col_off = (int)geo_cols.size();
      G.grp_off = (int)geo_kmin.size();
      G.row_off = (int)geo_rows.size();
      for (int d = 0; d < 3; d++) G.lo[d] = T.lo[d];
      const int nkz = 1 - T.lo[2], njy = 1 - T.lo[1];
      std::vector<int> i0((size_t)nkz * njy);
      for (int kz = 0; kz < nkz; kz++) {
        const double kr = (T.lo[2] + kz) * y.dh[2][2];
        const double krem = T.disr * T.disr - kr * kr;
        const int js = ortho_start_host(krem, y.dh_inv[1][1]);
        for (int jy = 0; jy < njy; jy++) {
          const int j1 = T.lo[1] + jy;
          int v = kNoRow;
          if (j1 >= js) {
            const double jr = j1 * y.dh[1][1];
            const double jrem = krem - jr * jr;
            v = ortho_start_host(jrem, y.dh_inv[0][0]);
          }
          i0[(size_t)kz * njy + jy] = v;
        }
      }
      for (int j = T.lo[1]; j <= 1 - T.lo[1]; j++) {
        geo_rows.push_back((int)geo_cols.size() - G.col_off);
        const int jy = (j <= 0 ? j : 1 - j) - T.lo[1];
        const int ic = i0[(size_t)(nkz - 1) * njy + jy];
        if (ic == kNoRow) continue;
        for (int i = ic; i <= 1 - ic; i++) {
          const int i1 = i <= 0 ? i : 1 - i;
          int kz = 0;
          while (i0[(size_t)kz * njy + jy] > i1) kz++;  // the central plane holds it
          for (int z = kz; z < nkz; z++)
            if (i0[(size_t)z * njy + jy] > i1)
              fatal("truncation region of level %d is not convex in z", T.level + 1);
          geo_cols.push_back((i - T.lo[0]) | ((j - T.lo[1]) << 8) | (kz << 16));
          G.npts += 2 - 2 * (T.lo[2] + kz);
        }
      }
      G.ncol = (int)geo_cols.size() - G.col_off;
      geo_rows.push_back(G.ncol);
      for (int c0 = 0; c0 < G.ncol; c0 += 64) {
        int kmin = 1;
        for (int c = c0; c < std::min(c0 + 64, G.ncol); c++)
          kmin = std::min(kmin, T.lo[2] + (geo_cols[G.col_off + c] >> 16));
        geo_kmin.push_back(kmin);
      }
      it = class_of.emplace(key, (int)geo.size()).first;
      geo.push_back(G);
    }
    T.geo = it->second;
  }
  tl->ngeo = (int)geo.size();
  tl->class_npts.assign(ntasks, 0);
  for (int s = 0; s < ntasks; s++)
    if (tasks[s].geo >= 0) tl->class_npts[s] = geo[tasks[s].geo].npts;
  tl->cls.resize(ntasks);
  tl->task_level.resize(ntasks);
  for (int s = 0; s < ntasks; s++) tl->task_level[s] = tasks[s].level;
  tl->max_ext_all = 0;
  for (int s = 0; s < ntasks; s++) {
    const Task &T = tasks[s];
    tl->cls[s] = (T.flags & kTaskSkip) ? kClsSkip
                 : (T.flags & kTaskOrtho) ? (T.geo >= 0 ? kClsOrthoClassed : kClsOrthoOther)
                                          : kClsGeneral;
    if (T.flags & kTaskOrtho)
      for (int d = 0; d < 3; d++) tl->max_ext_all = std::max(tl->max_ext_all, 2 - 2 * T.lo[d]);
  }

  // --- per-level ranges of the orthorhombic tasks ---
  tl->olev.assign(nlevels, OrthoLevel{});
  tl->lp0s.resize(ntasks);
  for (int s = 0; s < ntasks; s++) tl->lp0s[s] = (signed char)(tasks[s].la_max + tasks[s].lb_max);
  tl->highl_before.assign(ntasks + 1, 0);
  for (int s = 0; s < ntasks; s++)
    tl->highl_before[s + 1] =
        tl->highl_before[s] +
        (lane_cxyz_task(tasks[s].la_max, tasks[s].lb_max) ? 0 : 1);
  for (int s = 0; s < ntasks; s++) {
    const Task &T = tasks[s];
    if (!(T.flags & kTaskOrtho)) continue;
    OrthoLevel &o = tl->olev[T.level];
    if (o.end == 0) o.begin = s;
    o.end = s + 1;
    for (int d = 0; d < 3; d++) o.max_ext = std::max(o.max_ext, 2 - 2 * T.lo[d]);
    if (T.geo < 0)
      o.classed = false;
    else
      o.points += geo[T.geo].npts;
  }
  for (auto &o : tl->olev)
    if (o.end > o.begin && o.classed) tl->class_points += o.points;

  // --- collocation tiles: per classed level, tiles of full rows; each task
  //     becomes one item per tile its cube reaches (runs of rows j and planes
  //     k whose grid index stays in one tile without wrapping) ---
  std::vector<TileItem> items;
  tl->tiles.clear();
  for (auto &kv : tl->tile_plans) kv.second.d.release();
  tl->tile_plans.clear();
  for (auto &kv : tl->merged_plans) kv.second.d.release();
  tl->merged_plans.clear();
  for (auto &kv : tl->gtile_plans) kv.second.d.release();
  tl->gtile_plans.clear();
  for (int l = 0; l < nlevels; l++) {
    OrthoLevel &o = tl->olev[l];
    o.tiled = false;
    if (o.end <= o.begin || !o.classed) continue;
    const Layout &y = tl->layouts[l];
    const long long nx = y.npts_local[0], ny = y.npts_local[1], nz = y.npts_local[2];
    // the tile shares LDS with at least four wave regions of the largest
    // tables (lp <= 8, this level's cube extent)
    const long long wave_max = 8LL * (165 + 3LL * o.max_ext * 9) + 16;
    const long long cap = std::min(kTileDoubles, (lds_limit() - 4 * wave_max) / 8);
    if (nx * ny * nz <= cap) {
      o.by = (int)ny;
      o.bz = (int)nz;
    } else if (cap / (nx * nz) >= 4) {
      o.by = (int)(cap / (nx * nz));
      o.bz = (int)nz;
    } else if (nx * 8 * 2 <= cap) {
      o.by = (int)std::min<long long>(8, ny);
      o.bz = (int)std::min<long long>(nz, cap / (nx * o.by));
    } else {
      continue;  // rows too long for a tile: straight into HBM
    }
    o.tiled = true;
    const bool full_y = o.by == ny, full_z = o.bz == nz;
    const int nty = (int)((ny + o.by - 1) / o.by);
    struct Seg { int a, b, t; };
    auto segments = [&](int lo, int cw, int nglob, int nloc, int bsz, bool full, std::vector<Seg> &out) {
      out.clear();
      int prev = -2;
      for (int c = lo; c <= 1 - lo; c++) {
        const int g = pmod(cw + c, nglob);
        if (g >= nloc) fatal("level %d: cube leaves the local grid", l + 1);
        const int t = full ? 0 : g / bsz;
        if (out.empty() || t != out.back().t || (!full && g != prev + 1))
          out.push_back(Seg{c, c, t});
        else
          out.back().b = c;
        prev = g;
      }
    };
    struct Keyed { long long key; TileItem it; };
    std::vector<Keyed> lev_items;
    std::vector<Seg> ys, zs;
    for (int s2 = o.begin; s2 < o.end; s2++) {
      const Task &T = tasks[s2];
      segments(T.lo[1], pmod(T.center[1] - y.shift_local[1], y.npts_global[1]),
               y.npts_global[1], (int)ny, o.by, full_y, ys);
      segments(T.lo[2], pmod(T.center[2] - y.shift_local[2], y.npts_global[2]),
               y.npts_global[2], (int)nz, o.bz, full_z, zs);
      for (const Seg &a : ys)
        for (const Seg &b : zs) {
          TileItem it{};
          it.task = s2;
          it.j0 = (short)a.a;
          it.j1 = (short)a.b;
          it.k0 = (short)b.a;
          it.k1 = (short)b.b;
          const long long tile = (long long)b.t * nty + a.t;
          lev_items.push_back(Keyed{tile << 8 | tl->lp0s[s2], it});
        }
    }
    std::stable_sort(lev_items.begin(), lev_items.end(),
                     [](const Keyed &a, const Keyed &b) { return a.key < b.key; });
    for (size_t a = 0; a < lev_items.size();) {
      const long long tile = lev_items[a].key >> 8;
      size_t b = a;
      while (b < lev_items.size() && (lev_items[b].key >> 8) == tile) b++;
      TileInfo ti{};
      ti.level = l;
      ti.gy0 = (int)(tile % nty) * o.by;
      ti.gz0 = (int)(tile / nty) * o.bz;
      const int first = (int)items.size();
      for (size_t q = a; q < b; q++) items.push_back(lev_items[q].it);
      for (int v = 0; v <= kLp0Slots; v++) {
        int q = first;
        while (q < (int)items.size() && tl->lp0s[items[q].task] < v) q++;
        ti.start[v] = q;
      }
      tl->tiles.push_back(ti);
      a = b;
    }
  }
  // --- per-level ranges of the general tasks (sorted by level, then lp0) ---
  tl->glev.assign(nlevels, OrthoLevel{});
  for (int s = 0; s < ntasks; s++) {
    const Task &T = tasks[s];
    if ((T.flags & (kTaskOrtho | kTaskSkip)) != 0) continue;
    OrthoLevel &g = tl->glev[T.level];
    if (g.end == 0) g.begin = s;
    g.end = s + 1;
    const int ni = T.hi[0] - T.lo[0] + 1, nj = T.hi[1] - T.lo[1] + 1, nk = T.hi[2] - T.lo[2] + 1;
    g.gni = std::max(g.gni, ni);
    g.gnj = std::max(g.gnj, nj);
    g.gnk = std::max(g.gnk, nk);
    g.grows = std::max(g.grows, (long long)nj * nk);
    if (ni > 65535 || nj > 4096)
      fatal("task %d: index box %d x %d x %d is too large for the general kernels", s + 1, ni, nj, nk);
  }
  // --- collocation tiles of the general tasks: full rows of stride nx | 1;
  //     items are runs of rows j and planes k of a task's index box whose
  //     grid index stays inside the border-masked grid and one tile ---
  std::vector<TileItem> gitems;
  tl->gtiles.clear();
  for (int l = 0; l < nlevels; l++) {
    OrthoLevel &g = tl->glev[l];
    g.tiled = false;
    if (g.end <= g.begin) continue;
    const Layout &y = tl->layouts[l];
    const long long nx = y.npts_local[0], ny = y.npts_local[1], nz = y.npts_local[2];
    const long long sx = nx | 1;
    const long long wave = 8LL * (84 + 343) + 16;  // lp <= 6: packed cxyz + dense cijk
    const long long cap = std::min(kTileDoubles, (lds_limit() - 8 * wave) / 8);
    if (sx * ny * nz <= cap) {
      g.by = (int)ny;
      g.bz = (int)nz;
    } else if (cap / (sx * nz) >= 4) {
      g.by = (int)(cap / (sx * nz));
      g.bz = (int)nz;
    } else if (sx * 8 * 2 <= cap) {
      g.by = (int)std::min<long long>(8, ny);
      g.bz = (int)std::min<long long>(nz, cap / (sx * g.by));
    } else {
      continue;
    }
    bool fits = true;
    for (int s2 = g.begin; s2 < g.end && fits; s2++)
      for (int d = 1; d < 3; d++)
        fits = fits && tasks[s2].lo[d] > -32768 && tasks[s2].hi[d] < 32767;
    if (!fits) continue;
    g.tiled = true;
    g.max_waves = kGenTileWaves;
    const int nty = (int)((ny + g.by - 1) / g.by);
    struct Seg { int a, b, t; };
    auto segments = [&](const Task &T, int d, int bsz, std::vector<Seg> &out) {
      out.clear();
      int prev = -2;
      bool open = false;
      for (int c = T.lo[d]; c <= T.hi[d]; c++) {
        const int gg = pmod(c - y.shift_local[d], y.npts_global[d]);
        if (gg < T.blo[d] || gg > T.bhi[d]) {
          open = false;
          continue;
        }
        const int t = gg / bsz;
        if (!open || t != out.back().t || gg != prev + 1)
          out.push_back(Seg{c, c, t});
        else
          out.back().b = c;
        open = true;
        prev = gg;
      }
    };
    struct Keyed { long long key; TileItem it; };
    std::vector<Keyed> lev_items;
    std::vector<Seg> ys, zs;
    for (int s2 = g.begin; s2 < g.end; s2++) {
      const Task &T = tasks[s2];
      segments(T, 1, g.by, ys);
      segments(T, 2, g.bz, zs);
      for (const Seg &a : ys)
        for (const Seg &b : zs) {
          TileItem it{};
          it.task = s2;
          it.j0 = (short)a.a;
          it.j1 = (short)a.b;
          it.k0 = (short)b.a;
          it.k1 = (short)b.b;
          lev_items.push_back(Keyed{((long long)b.t * nty + a.t) << 8 | tl->lp0s[s2], it});
        }
    }
    std::stable_sort(lev_items.begin(), lev_items.end(),
                     [](const Keyed &a, const Keyed &b) { return a.key < b.key; });
    for (size_t a = 0; a < lev_items.size();) {
      const long long tile = lev_items[a].key >> 8;
      size_t b = a;
      while (b < lev_items.size() && (lev_items[b].key >> 8) == tile) b++;
      TileInfo ti{};
      ti.level = l;
      ti.gy0 = (int)(tile % nty) * g.by;
      ti.gz0 = (int)(tile / nty) * g.bz;
      const int first = (int)gitems.size();
      for (size_t q = a; q < b; q++) gitems.push_back(lev_items[q].it);
      for (int v = 0; v <= kLp0Slots; v++) {
        int q = first;
        while (q < (int)gitems.size() && tl->lp0s[gitems[q].task] < v) q++;
        ti.start[v] = q;
      }
      tl->gtiles.push_back(ti);
      a = b;
    }
  }
  upload(tl->d_gitems, gitems);

  for (auto *lv : {&tl->olev, &tl->glev})
    for (auto &o : *lv)
      for (int v = 0; v < kLp0Slots; v++) {
        int s = o.begin;
        while (s < o.end && tl->lp0s[s] < v) s++;
        o.lp_start[v] = s;
      }

  // --- general cells: per level and lp <= 6 the change of variables cxyz
  //     -> cijk (src/grid/ref/grid_ref_collint.h:697-762) as a matrix, and
  //     its transpose ---
  if (tl->has_general) {
    std::vector<double> gmat;
    std::vector<int> goff((size_t)nlevels * kGmatSlots, 0);
    // lp up to the largest of the list plus the l shift of integrate with
    // forces, virial and tau (5), at most 11
    const int lp_top = std::min(kGmatSlots - 1, tl->max_lp + 5);
    for (int l = 0; l < nlevels; l++) {
      if (tl->glev[l].end <= tl->glev[l].begin) continue;
      for (int lp = 0; lp <= lp_top; lp++) {
        goff[(size_t)l * kGmatSlots + lp] = (int)gmat.size();
        append_general_matrix(tl->layouts[l].dh, lp, gmat);
      }
    }
    if (gmat.size() >= (1ULL << 31)) fatal("general cell matrices exceed 2^31 doubles");
    upload(tl->d_gmat, gmat);
    upload(tl->d_gmat_off, goff);
  }

  // --- collocation tile plans of every l shift a grid_func can ask for
  //     (0..4), so that a collocate call allocates nothing for them ---
  for (int sh = 0; sh <= 4; sh++) {
    if (!tl->tiles.empty()) tile_plan(tl, sh);
    if (!tl->gtiles.empty()) tile_plan(tl, sh, true);
  }

  // --- force CSR: atom -> (2*task + side) ---
  std::vector<int> fptr(natoms + 1, 0), fent(2 * (size_t)ntasks);
  for (int s = 0; s < ntasks; s++) {
    fptr[tasks[s].iatom + 1]++;
    fptr[tasks[s].jatom + 1]++;
  }
  for (int a = 0; a < natoms; a++) fptr[a + 1] += fptr[a];
  std::vector<int> fill(fptr.begin(), fptr.end() - 1);
  for (int s = 0; s < ntasks; s++) {
    fent[fill[tasks[s].iatom]++] = 2 * s;
    fent[fill[tasks[s].jatom]++] = 2 * s + 1;
  }

  upload(tl->d_tasks, tasks);
  upload(tl->d_groups, groups);
  upload(tl->d_sphi, sphi_all);
  upload(tl->d_fptr, fptr);
  upload(tl->d_fent, fent);
  upload(tl->d_geo, geo);
  upload(tl->d_geo_cols, geo_cols);
  upload(tl->d_geo_kmin, geo_kmin);
  upload(tl->d_geo_rows, geo_rows);
  upload(tl->d_items, items);
  tl->d_cart.get(sizeof(double) * std::max<long long>(tl->cart_len, 1));
  tl->d_hab_cart.get(sizeof(double) * std::max<long long>(tl->cart_len, 1));
  tl->d_task_out.get(sizeof(double) * kTaskOutStride * (size_t)ntasks);
  tl->grid_dev.resize(nlevels);
  tl->points = -1;
}

// Largest dynamic LDS allocation per workgroup on the current device.
int lds_limit() {
  static int limit = 0;
  if (!limit) {
    int dev = 0, v = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    // minus room for the kernels' static LDS variables (a work-queue slot,
    // the per-task kernels' task record)
    limit = (v > 0 ? v : 65536) - 512;
  }
  return limit;
}

// LDS plan for one launch; ld are the l-range shifts of cab.
LdsPlan make_plan(const TaskListImpl *tl, const int ld[4], bool integrate) {
  LdsPlan p{};
  const int la = tl->max_la + ld[0], lb = tl->max_lb + ld[2];
  const int lp = std::min(tl->max_lp + ld[0] + ld[2], la + lb);
  p.np = lp + 1;
  p.ext = 2 * tl->max_cmax + 2;
  p.mapn = std::max(2 * tl->max_cmax + 1, tl->max_range);
  const int need_rows = std::max({tl->max_rows_ortho, tl->max_rows_general, 1});
  const int n_cab = ncoset(la) * ncoset(lb);
  const int n_alpha = 3 * (la + 1) * (lb + 1) * p.np;
  const int n_cube = p.np * p.np * p.np;
  auto layout = [&](int rowcap, int plcap) {
    p.rowcap = rowcap;
    p.plcap = plcap;
    int o = 0;
    p.o_cxyz = o; o += n_cube;
    p.o_cijk = o; o += tl->has_general ? n_cube : 0;
    p.o_red = o; o += 4 * kTaskOutStride;
    // Setup/finish phase (cab, alpha) and walking phase (tables, plane and
    // row coefficients) never overlap in time and share one region.
    const int u = o;
    p.o_cab = u;
    p.o_alpha = u + n_cab;
    const int phase_a = n_cab + n_alpha;
    p.o_tab = u;
    p.o_cxy = p.o_tab + (tl->has_ortho ? 3 * p.np * p.ext : 0);
    p.o_cxrow = p.o_cxy + plcap * p.np * p.np;
    p.o_stage = p.o_cxrow + rowcap * p.np;
    const int lpb = lpb_bucket(lp);
    const int stage = (integrate && lpb <= 8) ? 64 * ((lpb + 1) * (lpb + 2) / 2 + lpb + 1) : 0;
    p.o_row4 = (p.o_stage + stage + 1) & ~1;  // 16-byte aligned
    const int phase_b = p.o_row4 + 2 * (rowcap + 1) - u;
    p.n_doubles = u + std::max(phase_a, phase_b);
    int i = 0;
    p.i_map = i; i += 3 * p.mapn;
    p.i_plane = i; i += 3 * (plcap + 1);
    p.i_row = i; i += 4 * (rowcap + 1);
    p.i_misc = i; i += 2 + 8;
    p.n_ints = i;
    p.bytes = 8 * p.n_doubles + 4 * p.n_ints;
  };
  const int planes = std::max(1, tl->max_planes);
  const int limit = lds_limit();
  // LDS decides how many workgroups share a CU (160 KiB each): keep the
  // chunk arrays small; fewer planes per chunk only costs a few more syncs.
  layout(std::max(need_rows, 96), std::min(12, planes));
  if (p.bytes > 24 * 1024) layout(need_rows, std::min(4, planes));
  if (p.bytes > limit)
    fatal("task list needs %d bytes of LDS per task (lp=%d, cube half-width %d); "
          "the device limit is %d", p.bytes, lp, tl->max_cmax, limit);
  return p;
}

int num_cus() {
  static int n = 0;
  if (!n) {
    int dev = 0, v = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
    n = v > 0 ? v : 1;
  }
  return n;
}

enum LaunchMode { kCollocate = 0, kIntegrate = 1, kCount = 2 };

// LDS plan of a wave-per-task launch over orthorhombic tasks of one level
// with lp <= `lp`.  Setup/finish kernels hold dense cxyz and {cab, alpha} per
// wave; point kernels hold the per-axis tables, and collocation keeps the
// whole level grid in LDS when it fits next to at least four wave regions.
enum PlanKind { kPlanSetup, kPlanPoints, kPlanWindow, kPlanMfma, kPlanFinish };

// level -1: a merged launch over all levels (not for kPlanWindow).
WavePlan make_wave_plan(const TaskListImpl *tl, int level, int lp, const int ld[4], int kind) {
  const OrthoLevel &o = tl->olev[std::max(level, 0)];
  WavePlan w{};
  w.level = level;
  w.np = lp + 1;
  w.ext = level < 0 ? tl->max_ext_all : o.max_ext;
  int nd;
  if (kind == kPlanSetup || kind == kPlanFinish) {
    // cab [ncoset(la)][ncoset(lb)] + alpha [3][lb+1][la+1][lp+1] for la + lb <= lp;
    // the finish adds its x sums [la+1][lb+1][(lp+1)(lp+2)/2] and keeps cxyz
    // packed, the setup keeps cxyz dense [lp+1]^3
    int n_setup = 0;
    for (int la = 0; la <= tl->max_la + ld[0]; la++)
      for (int lb = 0; lb <= tl->max_lb + ld[2] && la + lb <= lp; lb++)
        n_setup = std::max(n_setup, ncoset(la) * ncoset(lb) + 3 * (la + 1) * (lb + 1) * w.np +
                                        (kind == kPlanFinish ? (la + 1) * (lb + 1) * w.np * (w.np + 1) / 2 : 0));
    w.o_cxyz = 0;
    w.o_u = kind == kPlanFinish ? (lp + 1) * (lp + 2) * (lp + 3) / 6 : w.np * w.np * w.np;
    nd = w.o_u + n_setup;
  } else if (kind == kPlanMfma) {
    // per-axis tables, then the staging of 64 columns (40 doubles each)
    w.o_u = 0;
    w.o_cxyz = 3 * w.ext * w.np;
    nd = w.o_cxyz + 64 * 40;
  } else {
    // packed cxyz of the task (collocate) + per-axis tables
    w.o_cxyz = 0;
    w.o_u = (lp + 1) * (lp + 2) * (lp + 3) / 6;
    nd = w.o_u + 3 * w.ext * w.np;
  }
  w.wave_bytes = (8 * nd + 15) & ~15;
  const int limit = lds_limit();
  if (kind == kPlanWindow) {
    long long npts = 0;  // the largest tile of the launch
    if (level >= 0) {
      npts = (long long)tl->npts_local[3 * level] * o.by * o.bz;
    } else {
      for (int l = 0; l < tl->nlevels; l++)
        if (tl->olev[l].tiled)
          npts = std::max(npts, (long long)tl->npts_local[3 * l] * tl->olev[l].by * tl->olev[l].bz);
    }
    // launch bounds of the tile kernels, and the tuned cap of the level
    const int most = std::min(lp <= 4 ? 12 : 8, level >= 0 ? o.max_waves : 12);
    for (int nw : {12, 10, 8, 6, 4})
      if (nw <= most && 8 * npts + (long long)nw * w.wave_bytes <= limit) {
        w.window = (int)npts;
        w.by = o.by;
        w.bz = o.bz;
        w.nwaves = nw;
        break;
      }
    if (!w.window) fatal("level %d: no room for the collocation tile in LDS", level + 1);
    if (level < 0) w.by = w.bz = 0;  // per unit
  }
  if (!w.window) {
    w.nwaves = 1;
    for (int nw : {4, 2})
      if (nw * w.wave_bytes <= limit) {
        w.nwaves = nw;
        break;
      }
  }
  w.bytes = 8 * w.window + w.nwaves * w.wave_bytes;
  if (w.bytes > limit)
    fatal("level %d: %d bytes of LDS per workgroup needed, the device has %d", level + 1,
          w.bytes, limit);
  return w;
}

// Doubles per task of the packed cxyz buffer for a call with lp shift sh.
int cxyz_stride(const TaskListImpl *tl, int sh) {
  const int lp = std::min(11, tl->max_lp + sh);
  return (lp + 1) * (lp + 2) * (lp + 3) / 6;
}

// Workgroups of the tile kernels for lp shift sh: the items of each tile in
// an lp bucket are dealt to a number of workgroups proportional to their
// count, about four workgroups per CU over a level (one is resident per CU;
// several rounds keep the last one short).
TilePlan &tile_plan(TaskListImpl *tl, int sh, bool general) {
  auto &plans = general ? tl->gtile_plans : tl->tile_plans;
  const std::vector<TileInfo> &tiles = general ? tl->gtiles : tl->tiles;
  const std::vector<OrthoLevel> &lev = general ? tl->glev : tl->olev;
  auto it = plans.find(sh);
  if (it != plans.end()) return it->second;
  TilePlan &plan = plans[sh];
  plan.range.assign(tl->nlevels, {std::make_pair(0, 0), std::make_pair(0, 0), std::make_pair(0, 0)});
  std::vector<TileUnit> units;
  auto slot = [](int v) { return std::min(std::max(v, 0), kLp0Slots); };
  for (int l = 0; l < tl->nlevels; l++) {
    for (int r = 0; r < 3; r++) {
      // lp <= 4, lp 5..6, lp 7..8
      const int v0 = slot(r == 0 ? 0 : (r == 1 ? 5 - sh : 7 - sh));
      const int v1 = slot(r == 0 ? 5 - sh : (r == 1 ? 7 - sh : 9 - sh));
      long long total = 0;
      for (const TileInfo &t : tiles)
        if (t.level == l) total += t.start[v1] - t.start[v0];
      const int first = (int)units.size();
      if (total > 0) {
        const long long rounds = (long long)lev[l].rounds * num_cus();
        const long long per = std::max<long long>(32, (total + rounds - 1) / rounds);
        for (const TileInfo &t : tiles) {
          if (t.level != l) continue;
          const int a = t.start[v0], b = t.start[v1];
          if (b <= a) continue;
          const int nw = (int)std::max<long long>(1, ((long long)(b - a) + per / 2) / per);
          for (int w = 0; w < nw; w++)
            units.push_back(TileUnit{a, b, w, nw, t.gy0, t.gz0, l, lev[l].by, lev[l].bz});
        }
      }
      plan.range[l][r] = std::make_pair(first, (int)units.size() - first);
    }
  }
  // merged launches: the units of every level, bucket by bucket, the units
  // with the most items first (the work queue deals them in this order)
  for (int r = 0; r < 3; r++) {
    const int first = (int)units.size();
    for (int l = 0; l < tl->nlevels; l++)
      for (int u = 0; u < plan.range[l][r].second; u++)
        units.push_back(units[plan.range[l][r].first + u]);
    auto items_of = [](const TileUnit &u) {
      return (u.iend - u.ibeg - u.first + u.stride - 1) / u.stride;
    };
    std::stable_sort(units.begin() + first, units.end(),
                     [&](const TileUnit &a, const TileUnit &b) { return items_of(a) > items_of(b); });
    plan.mrange[r] = std::make_pair(first, (int)units.size() - first);
  }
  upload(plan.d, units);
  return plan;
}

// Orthorhombic tasks, level by level: wave-per-task kernels for the levels
// whose tasks all have a geometry class, for lp <= 6; the per-task kernels
// for the rest.  Collocate: cab -> cxyz pass, then the point kernels;
// integrate: the point kernels, then the cxyz -> hab/forces pass.
void launch_ortho(TaskListImpl *tl, PointParams P, hipStream_t s, int mode) {
  const int sh = P.ld[0] + P.ld[2];
  const TilePlan *plan = mode == kCollocate ? &tile_plan(tl, sh) : nullptr;
  if (plan) P.units = (const TileUnit *)plan->d.p;
  for (int l = 0; l < tl->nlevels; l++) {
    const OrthoLevel &o = tl->olev[l];
    if (o.end <= o.begin) continue;
    auto first = [&](int v) { return o.lp_start[std::min(std::max(v, 0), kLp0Slots - 1)]; };
    // lp <= 4 | lp 5..6 | lp 7..8 (collocate) or 7..11 (integrate) | larger
    int c1 = first(5 - sh), c2 = first(7 - sh);
    int c3 = mode == kCollocate ? first(9 - sh) : first(12 - sh);
    if (!wave_kernels() || !o.classed) c1 = c2 = c3 = o.begin;
    if (c3 > o.begin) {
      const int lp_all = std::min(mode == kCollocate ? 8 : 11, tl->lp0s[c3 - 1] + sh);
      PointParams Q = P;
      Q.task_base = o.begin;
      Q.ntasks = c3 - o.begin;
      Q.wp = make_wave_plan(tl, l, lp_all, P.ld, kPlanSetup);
      if (mode == kCollocate) {
        TimedLaunch tm(6, s);
        // GRID_FUNC_AB without l shifts: lane-per-task setup for la, lb <= 2
        const bool ab = P.func == 100 && P.ld[0] == 0 && P.ld[1] == 0 && P.ld[2] == 0 &&
                        P.ld[3] == 0;
        Q.lane_cxyz = ab ? 1 : 0;
        if (ab) HIP_CHECK(launch_cxyz_lane(Q, s));
        if (!ab || tl->highl_before[c3] > tl->highl_before[o.begin])
          HIP_CHECK(launch_cxyz_wave(Q, s));
      }
      const int cuts[3][2] = {{o.begin, c1}, {c1, c2}, {c2, c3}};
      for (int r = 0; r < 3; r++) {
        const int a = cuts[r][0], b = cuts[r][1];
        if (b <= a) continue;
        const int lpb = r == 0 ? 4 : (r == 1 ? 6 : (mode == kIntegrate ? 11 : 8));
        const int lp = std::min(lpb, tl->lp0s[b - 1] + sh);
        PointParams R = P;
        R.task_base = a;
        R.ntasks = b - a;
        if (mode == kIntegrate) {
          R.wp = make_wave_plan(tl, l, lp, P.ld, r == 2 ? kPlanMfma : kPlanPoints);
          TimedLaunch tm(9, s);
          HIP_CHECK(r == 2 ? launch_integrate_mfma(R, s) : launch_integrate_wave(R, lpb, s));
        } else if (o.tiled) {
          R.wp = make_wave_plan(tl, l, lp, P.ld, kPlanWindow);
          R.unit_base = plan->range[l][r].first;
          TimedLaunch tm(7, s);
          HIP_CHECK(launch_collocate_tiles(R, lpb, plan->range[l][r].second, s));
        } else {
          R.wp = make_wave_plan(tl, l, lp, P.ld, kPlanPoints);
          TimedLaunch tm(8, s);
          HIP_CHECK(launch_collocate_wave(R, lpb, s));
        }
      }
      if (mode == kIntegrate) {
        TimedLaunch tm(10, s);
        // one wave per task: its hab atomics are contiguous per task (a
        // task per lane scatters them over 64 blocks and measured slower);
        // lp 7..11: a workgroup per task (the tables fill most of LDS)
        Q.lane_cxyz = 0;
        for (int v = sh; v <= 6; v++) {  // one launch per lp (LDS plan of that lp)
          const int a = std::max(o.begin, first(v - sh));
          const int b = std::min(c2, v == 6 ? c2 : first(v + 1 - sh));
          if (b <= a) continue;
          PointParams F = Q;
          F.task_base = a;
          F.ntasks = b - a;
          if (finish_lane_ok(F, v)) {
            HIP_CHECK(launch_finish_lane(F, s));
            continue;
          }
          F.wp = make_wave_plan(tl, l, v, P.ld, kPlanFinish);
          HIP_CHECK(launch_finish_wave(F, s));
        }
        if (c3 > c2) {
          PointParams F = Q;
          F.task_base = c2;
          F.ntasks = c3 - c2;
          F.wp = make_wave_plan(tl, l, lp_all, P.ld, kPlanFinish);
          HIP_CHECK(launch_finish_block(F, s));
        }
      }
    }
    if (o.end > c3) {
      PointParams Q = P;
      Q.task_base = c3;
      Q.ntasks = o.end - c3;
      TimedLaunch tm(11, s);
      HIP_CHECK(mode == kIntegrate ? launch_integrate(Q, true, s) : launch_collocate(Q, true, s));
    }
  }
}

// LDS plan of the general point kernels: packed cxyz and dense cijk per
// wave, then the collocation staging: 64 x 9 values and 64 x 9 indices.
WavePlan general_plan(int lp) {
  WavePlan w{};
  w.np = lp + 1;
  w.o_cxyz = 0;
  w.o_u = w.np * w.np * w.np;
  w.wave_bytes = (16 * w.o_u + 8 * 64 * 9 + 4 * 64 * 9 + 15) & ~15;
  w.nwaves = 4;
  w.bytes = w.nwaves * w.wave_bytes;
  return w;
}

// LDS plan of the general column walk (integrate) over the general tasks of
// `level` (-1: all levels) with lp <= `lp`: the task's sums (and for lp > 6
// the MFMA staging), the ratio tables, then the row table of a slab of
// planes, the footprint and the column starts of the rows j.
WavePlan general_col_plan(const TaskListImpl *tl, int level, int lp) {
  int ni = 1, nj = 1, nk = 1;
  long long rows = 1;
  for (int l = 0; l < tl->nlevels; l++) {
    if (level >= 0 && l != level) continue;
    const OrthoLevel &g = tl->glev[l];
    if (g.end <= g.begin) continue;
    ni = std::max(ni, g.gni);
    nj = std::max(nj, g.gnj);
    nk = std::max(nk, g.gnk);
    rows = std::max(rows, g.grows);
  }
  WavePlan w{};
  w.level = level;
  w.np = lp + 1;
  const int nc = (lp + 1) * (lp + 2) * (lp + 3) / 6;
  w.o_cxyz = 0;
  w.o_u = nc;
  const int nd = nc + (lp > 6 ? 64 * 40 : 0);
  w.o_q = 8 * nd;
  w.q_cap = ni + nj + 2 * nk;  // qi, qj, qk, then dk per plane
  w.o_rows = w.o_q + 8 * w.q_cap;
  // a slab holds at least one plane; 1024 rows (4 KiB) keep the wave region
  // small (more waves per CU) and cover most tasks in one or two slabs
  w.rows_cap = (int)std::max<long long>(nj, std::min<long long>(rows, 1024));
  w.o_jinfo = (w.o_rows + 4 * w.rows_cap + 15) & ~15;
  w.j_cap = nj;
  w.o_cstart = w.o_jinfo + 16 * nj;
  w.wave_bytes = (w.o_cstart + 4 * (nj + 1) + 15) & ~15;
  const int limit = lds_limit();
  w.nwaves = 0;
  for (int nw : {4, 2, 1})
    if (nw * w.wave_bytes <= limit) {
      w.nwaves = nw;
      break;
    }
  if (!w.nwaves)
    fatal("general integrate needs %d bytes of LDS per task, the device has %d", w.wave_bytes,
          limit);
  w.bytes = w.nwaves * w.wave_bytes;
  return w;
}

// LDS plan of a general tile launch over `level` (-1: every tiled general
// level): the largest tile (row stride nx | 1) and wave regions holding the
// packed cxyz and the dense cijk.
bool general_tile_plan(const TaskListImpl *tl, int level, int lp, WavePlan &w) {
  const int cap = level >= 0 ? tl->glev[level].max_waves : kGenTileWaves;
  w = WavePlan{};
  w.level = level;
  w.np = lp + 1;
  w.o_cxyz = 0;
  w.o_u = (lp + 1) * (lp + 2) * (lp + 3) / 6;
  w.wave_bytes = (8 * (w.o_u + w.np * w.np * w.np) + 15) & ~15;
  long long win = 0;
  for (int l = 0; l < tl->nlevels; l++)
    if ((level < 0 || l == level) && tl->glev[l].tiled)
      win = std::max(win, (long long)(tl->npts_local[3 * l] | 1) * tl->glev[l].by * tl->glev[l].bz);
  if (win <= 0) return false;
  const int most = std::min(lp <= 4 ? kGenTileWaves : 8, cap);
  for (int nw : {16, 12, 10, 8, 6, 4})
    if (nw <= most && 8 * win + (long long)nw * w.wave_bytes <= lds_limit()) {
      w.window = (int)win;
      w.nwaves = nw;
      w.bytes = 8 * w.window + nw * w.wave_bytes;
      return true;
    }
  return false;
}

// General tiles of one level (or all levels: level -1, unit range of the
// merged arrangement), bucket r of lp shift sh.
void launch_general_tile_bucket(TaskListImpl *tl, const PointParams &P, int level, int r, int lp,
                                hipStream_t s) {
  const int sh = P.ld[0] + P.ld[2];
  const TilePlan &tp = tile_plan(tl, sh, true);
  const std::pair<int, int> rg = level < 0 ? tp.mrange[r] : tp.range[level][r];
  if (rg.second <= 0) return;
  const int lpb = r == 0 ? 4 : 6;
  PointParams R = P;
  R.units = (const TileUnit *)tp.d.p;
  R.items = (const TileItem *)tl->d_gitems.p;
  R.unit_base = rg.first;
  if (!general_tile_plan(tl, level, std::min(lpb, lp), R.wp)) {
    if (level >= 0) fatal("level %d: no room for the general collocation tile in LDS", level + 1);
    for (int l = 0; l < tl->nlevels; l++)  // the largest tile and tables do not fit together
      if (tl->glev[l].tiled) launch_general_tile_bucket(tl, P, l, r, lp, s);
    return;
  }
  TimedLaunch tm(13, s);
  HIP_CHECK(launch_general_tiles(R, lpb, rg.second, s));
}

// General tasks, level by level: wave-per-task kernels for lp <= 6 (cab ->
// cxyz pass, point kernel; integrate: point kernel, cxyz -> hab pass), the
// per-task kernels for larger lp.
void launch_general(TaskListImpl *tl, PointParams P, hipStream_t s, int mode) {
  const int sh = P.ld[0] + P.ld[2];
  for (int l = 0; l < tl->nlevels; l++) {
    const OrthoLevel &g = tl->glev[l];
    if (g.end <= g.begin) continue;
    auto first = [&](int v) { return g.lp_start[std::min(std::max(v, 0), kLp0Slots - 1)]; };
    // lp <= 4 | lp 5..6 | lp 7..11 (integrate) | larger
    int c1 = first(5 - sh), c2 = first(7 - sh);
    int c3 = mode == kIntegrate ? first(12 - sh) : c2;
    if (!wave_kernels()) c1 = c2 = c3 = g.begin;
    if (c2 > g.begin) {
      const int lp_all = std::min(6, tl->lp0s[c2 - 1] + sh);
      PointParams Q = P;
      Q.task_base = g.begin;
      Q.ntasks = c2 - g.begin;
      Q.wp = make_wave_plan(tl, l, lp_all, P.ld, kPlanSetup);
      if (mode == kCollocate) {
        TimedLaunch tm(6, s);
        const bool ab = P.func == 100 && P.ld[0] == 0 && P.ld[1] == 0 && P.ld[2] == 0 &&
                        P.ld[3] == 0;
        Q.lane_cxyz = ab ? 1 : 0;
        if (ab) HIP_CHECK(launch_cxyz_lane(Q, s));
        if (!ab || tl->highl_before[c2] > tl->highl_before[g.begin])
          HIP_CHECK(launch_cxyz_wave(Q, s));
        // the general kernels take cijk: lp <= 4 a task per lane, lp 5..6 a
        // task per wave
        const int c5 = std::min(c2, std::max(g.begin, first(5 - sh)));
        PointParams Qa = Q, Qb = Q;
        Qa.ntasks = c5 - g.begin;
        Qb.task_base = c5;
        Qb.ntasks = c2 - c5;
        HIP_CHECK(launch_general_cijk_lane(Qa, s));
        HIP_CHECK(launch_general_cijk(Qb, s));
      }
      if (mode == kIntegrate) {
        // one launch per lp: the lane sums take the registers that lp needs
        for (int v = 0; v <= 6; v++) {
          const int a = first(v - sh), b = v == 6 ? c2 : first(v + 1 - sh);
          if (v < sh || b <= a) continue;
          PointParams R = P;
          R.task_base = a;
          R.ntasks = b - a;
          R.wp = general_col_plan(tl, l, v);
          TimedLaunch tm(13, s);
          HIP_CHECK(launch_general_integrate(R, v, v, s));
        }
      }
      const int cuts[2][2] = {{g.begin, c1}, {c1, c2}};
      for (int r = 0; r < 2 && mode == kCollocate; r++) {
        const int a = cuts[r][0], b = cuts[r][1];
        if (b <= a) continue;
        const int lpb = r == 0 ? 4 : 6;
        if (g.tiled) {
          launch_general_tile_bucket(tl, P, l, r, tl->lp0s[b - 1] + sh, s);
          continue;
        }
        PointParams R = P;
        R.task_base = a;
        R.ntasks = b - a;
        R.wp = general_plan(std::min(lpb, tl->lp0s[b - 1] + sh));
        TimedLaunch tm(13, s);
        HIP_CHECK(launch_general_collocate(R, lpb, s));
      }
      if (mode == kIntegrate) {
        // one launch per lp, each with the LDS plan of that lp (the small-l
        // tasks, most of them, then run many more waves per CU)
        TimedLaunch tm(10, s);
        Q.lane_cxyz = 0;
        for (int v = sh; v <= 6; v++) {
          const int a = std::max(g.begin, first(v - sh));
          const int b = std::min(c2, v == 6 ? c2 : first(v + 1 - sh));
          if (b <= a) continue;
          PointParams F = Q;
          F.task_base = a;
          F.ntasks = b - a;
          if (finish_lane_ok(F, v)) {
            HIP_CHECK(launch_finish_lane(F, s));
            continue;
          }
          F.wp = make_wave_plan(tl, l, v, P.ld, kPlanFinish);
          HIP_CHECK(launch_finish_wave(F, s));
        }
      }
    }
    if (c3 > c2) {  // integrate, lp 7..11: the column sums on the matrix cores
      PointParams R = P;
      R.task_base = c2;
      R.ntasks = c3 - c2;
      const int lp = std::min(11, tl->lp0s[c3 - 1] + sh);
      R.wp = general_col_plan(tl, l, lp);
      {
        TimedLaunch tm(13, s);
        HIP_CHECK(launch_general_integrate(R, 7, 11, s));
      }
      TimedLaunch tm(10, s);
      R.wp = make_wave_plan(tl, l, lp, P.ld, kPlanFinish);
      HIP_CHECK(launch_finish_block(R, s));
    }
    if (g.end > c3) {
      PointParams Q = P;
      Q.task_base = c3;
      Q.ntasks = g.end - c3;
      TimedLaunch tm(11, s);
      HIP_CHECK(mode == kIntegrate ? launch_integrate(Q, false, s) : launch_collocate(Q, false, s));
    }
  }
}

// Task index lists of the merged launches for lp shift sh.
MergedPlan &merged_plan(TaskListImpl *tl, int sh) {
  auto it = tl->merged_plans.find(sh);
  if (it != tl->merged_plans.end()) return it->second;
  MergedPlan &plan = tl->merged_plans[sh];
  std::vector<int> lists[kMNum];
  const bool wave = wave_kernels();
  for (int t = 0; t < tl->ntasks; t++) {
    const int lp = tl->lp0s[t] + sh;
    const int c = tl->cls[t];
    if (c == kClsSkip) continue;
    if (c == kClsGeneral) {
      if (wave && lp <= 6) {
        lists[kMSetup].push_back(t);
        lists[lp <= 4 ? kMGenA : kMGenB].push_back(t);
        if (!tl->glev[tl->task_level[t]].tiled) lists[lp <= 4 ? kMGenDirA : kMGenDirB].push_back(t);
      } else {
        lists[kMGenRest].push_back(t);
        lists[(wave && lp <= 11) ? kMGenHl : kMGenRestI].push_back(t);
      }
    } else if (c == kClsOrthoClassed && wave && lp <= 6) {
      lists[kMSetup].push_back(t);
      lists[lp <= 4 ? kMOrthoA : kMOrthoB].push_back(t);
      if (!tl->olev[tl->task_level[t]].tiled) lists[lp <= 4 ? kMDirectA : kMDirectB].push_back(t);
    } else if (c == kClsOrthoClassed && wave && lp <= 11) {
      if (lp <= 8) {
        lists[kMSetupHi].push_back(t);
        lists[kMOrthoC].push_back(t);
        if (!tl->olev[tl->task_level[t]].tiled) lists[kMDirectC].push_back(t);
      } else {
        lists[kMOrthoRestC].push_back(t);
      }
      lists[kMOrthoHl].push_back(t);  // integrate
    } else {
      lists[kMOrthoRest].push_back(t);
      lists[kMOrthoRestC].push_back(t);
    }
  }
  // the persistent integrate launches take tasks from a work queue in list
  // order: the largest first, so the smallest fill the tail
  for (int m : {kMOrthoA, kMOrthoB})
    std::stable_sort(lists[m].begin(), lists[m].end(),
                     [&](int a, int b) { return tl->class_npts[a] > tl->class_npts[b]; });
  std::vector<int> all;
  for (int m = 0; m < kMNum; m++) {
    plan.off[m] = (int)all.size();
    int mx = 0;
    for (int t : lists[m]) mx = std::max(mx, tl->lp0s[t] + sh);
    plan.lp[m] = mx;
    all.insert(all.end(), lists[m].begin(), lists[m].end());
  }
  plan.off[kMNum] = (int)all.size();
  upload(plan.d, all);
  return plan;
}

// Batched lists: every launch kind once over all levels (task index lists);
// collocation adds straight into HBM (no LDS tiles).
void launch_merged(TaskListImpl *tl, PointParams P, hipStream_t s, int mode) {
  const int sh = P.ld[0] + P.ld[2];
  MergedPlan &mp = merged_plan(tl, sh);
  // work-queue cursors of the persistent launches (at most one per tile
  // bucket and level, plus the integrate launches), zeroed by a kernel (no
  // memset node in a captured graph); the buffer is sized once per list
  const int slots = 2 * (3 * tl->nlevels + 8);
  int *cursors = nullptr;
  int next_slot = 0;
  if (tl->work_queue) {
    cursors = (int *)tl->d_queue.get(sizeof(int) * slots);
    HIP_CHECK(launch_fill((double *)cursors, slots / 2, 0.0, s));
  }
  auto queue_of = [&]() -> int * {
    if (!cursors) return nullptr;
    if (next_slot >= slots) fatal("merged call: more than %d persistent launches", slots);
    return cursors + next_slot++;
  };
  const int *idx = (const int *)mp.d.p;
  auto sub = [&](int m) {
    PointParams Q = P;
    Q.task_idx = idx + mp.off[m];
    Q.task_base = 0;
    Q.ntasks = mp.off[m + 1] - mp.off[m];
    return Q;
  };
  PointParams Q = sub(kMSetup);
  {
    Q.wp = make_wave_plan(tl, -1, std::min(6, mp.lp[kMSetup]), P.ld, kPlanSetup);
    if (mode == kCollocate) {
      TimedLaunch tm(6, s);
      const bool ab = P.func == 100 && P.ld[0] == 0 && P.ld[1] == 0 && P.ld[2] == 0 &&
                      P.ld[3] == 0;
      Q.lane_cxyz = ab ? 1 : 0;
      if (ab) HIP_CHECK(launch_cxyz_lane(Q, s));
      HIP_CHECK(launch_cxyz_wave(Q, s));
      PointParams H = sub(kMSetupHi);  // lp 7..8
      if (H.ntasks > 0) {
        H.wp = make_wave_plan(tl, -1, std::min(8, mp.lp[kMSetupHi]), P.ld, kPlanSetup);
        H.lane_cxyz = 0;
        HIP_CHECK(launch_cxyz_wave(H, s));
      }
      PointParams G = sub(kMGenA);  // cxyz -> cijk: lp <= 4 a task per lane
      HIP_CHECK(launch_general_cijk_lane(G, s));
      G = sub(kMGenB);              // lp 5..6 a task per wave
      HIP_CHECK(launch_general_cijk(G, s));
    }
    if (mode == kIntegrate) {
      for (int m : {kMOrthoA, kMOrthoB}) {
        PointParams R = sub(m);
        if (R.ntasks <= 0) continue;
        const int lpb = m == kMOrthoA ? 4 : 6;
        R.wp = make_wave_plan(tl, -1, std::min(lpb, mp.lp[m]), P.ld, kPlanPoints);
        R.queue = queue_of();
        R.qchunk = 4;
        TimedLaunch tm(9, s);
        HIP_CHECK(launch_integrate_wave(R, lpb, s));
      }
    } else {
      // tiled levels through their LDS tiles (one launch per bucket over all
      // levels), the others straight into HBM
      const TilePlan &tp = tile_plan(tl, sh);
      long long max_window = 0;
      for (int l = 0; l < tl->nlevels; l++)
        if (tl->olev[l].tiled)
          max_window = std::max(max_window, (long long)tl->npts_local[3 * l] * tl->olev[l].by *
                                                tl->olev[l].bz);
      for (int r = 0; r < 3; r++) {
        if (tp.mrange[r].second <= 0) continue;
        const int lpb = r == 0 ? 4 : (r == 1 ? 6 : 8);
        const int lp = std::min(lpb, mp.lp[r == 0 ? kMOrthoA : (r == 1 ? kMOrthoB : kMOrthoC)]);
        PointParams R = P;
        R.units = (const TileUnit *)tp.d.p;
        const int np = lp + 1;
        const long long wave = 8LL * ((lp + 1) * (lp + 2) * (lp + 3) / 6 + 3LL * tl->max_ext_all * np) + 16;
        if (8 * max_window + 4 * wave <= lds_limit()) {
          R.unit_base = tp.mrange[r].first;
          R.wp = make_wave_plan(tl, -1, lp, P.ld, kPlanWindow);
          R.queue = queue_of();
          TimedLaunch tm(7, s);
          HIP_CHECK(launch_collocate_tiles(R, lpb, tp.mrange[r].second, s));
        } else {
          // the largest tile and the largest tables do not fit together:
          // one launch per level
          for (int l = 0; l < tl->nlevels; l++) {
            if (!tl->olev[l].tiled || tp.range[l][r].second <= 0) continue;
            R.unit_base = tp.range[l][r].first;
            R.wp = make_wave_plan(tl, l, lp, P.ld, kPlanWindow);
            R.queue = queue_of();
            TimedLaunch tm(7, s);
            HIP_CHECK(launch_collocate_tiles(R, lpb, tp.range[l][r].second, s));
          }
        }
      }
      for (int m : {kMDirectA, kMDirectB, kMDirectC}) {
        PointParams R = sub(m);
        if (R.ntasks <= 0) continue;
        const int lpb = m == kMDirectA ? 4 : (m == kMDirectB ? 6 : 8);
        R.wp = make_wave_plan(tl, -1, std::min(lpb, mp.lp[m]), P.ld, kPlanPoints);
        TimedLaunch tm(8, s);
        HIP_CHECK(launch_collocate_wave(R, lpb, s));
      }
    }
    if (mode == kCollocate)
      for (int r = 0; r < 2; r++) launch_general_tile_bucket(tl, P, -1, r, mp.lp[r == 0 ? kMGenA : kMGenB], s);
    for (int m : (mode == kIntegrate ? std::array<int, 2>{kMGenA, kMGenB}
                                     : std::array<int, 2>{kMGenDirA, kMGenDirB})) {
      PointParams R = sub(m);
      if (R.ntasks <= 0) continue;
      const int lpb = (m == kMGenA || m == kMGenDirA) ? 4 : 6;
      const int lp = std::min(lpb, mp.lp[m]);
      R.wp = mode == kIntegrate ? general_col_plan(tl, -1, lp) : general_plan(lp);
      TimedLaunch tm(13, s);
      HIP_CHECK(mode == kIntegrate ? launch_general_integrate(R, lpb == 4 ? 0 : 5, lpb, s)
                                   : launch_general_collocate(R, lpb, s));
    }
    if (mode == kIntegrate) {
      PointParams R = sub(kMGenHl);
      if (R.ntasks > 0) {
        const int lp = std::min(11, mp.lp[kMGenHl]);
        R.wp = general_col_plan(tl, -1, lp);
        {
          TimedLaunch tm(13, s);
          HIP_CHECK(launch_general_integrate(R, 7, 11, s));
        }
        TimedLaunch tm(10, s);
        R.wp = make_wave_plan(tl, -1, lp, P.ld, kPlanFinish);
        HIP_CHECK(launch_finish_block(R, s));
      }
    }
    if (mode == kIntegrate) {
      PointParams H = sub(kMOrthoHl);
      if (H.ntasks > 0) {
        H.wp = make_wave_plan(tl, -1, std::min(11, mp.lp[kMOrthoHl]), P.ld, kPlanMfma);
        TimedLaunch tm(9, s);
        HIP_CHECK(launch_integrate_mfma(H, s));
      }
      TimedLaunch tm(10, s);
      Q.lane_cxyz = 0;
      Q.wp = make_wave_plan(tl, -1, std::min(6, mp.lp[kMSetup]), P.ld, kPlanFinish);
      HIP_CHECK(launch_finish_wave(Q, s));
      if (H.ntasks > 0) {
        H.wp = make_wave_plan(tl, -1, std::min(11, mp.lp[kMOrthoHl]), P.ld, kPlanFinish);
        HIP_CHECK(launch_finish_block(H, s));
      }
    }
  }
  for (int m : {mode == kIntegrate ? kMOrthoRest : kMOrthoRestC,
                mode == kIntegrate ? kMGenRestI : kMGenRest}) {
    PointParams R = sub(m);
    if (R.ntasks <= 0) continue;
    TimedLaunch tm(11, s);
    const bool ortho = m == kMOrthoRest || m == kMOrthoRestC;
    HIP_CHECK(mode == kIntegrate ? launch_integrate(R, ortho, s) : launch_collocate(R, ortho, s));
  }
}

// Launches the orthorhombic and the general task ranges (tasks are sorted
// by class: orthorhombic, general, skipped).
void launch_ranges(TaskListImpl *tl, PointParams P, hipStream_t s, int mode) {
  if (tl->merged && mode != kCount) {
    launch_merged(tl, P, s, mode);
    return;
  }
  if (mode == kCount) {
    // tasks with a geometry class are counted on the host
    for (int l = 0; l < tl->nlevels; l++) {
      const OrthoLevel &o = tl->olev[l];
      if (o.end <= o.begin || o.classed) continue;
      PointParams Q = P;
      Q.task_base = o.begin;
      Q.ntasks = o.end - o.begin;
      HIP_CHECK(launch_collocate(Q, true, s));
    }
  } else if (tl->n_ortho > 0) {
    launch_ortho(tl, P, s, mode);
  }
  if (tl->n_general > 0 && mode != kCount) {
    launch_general(tl, P, s, mode);
  } else if (tl->n_general > 0) {
    P.task_base = (int)tl->n_ortho;
    P.ntasks = (int)tl->n_general;
    HIP_CHECK(launch_collocate(P, false, s));
  }
}

// Per-call device buffers of the allocation-free entry point, in its static
// scratch array (nullptr members: the list's own buffers).
struct CallScratch {
  const Level *levels = nullptr;
  double *cart = nullptr;
  double *cxyz = nullptr;
};

// Geometry classes and the per-task cxyz buffer of the wave-per-task kernels.
void set_wave_buffers(TaskListImpl *tl, PointParams &P, const CallScratch *cs = nullptr) {
  P.geo = (const GeoClass *)tl->d_geo.p;
  P.geo_cols = (const int *)tl->d_geo_cols.p;
  P.geo_kmin = (const int *)tl->d_geo_kmin.p;
  P.geo_rows = (const int *)tl->d_geo_rows.p;
  P.items = (const TileItem *)tl->d_items.p;
  P.cxyz_stride = cxyz_stride(tl, P.ld[0] + P.ld[2]);
  P.gmat = (const double *)tl->d_gmat.p;
  P.gmat_off = (const int *)tl->d_gmat_off.p;
  P.cxyz_all = (cs && cs->cxyz) ? cs->cxyz
                                 : (double *)tl->d_cxyz.get(sizeof(double) *
                                                            (size_t)P.cxyz_stride *
                                                            (size_t)std::max(tl->ntasks, 1));
}

long long level_points(const TaskListImpl *tl, int l) {
  return (long long)tl->npts_local[3 * l] * tl->npts_local[3 * l + 1] * tl->npts_local[3 * l + 2];
}

// Level table of a call (device pointer); with `gridf`, level l of the
// potential rounded to FP32 sits at gridf + (points of the levels before l).
const Level *set_levels(TaskListImpl *tl, double *const *grids, hipStream_t s,
                        const float *gridf = nullptr) {
  if (tl->static_levels) return tl->static_levels;
  TaskListImpl::LevelSlot &ls = tl->level_slots[tl->next_level_slot];
  tl->next_level_slot = (tl->next_level_slot + 1) % TaskListImpl::kLevelSlots;
  if (!ls.h) {
    HIP_CHECK(hipHostMalloc((void **)&ls.h, sizeof(Level) * (size_t)std::max(tl->nlevels, 1)));
    HIP_CHECK(hipEventCreateWithFlags(&ls.copied, hipEventDisableTiming));
  } else {
    HIP_CHECK(hipEventSynchronize(ls.copied));  // the slot's previous copy is done
  }
  ls.d.get(sizeof(Level) * (size_t)std::max(tl->nlevels, 1));
  long long off = 0;
  for (int l = 0; l < tl->nlevels; l++) {
    Level &L = ls.h[l];
    L.lay = tl->layouts[l];
    L.grid = grids[l];
    L.npts = (long long)L.lay.npts_local[0] * L.lay.npts_local[1] * L.lay.npts_local[2];
    L.gridf = gridf ? gridf + off : nullptr;
    off += L.npts;
  }
  HIP_CHECK(hipMemcpyAsync(ls.d.p, ls.h, sizeof(Level) * tl->nlevels, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipEventRecord(ls.copied, s));
  tl->last_level_slot = (int)(&ls - tl->level_slots);
  return (const Level *)ls.d.p;
}

// After the last kernel of a call: the slot is free once that kernel is done
// (also when the next call of the ring comes on another stream).
void end_levels(TaskListImpl *tl, hipStream_t s) {
  if (tl->static_levels) return;
  HIP_CHECK(hipEventRecord(tl->level_slots[tl->last_level_slot].copied, s));
}

void count_stats(const TaskListImpl *tl, bool collocate, int lpshift) {
  std::lock_guard<std::mutex> g(S.mu);
  for (int o = 0; o < 2; o++)
    for (int l = 0; l < kCounterLp; l++) {
      if (!tl->hist[o][l]) continue;
      const int kern = (collocate ? 0 : 1) + (o ? 0 : 2);
      S.counters[kern][std::min(l + lpshift, kCounterLp - 1)] += tl->hist[o][l];
    }
}

int decontract_chunk(const TaskListImpl *tl) {
  // columns of the Cartesian block per LDS pass: nsb*cach doubles <= 8 KiB
  // per wave (one wave per group, four per workgroup)
  const int nsb = std::max(tl->max_nsb, 1);
  return std::max(1, std::min(std::max(tl->max_ncoa, 1), 1024 / nsb));
}

// LDS plan of the staged (de)contraction kernels: nc columns of the
// Cartesian block per pass, doubles per wave (decontract, contract), waves
// per workgroup and resident workgroups; ok = false when even one column
// does not fit (the unstaged kernels take those lists).
struct StagePlan {
  bool ok = false;
  bool pf_d = false, pf_c = false;  // operands fit the prefetch registers
  int nc_d = 1, wd_d = 0, nc_c = 1, wd_c = 0, nwaves = 4, blocks_d = 1, blocks_c = 1;
};

StagePlan stage_plan(const TaskListImpl *tl) {
  StagePlan p;
  const long long A = std::max(tl->max_nsa, 1), B = std::max(tl->max_nsb, 1);
  const long long CA = std::max(tl->max_ncoa, 1), CB = std::max(tl->max_ncob, 1);
  const long long budget = 2048;  // doubles per wave (16 KiB)
  // the kernels' index arithmetic (fdiv) holds for dimensions up to 256 and
  // matrices of fewer than 4096 elements
  if (std::max(std::max(A, B), std::max(CA, CB)) > 256) return p;
  auto fit = [&](long long fixed, long long per, int &nc, int &wd) {
    const long long n = std::max<long long>(1, std::min(CA, (budget - fixed) / per));
    nc = (int)n;
    wd = (int)((fixed + n * per + 1) & ~1LL);
  };
  fit(A * B + B * CB, A + B, p.nc_d, p.wd_d);
  fit(B * CB + A * B, CB + B + A, p.nc_c, p.wd_c);
  const long long ncm = std::max(p.nc_d, p.nc_c);
  if (A * B >= 4096 || B * CB >= 4096 || std::max(std::max(A, B), CB) * ncm >= 4096) return p;
  const long long wd = std::max(p.wd_d, p.wd_c);
  const long long limit = lds_limit();
  if (8 * wd > limit) return p;
  p.nwaves = (int)std::max<long long>(1, std::min<long long>(4, 65536 / (8 * wd)));
  auto resident = [&](int wdk) {
    const long long per_wg = 8LL * wdk * p.nwaves;
    return (int)std::max<long long>(1, std::min<long long>(32 / p.nwaves, limit / per_wg));
  };
  p.blocks_d = num_cus() * resident(p.wd_d);
  p.blocks_c = num_cus() * resident(p.wd_c);
  const long long cap = 64LL * stage_regs();
  p.pf_d = A * B <= cap && B * CB <= cap && A * std::min<long long>(p.nc_d, CA) <= cap;
  p.pf_c = CB * std::min<long long>(p.nc_c, CA) <= cap && B * CB <= cap &&
           A * std::min<long long>(p.nc_c, CA) <= cap;
  p.ok = true;
  return p;
}

void run_decontract(TaskListImpl *tl, const double *pab, hipStream_t s,
                    double *cart = nullptr) {
  TimedLaunch tm(2, s);
  const StagePlan sp = stage_plan(tl);
  if (sp.ok) {
    HIP_CHECK(launch_decontract_staged((const Group *)tl->d_groups.p, tl->ngroups, pab,
                                       (const double *)tl->d_sphi.p,
                                       cart ? cart : (double *)tl->d_cart.p, tl->cart_len,
                                       sp.nc_d, sp.wd_d, sp.nwaves, sp.blocks_d, sp.pf_d, s));
    return;
  }
  const int cach = decontract_chunk(tl);
  HIP_CHECK(launch_decontract((const Group *)tl->d_groups.p, tl->ngroups, pab, 0,
                              (const double *)tl->d_sphi.p, cart ? cart : (double *)tl->d_cart.p,
                              tl->cart_len, cach, std::max(tl->max_nsb, 1), s));
}

void run_collocate(TaskListImpl *tl, int func, const double *pab, double *const *grids,
                   hipStream_t s, const CallScratch *cs = nullptr) {
  int ld[4];
  if (!func_ldiffs(func, ld)) fatal("unknown grid_func %d", func);
  const Level *levels = cs ? cs->levels : set_levels(tl, grids, s);
  {
    TimedLaunch tm(4, s);
    for (int l = 0; l < tl->nlevels; l++)
      HIP_CHECK(launch_fill(grids[l], level_points(tl, l), 0.0, s));
  }
  double *cart = cs ? cs->cart : (double *)tl->d_cart.p;
  run_decontract(tl, pab, s, cart);
  PointParams P{};
  P.tasks = (const Task *)tl->d_tasks.p;
  P.ntasks = tl->ntasks;
  P.groups = (const Group *)tl->d_groups.p;
  P.levels = levels;
  P.cart = cart;
  P.cart_len = tl->cart_len;
  P.func = func;
  P.fp32 = precision() == GRID_MI355X_FP32_MIXED;
  for (int i = 0; i < 4; i++) P.ld[i] = ld[i];
  P.plan = make_plan(tl, ld, false);
  set_wave_buffers(tl, P, cs);
  {
    TimedLaunch tm(0, s);
    launch_ranges(tl, P, s, kCollocate);
  }
  if (!cs) end_levels(tl, s);
  count_stats(tl, true, ld[0] + ld[2]);
}

// Per-list virial of a batched list: virial[l * 9 ..] sums the tasks
// ent[ptr[l] .. ptr[l + 1]).
struct ListVirial {
  const int *ptr, *ent;
  int nlists;
};

void run_integrate(TaskListImpl *tl, bool tau, const double *pab, const double *const *grids,
                   double *hab, long long hab_len, double *forces, double *virial,
                   hipStream_t s, const ListVirial *lv = nullptr) {
  const bool do_f = forces != nullptr, do_v = virial != nullptr;
  if (do_v && !do_f)
    fatal("virial requires forces (src/grid/ref/grid_ref_integrate.c:109)");
  if ((do_f || do_v) && !pab) fatal("forces/virial require pab_blocks");
  const bool fp32 = precision() == GRID_MI355X_FP32_MIXED;
  int ld[4];
  integrate_ldiffs(do_f, do_v, tau, ld);
  float *gridf = nullptr;
  if (fp32) {
    long long total = 0;
    for (int l = 0; l < tl->nlevels; l++) total += level_points(tl, l);
    gridf = (float *)tl->d_gridf.get(sizeof(float) * (size_t)std::max(total, 1LL));
    TimedLaunch tm(12, s);
    long long off = 0;
    for (int l = 0; l < tl->nlevels; l++) {
      HIP_CHECK(launch_to_float(grids[l], gridf + off, level_points(tl, l), s));
      off += level_points(tl, l);
    }
  }
  const Level *levels = set_levels(tl, (double *const *)grids, s, gridf);
  if (do_f) run_decontract(tl, pab, s);
  {
    TimedLaunch tm(4, s);
    HIP_CHECK(launch_fill((double *)tl->d_hab_cart.p, tl->cart_len, 0.0, s));
  }
  PointParams P{};
  P.tasks = (const Task *)tl->d_tasks.p;
  P.ntasks = tl->ntasks;
  P.groups = (const Group *)tl->d_groups.p;
  P.levels = levels;
  P.cart = (const double *)tl->d_cart.p;
  P.hab_cart = (double *)tl->d_hab_cart.p;
  P.task_out = (double *)tl->d_task_out.p;
  P.cart_len = tl->cart_len;
  for (int i = 0; i < 4; i++) P.ld[i] = ld[i];
  P.compute_tau = tau;
  P.do_forces = do_f;
  P.do_virial = do_v;
  P.fp32 = fp32;
  P.plan = make_plan(tl, ld, true);
  set_wave_buffers(tl, P);
  if (do_f && tl->n_skip > 0) {  // skipped tasks contribute no forces
    const long long first = (long long)(tl->n_ortho + tl->n_general) * kTaskOutStride;
    HIP_CHECK(launch_fill((double *)tl->d_task_out.p + first,
                          (long long)tl->n_skip * kTaskOutStride, 0.0, s));
  }
  {
    TimedLaunch tm(1, s);
    launch_ranges(tl, P, s, kIntegrate);
  }
  {
    TimedLaunch tm(4, s);
    HIP_CHECK(launch_fill(hab, hab_len, 0.0, s));
  }
  {
    TimedLaunch tm(3, s);
    const StagePlan sp = stage_plan(tl);
    if (sp.ok) {
      HIP_CHECK(launch_contract_staged((const Group *)tl->d_groups.p, tl->ngroups,
                                       (const double *)tl->d_hab_cart.p,
                                       (const double *)tl->d_sphi.p, hab, hab_len, sp.nc_c,
                                       sp.wd_c, sp.nwaves, sp.blocks_c, sp.pf_c, s));
    } else {
      const int cach = decontract_chunk(tl);
      HIP_CHECK(launch_contract((const Group *)tl->d_groups.p, tl->ngroups,
                                (const double *)tl->d_hab_cart.p, (const double *)tl->d_sphi.p,
                                hab, hab_len, cach, std::max(tl->max_nsb, 1),
                                tl->max_nsa_nsb, s));
    }
  }
  if (do_f || do_v) {
    TimedLaunch tm(5, s);
    if (do_f)
      HIP_CHECK(launch_reduce_forces((const int *)tl->d_fptr.p, (const int *)tl->d_fent.p,
                                     (const double *)tl->d_task_out.p, forces, tl->natoms, s));
    if (do_v)
      HIP_CHECK(lv ? launch_reduce_virial_lists(lv->ptr, lv->ent, (const double *)tl->d_task_out.p,
                                                virial, lv->nlists, s)
                   : launch_reduce_virial((const double *)tl->d_task_out.p, tl->ntasks, virial, s));
  }
  end_levels(tl, s);
  count_stats(tl, false, ld[0] + ld[2]);
}

// Every matrix block addressed by the list must lie inside the buffer.
void check_buffer(const TaskListImpl *tl, const offload_buffer *b, const char *what) {
  if (!b) fatal("%s is NULL", what);
  if ((long long)(b->size / sizeof(double)) < tl->block_end)
    fatal("%s holds %zu doubles but the task list addresses %lld", what,
          b->size / sizeof(double), tl->block_end);
}

double *device_mirror(const offload_buffer *b, DevBuf &scratch) {
  if (b->device_buffer) return b->device_buffer;
  return (double *)scratch.get(b->size ? b->size : sizeof(double));
}

// Residency of offload_buffers (extension, include/grid_buffers.h): which
// copy holds the current data, and whether results the grid calls write into
// the buffer stay in HBM.  Buffers never named by the extension keep the
// reference behaviour: inputs are copied host -> device on every call and
// results device -> host.
struct Residency {
  bool resident = false;        // results stay on the device
  bool device_current = false;  // device_buffer holds the current data
};
std::mutex res_mu;
std::map<const offload_buffer *, Residency> res_tab;

Residency residency_of(const offload_buffer *b) {
  std::lock_guard<std::mutex> g(res_mu);
  auto it = res_tab.find(b);
  return it == res_tab.end() ? Residency{} : it->second;
}

void set_device_current(const offload_buffer *b, bool v) {
  std::lock_guard<std::mutex> g(res_mu);
  auto it = res_tab.find(b);
  if (it != res_tab.end()) {
    it->second.device_current = v;
  } else if (v) {
    res_tab[b].device_current = true;
  }
}

void forget_buffer(const offload_buffer *b) {
  std::lock_guard<std::mutex> g(res_mu);
  res_tab.erase(b);
}

// Device view of an input buffer: its device copy when that is current,
// else `bytes` copied host -> device into the buffer's mirror.
const double *input_view(const offload_buffer *b, DevBuf &scratch, size_t bytes, hipStream_t s) {
  if (b->device_buffer && residency_of(b).device_current) return b->device_buffer;
  double *d = device_mirror(b, scratch);
  HIP_CHECK(hipMemcpyAsync(d, b->host_buffer, bytes, hipMemcpyHostToDevice, s));
  return d;
}

// After a grid call wrote `bytes` of results at d (the buffer's device view):
// a resident buffer keeps them in HBM, any other gets them copied to host.
void output_done(offload_buffer *b, const double *d, size_t bytes, hipStream_t s) {
  if (b->device_buffer == d && residency_of(b).resident) {
    set_device_current(b, true);
    return;
  }
  HIP_CHECK(hipMemcpyAsync(b->host_buffer, d, bytes, hipMemcpyDeviceToHost, s));
  set_device_current(b, false);
}

// Zeroes a buffer the way its residency says (empty task lists: no device
// work unless the buffer is resident).
void zero_output(offload_buffer *b) {
  if (b->device_buffer && residency_of(b).resident) {
    const hipStream_t s = stream();
    HIP_CHECK(launch_fill(b->device_buffer, (long long)(b->size / sizeof(double)), 0.0, s));
    sync_stream(s);
    set_device_current(b, true);
    return;
  }
  memset(b->host_buffer, 0, b->size);
  set_device_current(b, false);
}
}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Batched task lists: the lists are merged into one list (levels, atoms,
// kinds, blocks and buffers concatenated) whose launches each cover every
// list, and a run replays a hipGraph of the merged calls.
// ---------------------------------------------------------------------------
namespace {

struct BatchImpl {
  static constexpr unsigned kMagic = 0x42415443u;
  unsigned magic = kMagic;
  int nlists = 0;
  TaskListImpl *merged = nullptr;
  std::vector<int> level_off, atom_off;  // per list (+ total)
  std::vector<long long> pab_off;        // per list (+ total), doubles
  std::vector<long long> grid_off;       // per merged level (+ total), doubles
  DevBuf pab, hab, rho, pot, forces, virial, vptr, vent;
  std::vector<double *> rho_lv, pot_lv;  // per merged level
  // A captured run with its own level tables (the FP32 potential pointer in
  // the integrate table depends on the run's precision).
  struct Graph {
    hipGraphExec_t exec = nullptr;
    DevBuf lev_col, lev_int;
  };
  std::map<std::array<int, 5>, Graph> graphs;  // (what, func, tau, forces, precision)
  void drop_graphs() {
    for (auto &kv : graphs) {
      if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
      kv.second.lev_col.release();
      kv.second.lev_int.release();
    }
    graphs.clear();
  }
  ~BatchImpl() {
    drop_graphs();
    DevBuf *bufs[] = {&pab, &hab, &rho, &pot, &forces, &virial, &vptr, &vent};
    for (DevBuf *b : bufs) b->release();
    delete merged;
    magic = 0;
  }
};

// Device buffers of the merged list that a captured graph addresses and that
// a call may (re)allocate on demand; a graph is valid while none has moved.
std::array<const void *, 7> captured_buffers(const TaskListImpl *tl) {
  return {tl->d_cxyz.p, tl->d_gridf.p, tl->d_cart.p, tl->d_hab_cart.p, tl->d_task_out.p,
          tl->d_gitems.p, tl->d_queue.p};
}

BatchImpl *checked_batch(const void *p) {
  BatchImpl *b = (BatchImpl *)p;
  if (!b || b->magic != BatchImpl::kMagic) fatal("invalid or freed batch handle");
  return b;
}

// Level table with fixed grid pointers, uploaded synchronously.
void static_levels(BatchImpl *b, DevBuf &dst, const std::vector<double *> &grids,
                   const float *gridf) {
  TaskListImpl *tl = b->merged;
  std::vector<Level> lv(tl->nlevels);
  long long off = 0;
  for (int l = 0; l < tl->nlevels; l++) {
    lv[l].lay = tl->layouts[l];
    lv[l].grid = grids[l];
    lv[l].npts = level_points(tl, l);
    lv[l].gridf = gridf ? gridf + off : nullptr;
    off += lv[l].npts;
  }
  upload(dst, lv);
}

// The merged calls of a run: per batch, collocate (what & 1) then integrate
// (what & 2).
void batch_enqueue(BatchImpl *b, int what, int func, bool tau, bool forces,
                   const BatchImpl::Graph *fixed, hipStream_t s) {
  TaskListImpl *tl = b->merged;
  double *pab = (double *)b->pab.p;
  if (what & 1) {
    tl->static_levels = fixed ? (const Level *)fixed->lev_col.p : nullptr;
    run_collocate(tl, func, pab, b->rho_lv.data(), s);
  }
  if (what & 2) {
    tl->static_levels = fixed ? (const Level *)fixed->lev_int.p : nullptr;
    const ListVirial lv{(const int *)b->vptr.p, (const int *)b->vent.p, b->nlists};
    run_integrate(tl, tau, forces ? pab : nullptr, b->pot_lv.data(), (double *)b->hab.p,
                  b->pab_off[b->nlists], forces ? (double *)b->forces.p : nullptr,
                  forces ? (double *)b->virial.p : nullptr, s, forces ? &lv : nullptr);
  }
  tl->static_levels = nullptr;
}

// ---------------------------------------------------------------------------
// Single-task entry points: the drop-ins for grid_cpu_collocate_pgf_product /
// grid_cpu_integrate_pgf_product, which src/grid/grid_api.F:110-490
// (collocate_pgf_product / integrate_pgf_product) binds.  One task runs on
// the per-task kernels; every array may live in host or device memory.
// ---------------------------------------------------------------------------
struct PgfContext {
  std::mutex mu;
  TaskListImpl tl;  // holds the maxima of the one task (LDS plan) only
  DevBuf rec, grid, pab, hab, out, hdab, hadb, a_hdab;
};

PgfContext &pgf_ctx() {
  static PgfContext *c = new PgfContext();  // lives until the process ends
  return *c;
}

bool device_pointer(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // plain host memory on some runtimes
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Device view of n doubles at p: p itself if it is device memory, else a
// copy in `mirror` (copied in when copy_in).
double *device_view(const double *p, size_t n, DevBuf &mirror, bool copy_in, hipStream_t s) {
  if (!p || device_pointer(p)) return const_cast<double *>(p);
  double *d = (double *)mirror.get(sizeof(double) * std::max<size_t>(n, 1));
  if (copy_in) HIP_CHECK(hipMemcpyAsync(d, p, sizeof(double) * n, hipMemcpyHostToDevice, s));
  return d;
}

void copy_back(double *dst, const double *view, size_t n, hipStream_t s) {
  if (dst && view != dst)
    HIP_CHECK(hipMemcpyAsync(dst, view, sizeof(double) * n, hipMemcpyDeviceToHost, s));
}

// dst[0..n) += src[0..n) (src on the host; dst host or device).
void add_into(double *dst, const double *src, int n, hipStream_t s) {
  if (!dst) return;
  double tmp[24];
  if (device_pointer(dst)) {
    HIP_CHECK(hipMemcpyAsync(tmp, dst, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < n; i++) tmp[i] += src[i];
    HIP_CHECK(hipMemcpyAsync(dst, tmp, sizeof(double) * n, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
  } else {
    for (int i = 0; i < n; i++) dst[i] += src[i];
  }
}

struct PgfArgs {
  bool orthorhombic;
  int border_mask, la_max, la_min, lb_max, lb_min;
  double zeta, zetb, rscale;
  const double (*dh)[3], (*dh_inv)[3];
  const double *ra, *rab;
  const int *npts_global, *npts_local, *shift_local, *border_width;
  double radius;
  int o1, o2, n1, n2;
};

// Task, group and level records of the one task (uploaded to c.rec), the
// task's kind (kTaskSkip, kTaskOrtho or general) and the LDS-plan maxima.
int pgf_records(PgfContext &c, const PgfArgs &a, Level &L, hipStream_t s, const Task **dtask,
                const Group **dgroup, const Level **dlevel) {
  if (a.la_max < 0 || a.lb_max < 0 || a.la_min < 0 || a.lb_min < 0 || a.la_min > a.la_max ||
      a.lb_min > a.lb_max)
    fatal("pgf_product: bad l range a %d..%d, b %d..%d", a.la_min, a.la_max, a.lb_min, a.lb_max);
  if (a.o1 < 0 || a.o2 < 0 || a.o1 + ncoset(a.la_max) > a.n1 || a.o2 + ncoset(a.lb_max) > a.n2)
    fatal("pgf_product: block [%d][%d] does not hold offsets o1=%d, o2=%d for la_max=%d, "
          "lb_max=%d", a.n2, a.n1, a.o1, a.o2, a.la_max, a.lb_max);
  Layout &y = L.lay;
  for (int i = 0; i < 3; i++) {
    y.npts_global[i] = a.npts_global[i];
    y.npts_local[i] = a.npts_local[i];
    y.shift_local[i] = a.shift_local[i];
    y.border_width[i] = a.border_width[i];
    if (y.npts_local[i] <= 0 || y.npts_global[i] <= 0) fatal("pgf_product: non-positive grid size");
    for (int j = 0; j < 3; j++) {
      y.dh[i][j] = a.dh[i][j];
      y.dh_inv[i][j] = a.dh_inv[i][j];
    }
  }
  L.npts = (long long)y.npts_local[0] * y.npts_local[1] * y.npts_local[2];
  L.gridf = nullptr;
  Task T;
  memset(&T, 0, sizeof T);
  T.la_max = a.la_max;
  T.la_min = a.la_min;
  T.lb_max = a.lb_max;
  T.lb_min = a.lb_min;
  T.o1 = a.o1;
  T.o2 = a.o2;
  T.zeta = a.zeta;
  T.zetb = a.zetb;
  T.rscale = a.rscale;
  T.radius = a.radius;
  T.geo = -1;
  for (int d = 0; d < 3; d++) {
    T.ra[d] = a.ra[d];
    T.rab[d] = a.rab[d];
  }
  const int kind = task_geometry(T, y, a.orthorhombic && a.border_mask == 0, a.border_mask, 1);
  TaskListImpl &tl = c.tl;
  tl.max_la = a.la_max;
  tl.max_lb = a.lb_max;
  tl.max_lp = a.la_max + a.lb_max;
  tl.has_ortho = kind == kTaskOrtho;
  tl.has_general = kind == 0;
  tl.max_cmax = tl.max_rows_ortho = tl.max_range = tl.max_rows_general = tl.max_planes = 0;
  if (kind == kTaskOrtho) {
    tl.max_cmax = T.cmax;
    tl.max_rows_ortho = 2 - 2 * T.lo[1];
    tl.max_planes = 2 - 2 * T.lo[2];
  } else if (kind == 0) {
    for (int d = 0; d < 3; d++) tl.max_range = std::max(tl.max_range, T.hi[d] - T.lo[d] + 1);
    tl.max_rows_general = T.hi[1] - T.lo[1] + 1;
    tl.max_planes = T.hi[2] - T.lo[2] + 1;
  }
  Group G;
  memset(&G, 0, sizeof G);
  G.ncoa = a.n1;
  G.ncob = a.n2;
  struct alignas(16) Rec {
    Task t;
    Group g;
    Level l;
  } r;
  r.t = T;
  r.g = G;
  r.l = L;
  Rec *d = (Rec *)c.rec.get(sizeof(Rec));
  // every single-task call ends synchronised, so no kernel still reads d
  HIP_CHECK(hipMemcpy(d, &r, sizeof(Rec), hipMemcpyHostToDevice));
  (void)s;
  *dtask = &d->t;
  *dgroup = &d->g;
  *dlevel = &d->l;
  return kind;
}

}  // namespace

extern "C" {

// ---- offload_buffer, src/offload/offload_buffer.c:52-118 ----
void offload_create_buffer(const int length, offload_buffer **buffer) {
  const size_t requested = sizeof(double) * (size_t)(length > 0 ? length : 0);
  if (*buffer != nullptr) {
    if (requested == 0) {
      if ((*buffer)->host_buffer) (void)hipHostFree((*buffer)->host_buffer);
      if ((*buffer)->device_buffer) (void)hipFree((*buffer)->device_buffer);
      (*buffer)->size = 0;
      (*buffer)->host_buffer = nullptr;
      (*buffer)->device_buffer = nullptr;
      return;
    }
    if (requested <= (*buffer)->size) return;
    set_device_current(*buffer, false);  // new, uninitialised memory
    if ((*buffer)->host_buffer) (void)hipHostFree((*buffer)->host_buffer);
    if ((*buffer)->device_buffer) (void)hipFree((*buffer)->device_buffer);
  } else {
    *buffer = (offload_buffer *)malloc(sizeof(offload_buffer));
    if (!*buffer) fatal("out of host memory");
  }
  (*buffer)->size = requested;
  (*buffer)->host_buffer = nullptr;
  (*buffer)->device_buffer = nullptr;
  if (requested == 0) return;
  ensure_device();
  HIP_CHECK(hipHostMalloc((void **)&(*buffer)->host_buffer, requested));
  HIP_CHECK(hipMalloc((void **)&(*buffer)->device_buffer, requested));
}

void offload_free_buffer(offload_buffer *buffer) {
  if (!buffer) return;
  forget_buffer(buffer);
  if (buffer->host_buffer) (void)hipHostFree(buffer->host_buffer);
  if (buffer->device_buffer) (void)hipFree(buffer->device_buffer);
  free(buffer);
}

double *offload_get_buffer_host_pointer(offload_buffer *buffer) {
  return buffer ? buffer->host_buffer : nullptr;
}

// ---- basis sets, src/grid/common/grid_basis_set.c:19-86 ----
void grid_create_basis_set(const int nset, const int nsgf, const int maxco, const int maxpgf,
                           const int *lmin, const int *lmax, const int *npgf,
                           const int *nsgf_set, const int *first_sgf, const double *sphi,
                           const double *zet, grid_basis_set **basis_set) {
  grid_basis_set *b = (grid_basis_set *)malloc(sizeof(grid_basis_set));
  if (!b) fatal("out of host memory");
  b->nset = nset;
  b->nsgf = nsgf;
  b->maxco = maxco;
  b->maxpgf = maxpgf;
  auto icopy = [&](const int *src) {
    int *d = (int *)malloc(sizeof(int) * (nset > 0 ? nset : 1));
    if (nset > 0) memcpy(d, src, sizeof(int) * nset);
    return d;
  };
  b->lmin = icopy(lmin);
  b->lmax = icopy(lmax);
  b->npgf = icopy(npgf);
  b->nsgf_set = icopy(nsgf_set);
  b->first_sgf = icopy(first_sgf);
  const size_t ns = (size_t)nsgf * maxco, nz = (size_t)nset * maxpgf;
  b->sphi = (double *)malloc(sizeof(double) * (ns ? ns : 1));
  if (ns) memcpy(b->sphi, sphi, sizeof(double) * ns);
  b->zet = (double *)malloc(sizeof(double) * (nz ? nz : 1));
  if (nz) memcpy(b->zet, zet, sizeof(double) * nz);
  *basis_set = b;
}

void grid_free_basis_set(grid_basis_set *b) {
  if (!b) return;
  free(b->lmin);
  free(b->lmax);
  free(b->npgf);
  free(b->nsgf_set);
  free(b->first_sgf);
  free(b->sphi);
  free(b->zet);
  free(b);
}

// ---- library, src/grid/common/grid_library.c:53-257 ----
void grid_library_init(void) {
  std::lock_guard<std::mutex> g(S.mu);
  if (S.initialized) {
    printf("Error: Grid library was already initialized.\n");
    abort();
  }
  memset(S.counters, 0, sizeof(S.counters));
  S.initialized = true;
}

void grid_library_finalize(void) {
  std::lock_guard<std::mutex> g(S.mu);
  if (!S.initialized) {
    printf("Error: Grid library is not initialized.\n");
    abort();
  }
  S.initialized = false;
}

void grid_library_set_config(const enum grid_backend backend, const bool validate,
                             const bool apply_cutoff) {
  S.config.backend = backend;
  S.config.validate = validate;
  S.config.apply_cutoff = apply_cutoff;
}

grid_library_config grid_library_get_config(void) { return S.config; }

void grid_library_counter_add(const int lp, const enum grid_backend backend,
                              const enum grid_library_kernel kern, const int increment) {
  if (lp < 0 || (int)kern < 0 || (int)kern > 3) fatal("bad counter index");
  (void)backend;  // every counter of this library belongs to the GPU backend
  std::lock_guard<std::mutex> g(S.mu);
  S.counters[kern][std::min(lp, kCounterLp - 1)] += increment;
}

void grid_library_print_stats(const int fortran_comm, void (*print_func)(const char *, int, int),
                              int output_unit) {
  (void)fortran_comm;
  if (!S.initialized) {
    printf("Error: Grid library is not initialized.\n");
    abort();
  }
  struct Row {
    long n;
    int kern, lp;
  };
  std::vector<Row> rows;
  double total = 0.0;
  {
    std::lock_guard<std::mutex> g(S.mu);
    for (int k = 0; k < 4; k++)
      for (int l = 0; l < kCounterLp; l++)
        if (S.counters[k][l]) {
          rows.push_back({S.counters[k][l], k, l});
          total += S.counters[k][l];
        }
  }
  if (rows.empty()) return;
  std::stable_sort(rows.begin(), rows.end(), [](const Row &a, const Row &b) { return a.n > b.n; });
  auto out = [&](const char *m) { print_func(m, (int)strlen(m), output_unit); };
  const char *rule =
      " -------------------------------------------------------------------------------\n";
  out("\n");
  out(rule);
  out(" -                                                                             -\n");
  out(" -                                GRID STATISTICS                              -\n");
  out(" -                                                                             -\n");
  out(rule);
  out(" LP    KERNEL             BACKEND                              COUNT     PERCENT\n");
  const char *names[] = {"collocate ortho", "integrate ortho", "collocate general",
                         "integrate general"};
  for (const Row &r : rows) {
    char buf[128];
    snprintf(buf, sizeof(buf), " %-5i %-17s  %-6s  %34li %10.2f%%\n", r.lp, names[r.kern],
             "GPU", r.n, 100.0 * r.n / total);
    out(buf);
  }
  out(rule);
}

// ---- task list, src/grid/grid_task_list.c:23-436 ----
void grid_create_task_list(
    const bool orthorhombic, const int ntasks, const int nlevels, const int natoms,
    const int nkinds, const int nblocks, const int *block_offsets,
    const double (*atom_positions)[3], const int *atom_kinds,
    const grid_basis_set **basis_sets, const int *level_list, const int *iatom_list,
    const int *jatom_list, const int *iset_list, const int *jset_list, const int *ipgf_list,
    const int *jpgf_list, const int *border_mask_list, const int *block_num_list,
    const double *radius_list, const double (*rab_list)[3], const int (*npts_global)[3],
    const int (*npts_local)[3], const int (*shift_local)[3], const int (*border_width)[3],
    const double (*dh)[3][3], const double (*dh_inv)[3][3], grid_task_list **task_list) {
  grid_gpu_create_task_list(orthorhombic, ntasks, nlevels, natoms, nkinds, nblocks,
                            block_offsets, (const double *)atom_positions, atom_kinds,
                            basis_sets, level_list, iatom_list, jatom_list, iset_list,
                            jset_list, ipgf_list, jpgf_list, border_mask_list, block_num_list,
                            radius_list, (const double *)rab_list, (const int *)npts_global,
                            (const int *)npts_local, (const int *)shift_local,
                            (const int *)border_width, (const double *)dh,
                            (const double *)dh_inv, (grid_gpu_task_list **)task_list);
}

void grid_free_task_list(grid_task_list *task_list) { grid_gpu_free_task_list(task_list); }

static void check_npts(const TaskListImpl *tl, int nlevels, const int *npts_local) {
  if (tl->nlevels != nlevels) fatal("nlevels %d does not match the task list (%d)", nlevels, tl->nlevels);
  for (int i = 0; i < 3 * nlevels; i++)
    if (tl->npts_local[i] != npts_local[i])
      fatal("npts_local of level %d does not match the task list", i / 3 + 1);
}

void grid_collocate_task_list(const grid_task_list *task_list, const enum grid_func func,
                              const int nlevels, const int (*npts_local)[3],
                              const offload_buffer *pab_blocks, offload_buffer **grids) {
  if (!task_list) return;
  TaskListImpl *tl = checked(task_list);
  if (tl->empty) {  // src/grid/grid_task_list.c:185-189
    for (int l = 0; l < nlevels; l++) zero_output(grids[l]);
    return;
  }
  check_npts(tl, nlevels, (const int *)npts_local);
  grid_gpu_collocate_task_list(task_list, func, nlevels, pab_blocks, grids);
}

void grid_integrate_task_list(const grid_task_list *task_list, const bool compute_tau,
                              const int natoms, const int nlevels, const int (*npts_local)[3],
                              const offload_buffer *pab_blocks, const offload_buffer **grids,
                              offload_buffer *hab_blocks, double (*forces)[3],
                              double (*virial)[3]) {
  if (!task_list) return;
  TaskListImpl *tl = checked(task_list);
  if (tl->empty) {  // src/grid/grid_task_list.c:290-311
    zero_output(hab_blocks);
    if (virial) memset(virial, 0, 9 * sizeof(double));
    if (forces) memset(forces, 0, 3 * sizeof(double) * (size_t)natoms);
    return;
  }
  check_npts(tl, nlevels, (const int *)npts_local);
  if (tl->natoms != natoms) fatal("natoms %d does not match the task list (%d)", natoms, tl->natoms);
  grid_gpu_integrate_task_list(task_list, compute_tau, nlevels, pab_blocks, grids, hab_blocks,
                               (double *)forces, (double *)virial);
}

// ---- backend level, src/grid/gpu/grid_gpu_task_list.h:25-59 ----
void grid_gpu_create_task_list(
    const bool ortho, const int ntasks, const int nlevels, const int natoms, const int nkinds,
    const int nblocks, const int *block_offsets, const double *atom_positions,
    const int *atom_kinds, const grid_basis_set **basis_sets, const int *level_list,
    const int *iatom_list, const int *jatom_list, const int *iset_list, const int *jset_list,
    const int *ipgf_list, const int *jpgf_list, const int *border_mask_list,
    const int *block_num_list, const double *radius_list, const double *rab_list,
    const int *npts_global, const int *npts_local, const int *shift_local,
    const int *border_width, const double *dh, const double *dh_inv,
    grid_gpu_task_list **ptr) {
  TaskListImpl *tl;
  if (*ptr) {
    // reuse the handle (src/grid/grid_task_list.c:42-63): rebuild in place
    TaskListImpl *old = checked(*ptr);
    old->~TaskListImpl();
    tl = new (old) TaskListImpl();
  } else {
    tl = new TaskListImpl();
  }
  build_task_list(tl, ortho, ntasks, nlevels, natoms, nkinds, nblocks, block_offsets,
                  atom_positions, atom_kinds, basis_sets, level_list, iatom_list, jatom_list,
                  iset_list, jset_list, ipgf_list, jpgf_list, border_mask_list, block_num_list,
                  radius_list, rab_list, npts_global, npts_local, shift_local, border_width, dh,
                  dh_inv);
  *ptr = tl;
}

void grid_gpu_free_task_list(grid_gpu_task_list *ptr) {
  if (!ptr) return;
  TaskListImpl *tl = checked(ptr);
  delete tl;
}

void grid_gpu_collocate_task_list(const grid_gpu_task_list *ptr, const enum grid_func func,
                                  const int nlevels, const offload_buffer *pab_blocks,
                                  offload_buffer **grids) {
  TaskListImpl *tl = checked(ptr);
  if (tl->empty) {
    for (int l = 0; l < nlevels; l++) zero_output(grids[l]);
    return;
  }
  if (nlevels != tl->nlevels) fatal("nlevels does not match the task list");
  const hipStream_t s = stream();
  check_buffer(tl, pab_blocks, "pab_blocks");
  const double *pab = input_view(pab_blocks, tl->pab_dev, pab_blocks->size, s);
  std::vector<double *> g(nlevels);
  for (int l = 0; l < nlevels; l++) {
    const size_t need = sizeof(double) * (size_t)tl->npts_local[3 * l] *
                        tl->npts_local[3 * l + 1] * tl->npts_local[3 * l + 2];
    if (grids[l]->size < need) fatal("grid buffer of level %d is too small", l + 1);
    g[l] = device_mirror(grids[l], tl->grid_dev[l]);
  }
  run_collocate(tl, func, pab, g.data(), s);
  for (int l = 0; l < nlevels; l++) output_done(grids[l], g[l], grids[l]->size, s);
  sync_stream(s);
}

void grid_gpu_integrate_task_list(const grid_gpu_task_list *ptr, const bool compute_tau,
                                  const int nlevels, const offload_buffer *pab_blocks,
                                  const offload_buffer **grids, offload_buffer *hab_blocks,
                                  double *forces, double *virial) {
  TaskListImpl *tl = checked(ptr);
  if (tl->empty) {
    zero_output(hab_blocks);
    if (virial) memset(virial, 0, 9 * sizeof(double));
    if (forces) memset(forces, 0, 3 * sizeof(double) * (size_t)tl->natoms);
    return;
  }
  if (nlevels != tl->nlevels) fatal("nlevels does not match the task list");
  const hipStream_t s = stream();
  std::vector<const double *> g(nlevels);
  for (int l = 0; l < nlevels; l++) {
    const size_t need = sizeof(double) * (size_t)tl->npts_local[3 * l] *
                        tl->npts_local[3 * l + 1] * tl->npts_local[3 * l + 2];
    if (grids[l]->size < need) fatal("grid buffer of level %d is too small", l + 1);
    g[l] = input_view(grids[l], tl->grid_dev[l], need, s);
  }
  const double *pab = nullptr;
  check_buffer(tl, hab_blocks, "hab_blocks");
  if (pab_blocks && (forces || virial)) {
    check_buffer(tl, pab_blocks, "pab_blocks");
    pab = input_view(pab_blocks, tl->pab_dev, pab_blocks->size, s);
  }
  double *hab = device_mirror(hab_blocks, tl->hab_dev);
  const long long hab_len = (long long)(hab_blocks->size / sizeof(double));
  double *fv = (double *)tl->forces_dev.get(sizeof(double) * (3 * (size_t)tl->natoms + 9));
  run_integrate(tl, compute_tau, pab, g.data(), hab, hab_len, forces ? fv : nullptr,
                virial ? fv + 3 * tl->natoms : nullptr, s);
  output_done(hab_blocks, hab, hab_blocks->size, s);
  if (forces)
    HIP_CHECK(hipMemcpyAsync(forces, fv, sizeof(double) * 3 * tl->natoms,
                             hipMemcpyDeviceToHost, s));
  if (virial)
    HIP_CHECK(hipMemcpyAsync(virial, fv + 3 * tl->natoms, sizeof(double) * 9,
                             hipMemcpyDeviceToHost, s));
  sync_stream(s);
}

// ---- extensions, include/grid_mi355x.h ----
void grid_mi355x_set_precision(int p) {
  if (p != GRID_MI355X_FP64 && p != GRID_MI355X_FP32_MIXED) fatal("unknown precision %d", p);
  S.precision = p;
}

int grid_mi355x_get_precision(void) { return precision(); }

void grid_mi355x_collocate_task_list_dev(const grid_task_list *task_list,
                                         const enum grid_func func, const int nlevels,
                                         const double *pab_blocks, double *const *grids) {
  TaskListImpl *tl = checked(task_list);
  const hipStream_t s = stream();
  if (tl->empty) {
    for (int l = 0; l < nlevels && l < (int)tl->npts_local.size() / 3; l++)
      HIP_CHECK(launch_fill(grids[l], (long long)tl->npts_local[3 * l] * tl->npts_local[3 * l + 1] *
                                          tl->npts_local[3 * l + 2], 0.0, s));
    return;
  }
  if (nlevels != tl->nlevels) fatal("nlevels does not match the task list");
  run_collocate(tl, func, pab_blocks, grids, s);
}

// No reference counterpart: grid_mi355x_collocate_task_list_dev without any
// device allocation, for callers that cannot allocate (graph capture, tight
// serving loops).  Per-call buffers come from the library's static scratch
// array (128 MiB per device): the level table, the decontracted density
// blocks and the per-task cxyz coefficients.  Lists that need more are
// refused with a message; calls that overlap on one device race.
void grid_mi355x_collocate_task_list_noalloc(const grid_task_list *task_list,
                                             const enum grid_func func, const int nlevels,
                                             const double *pab_blocks, double *const *grids) {
  TaskListImpl *tl = checked(task_list);
  if (nlevels != tl->nlevels) fatal("nlevels does not match the task list");
  if (tl->empty) {
    fatal("grid_mi355x_collocate_task_list_noalloc: the list is empty (zero the grids "
          "with the caller's own means)");
  }
  if (tl->merged) fatal("grid_mi355x_collocate_task_list_noalloc: not for batched lists");
  int ld[4];
  if (!func_ldiffs(func, ld)) fatal("unknown grid_func %d", (int)func);
  const int sh = ld[0] + ld[2];
  if ((!tl->tiles.empty() && !tl->tile_plans.count(sh)) ||
      (!tl->gtiles.empty() && !tl->gtile_plans.count(sh)))
    fatal("grid_mi355x_collocate_task_list_noalloc: no tile plan for l shift %d", sh);
  hipError_t err;
  long long cap = 0;
  double *base = noalloc_scratch(&cap, &err);
  if (!base) fatal("scratch array address: %s", hipGetErrorString(err));
  // layout: level table, decontracted blocks, cxyz (256-byte aligned parts)
  auto up = [](long long n) { return (n + 31) & ~31LL; };
  const long long n_lev = up(((long long)sizeof(Level) * tl->nlevels + 7) / 8);
  const long long n_cart = up(std::max(tl->cart_len, 1LL));
  const long long n_cxyz = up((long long)cxyz_stride(tl, sh) * std::max(tl->ntasks, 1));
  if (n_lev + n_cart + n_cxyz > cap)
    fatal("grid_mi355x_collocate_task_list_noalloc: the list needs %lld doubles of scratch "
          "(%d tasks), the maximum is %lld (128 MiB); use "
          "grid_mi355x_collocate_task_list_dev",
          n_lev + n_cart + n_cxyz, tl->ntasks, cap);
  CallScratch cs;
  cs.levels = (const Level *)base;
  cs.cart = base + n_lev;
  cs.cxyz = base + n_lev + n_cart;
  const hipStream_t s = stream();
  constexpr int kMaxNoAllocLevels = 32;
  if (tl->nlevels > kMaxNoAllocLevels)
    fatal("grid_mi355x_collocate_task_list_noalloc: at most %d grid levels", kMaxNoAllocLevels);
  Level lv[kMaxNoAllocLevels];
  for (int l = 0; l < tl->nlevels; l++) {
    lv[l].lay = tl->layouts[l];
    lv[l].grid = grids[l];
    lv[l].gridf = nullptr;
    lv[l].npts = level_points(tl, l);
  }
  HIP_CHECK(launch_write_levels(lv, tl->nlevels, (Level *)base, s));
  S.noalloc = true;
  run_collocate(tl, func, pab_blocks, grids, s, &cs);
  S.noalloc = false;
}

void grid_mi355x_integrate_task_list_dev(const grid_task_list *task_list, const bool compute_tau,
                                         const int natoms, const int nlevels,
                                         const double *pab_blocks, const double *const *grids,
                                         double *hab_blocks, const long hab_length,
                                         double *forces, double *virial) {
  TaskListImpl *tl = checked(task_list);
  const hipStream_t s = stream();
  if (tl->empty) {
    HIP_CHECK(launch_fill(hab_blocks, hab_length, 0.0, s));
    if (forces) HIP_CHECK(launch_fill(forces, 3LL * natoms, 0.0, s));
    if (virial) HIP_CHECK(launch_fill(virial, 9, 0.0, s));
    return;
  }
  if (nlevels != tl->nlevels) fatal("nlevels does not match the task list");
  if (natoms != tl->natoms) fatal("natoms does not match the task list");
  if (hab_length < tl->block_end)
    fatal("hab_blocks holds %ld doubles but the task list addresses %lld", hab_length,
          tl->block_end);
  run_integrate(tl, compute_tau, pab_blocks, grids, hab_blocks, hab_length, forces, virial, s);
}

// ---- single task, src/grid/cpu/grid_cpu_collocate.h / grid_cpu_integrate.h
//      (bound by src/grid/grid_api.F:138-172, 311-350) ----
void grid_mi355x_collocate_pgf_product(
    const bool orthorhombic, const int border_mask, const enum grid_func func, const int la_max,
    const int la_min, const int lb_max, const int lb_min, const double zeta, const double zetb,
    const double rscale, const double dh[3][3], const double dh_inv[3][3], const double ra[3],
    const double rab[3], const int npts_global[3], const int npts_local[3],
    const int shift_local[3], const int border_width[3], const double radius, const int o1,
    const int o2, const int n1, const int n2, const double *pab, double *grid) {
  ensure_device();
  PgfContext &c = pgf_ctx();
  std::lock_guard<std::mutex> g(c.mu);
  int ld[4];
  if (!func_ldiffs(func, ld)) fatal("unknown grid_func %d", (int)func);
  const hipStream_t s = stream();
  const PgfArgs a{orthorhombic, border_mask, la_max, la_min, lb_max, lb_min, zeta, zetb, rscale,
                  dh, dh_inv, ra, rab, npts_global, npts_local, shift_local, border_width,
                  radius, o1, o2, n1, n2};
  Level L;
  const size_t npts = (size_t)npts_local[0] * npts_local[1] * npts_local[2];
  double *gd = device_view(grid, npts, c.grid, true, s);
  L.grid = gd;
  const Task *dt;
  const Group *dg;
  const Level *dl;
  const int kind = pgf_records(c, a, L, s, &dt, &dg, &dl);
  if (kind != kTaskSkip) {
    PointParams P{};
    P.tasks = dt;
    P.ntasks = 1;
    P.groups = dg;
    P.levels = dl;
    P.cart = device_view(pab, (size_t)n1 * n2, c.pab, true, s);
    P.cart_len = (long long)n1 * n2;
    P.func = func;
    for (int i = 0; i < 4; i++) P.ld[i] = ld[i];
    P.fp32 = precision() == GRID_MI355X_FP32_MIXED;
    P.plan = make_plan(&c.tl, ld, false);
    HIP_CHECK(launch_collocate(P, kind == kTaskOrtho, s));
  }
  copy_back(grid, gd, npts, s);
  HIP_CHECK(hipStreamSynchronize(s));
}

void grid_mi355x_integrate_pgf_product(
    const bool orthorhombic, const bool compute_tau, const int border_mask, const int la_max,
    const int la_min, const int lb_max, const int lb_min, const double zeta, const double zetb,
    const double dh[3][3], const double dh_inv[3][3], const double ra[3], const double rab[3],
    const int npts_global[3], const int npts_local[3], const int shift_local[3],
    const int border_width[3], const double radius, const int o1, const int o2, const int n1,
    const int n2, const double *grid, double *hab, const double *pab, double *forces,
    double *virials, double *hdab, double *hadb, double *a_hdab) {
  // src/grid/ref/grid_ref_integrate.c:56-60
  const bool calc_forces = forces != nullptr || hdab != nullptr;
  const bool calc_virial = virials != nullptr || a_hdab != nullptr;
  if (calc_virial && !calc_forces) fatal("integrate_pgf_product: virial requires forces");
  if ((hdab || hadb || a_hdab) && compute_tau)
    fatal("integrate_pgf_product: hdab, hadb and a_hdab are not defined with compute_tau");
  if ((forces || virials) && !pab) fatal("integrate_pgf_product: forces/virials need pab");
  ensure_device();
  PgfContext &c = pgf_ctx();
  std::lock_guard<std::mutex> g(c.mu);
  const hipStream_t s = stream();
  const PgfArgs a{orthorhombic, border_mask, la_max, la_min, lb_max, lb_min, zeta, zetb, 1.0,
                  dh, dh_inv, ra, rab, npts_global, npts_local, shift_local, border_width,
                  radius, o1, o2, n1, n2};
  const size_t npts = (size_t)npts_local[0] * npts_local[1] * npts_local[2];
  const size_t nab = (size_t)n1 * n2;
  Level L;
  L.grid = device_view(grid, npts, c.grid, true, s);
  const Task *dt;
  const Group *dg;
  const Level *dl;
  const int kind = pgf_records(c, a, L, s, &dt, &dg, &dl);
  if (kind == kTaskSkip) {
    HIP_CHECK(hipStreamSynchronize(s));
    return;  // no point of the grid: hab, forces and derivatives unchanged
  }
  int ld[4];
  integrate_ldiffs(calc_forces, calc_virial, compute_tau, ld);
  PointParams P{};
  P.tasks = dt;
  P.ntasks = 1;
  P.groups = dg;
  P.levels = dl;
  P.cart = calc_forces ? device_view(pab, nab, c.pab, true, s) : nullptr;
  P.cart_len = (long long)nab;
  P.hab_cart = device_view(hab, nab, c.hab, true, s);
  for (int i = 0; i < 4; i++) P.ld[i] = ld[i];
  P.compute_tau = compute_tau;
  P.do_forces = calc_forces;
  P.do_virial = calc_virial;
  P.fp32 = precision() == GRID_MI355X_FP32_MIXED;
  P.plan = make_plan(&c.tl, ld, true);
  double *out = (double *)c.out.get(sizeof(double) * (24 + kTaskOutStride));
  P.pgf_out = out;
  P.task_out = out + 24;
  P.hdab = device_view(hdab, 3 * nab, c.hdab, true, s);
  P.hadb = device_view(hadb, 3 * nab, c.hadb, true, s);
  P.a_hdab = device_view(a_hdab, 9 * nab, c.a_hdab, true, s);
  HIP_CHECK(launch_fill(out, 24 + kTaskOutStride, 0.0, s));
  HIP_CHECK(launch_integrate(P, kind == kTaskOrtho, s));
  copy_back(hab, P.hab_cart, nab, s);
  copy_back(hdab, P.hdab, 3 * nab, s);
  copy_back(hadb, P.hadb, 3 * nab, s);
  copy_back(a_hdab, P.a_hdab, 9 * nab, s);
  double res[24];
  HIP_CHECK(hipMemcpyAsync(res, out, sizeof res, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (forces) add_into(forces, res, 6, s);
  if (virials) add_into(virials, res + 6, 18, s);
}

// ---- buffer packing and residency, src/task_list_methods.F:2561-2674 ----
void grid_mi355x_pack_buffer(const int npairs, const double *const *blocks,
                             const int *block_sizes, const int *pair_offsets,
                             offload_buffer *dest) {
  if (npairs <= 0) return;
  if (!dest || !blocks || !block_sizes || !pair_offsets) fatal("pack_buffer: NULL argument");
  const long long cap = (long long)(dest->size / sizeof(double));
  std::vector<long long> off(npairs);
  for (int i = 0; i < npairs; i++) {
    if (block_sizes[i] < 0 || pair_offsets[i] < 0 ||
        (long long)pair_offsets[i] + block_sizes[i] > cap)
      fatal("pack_buffer: pair %d (offset %d, %d doubles) outside the buffer of %lld", i + 1,
            pair_offsets[i], block_sizes[i], cap);
    off[i] = pair_offsets[i];
  }
  const hipStream_t s = stream();
  if (!device_pointer(blocks[0]) || !device_pointer(blocks[npairs - 1])) {
    // host blocks: rs_pack_buffer as the reference does it, into host_buffer
    for (int i = 0; i < npairs; i++)
      memcpy(dest->host_buffer + off[i], blocks[i], sizeof(double) * (size_t)block_sizes[i]);
    set_device_current(dest, false);
    return;
  }
  if (!dest->device_buffer) fatal("pack_buffer: the buffer has no device copy");
  ensure_device();
  static DevBuf d_ptr, d_size, d_off;
  static std::mutex mu;
  std::lock_guard<std::mutex> g(mu);
  upload(d_ptr, std::vector<const double *>(blocks, blocks + npairs));
  upload(d_size, std::vector<int>(block_sizes, block_sizes + npairs));
  upload(d_off, off);
  HIP_CHECK(launch_pack((const double *const *)d_ptr.p, (const int *)d_size.p,
                        (const long long *)d_off.p, npairs, dest->device_buffer, s));
  sync_stream(s);
  set_device_current(dest, true);
}

void grid_mi355x_unpack_buffer(const int npairs, double *const *blocks, const int *block_sizes,
                               const int *pair_offsets, const offload_buffer *src) {
  if (npairs <= 0) return;
  if (!src || !blocks || !block_sizes || !pair_offsets) fatal("unpack_buffer: NULL argument");
  const long long cap = (long long)(src->size / sizeof(double));
  std::vector<long long> off(npairs);
  for (int i = 0; i < npairs; i++) {
    if (block_sizes[i] < 0 || pair_offsets[i] < 0 ||
        (long long)pair_offsets[i] + block_sizes[i] > cap)
      fatal("unpack_buffer: pair %d (offset %d, %d doubles) outside the buffer of %lld", i + 1,
            pair_offsets[i], block_sizes[i], cap);
    off[i] = pair_offsets[i];
  }
  const hipStream_t s = stream();
  const bool dev_current = src->device_buffer && residency_of(src).device_current;
  if (!device_pointer(blocks[0]) || !device_pointer(blocks[npairs - 1])) {
    // host blocks: rs_unpack_buffer on the host copy (brought up to date first)
    if (dev_current) {
      HIP_CHECK(hipMemcpyAsync(src->host_buffer, src->device_buffer, src->size,
                               hipMemcpyDeviceToHost, s));
      sync_stream(s);
    }
    for (int i = 0; i < npairs; i++) {
      const double *p = src->host_buffer + off[i];
      for (int e = 0; e < block_sizes[i]; e++) blocks[i][e] += p[e];
    }
    return;
  }
  ensure_device();
  static DevBuf d_ptr, d_size, d_off, d_src;
  static std::mutex mu;
  std::lock_guard<std::mutex> g(mu);
  const double *view = dev_current ? src->device_buffer : nullptr;
  if (!view) {
    double *d = (double *)d_src.get(src->size ? src->size : sizeof(double));
    HIP_CHECK(hipMemcpyAsync(d, src->host_buffer, src->size, hipMemcpyHostToDevice, s));
    view = d;
  }
  upload(d_ptr, std::vector<double *>(blocks, blocks + npairs));
  upload(d_size, std::vector<int>(block_sizes, block_sizes + npairs));
  upload(d_off, off);
  HIP_CHECK(launch_unpack((double *const *)d_ptr.p, (const int *)d_size.p,
                          (const long long *)d_off.p, npairs, view, s));
  sync_stream(s);
}

void grid_mi355x_buffer_set_resident(offload_buffer *buffer, const bool resident) {
  if (!buffer) fatal("buffer_set_resident: NULL buffer");
  if (resident && !buffer->device_buffer)
    fatal("buffer_set_resident: the buffer has no device copy (created by another "
          "offload_create_buffer?)");
  std::lock_guard<std::mutex> g(res_mu);
  res_tab[buffer].resident = resident;
}

void grid_mi355x_buffer_sync_host(offload_buffer *buffer) {
  if (!buffer) fatal("buffer_sync_host: NULL buffer");
  if (!buffer->device_buffer || !residency_of(buffer).device_current) return;
  const hipStream_t s = stream();
  HIP_CHECK(hipMemcpyAsync(buffer->host_buffer, buffer->device_buffer, buffer->size,
                           hipMemcpyDeviceToHost, s));
  sync_stream(s);
}

void grid_mi355x_buffer_host_written(offload_buffer *buffer) {
  if (!buffer) fatal("buffer_host_written: NULL buffer");
  set_device_current(buffer, false);
}

int grid_mi355x_buffer_device_current(const offload_buffer *buffer) {
  return buffer && buffer->device_buffer && residency_of(buffer).device_current ? 1 : 0;
}

void *grid_mi355x_device_malloc(const size_t bytes) {
  ensure_device();
  void *p = nullptr;
  HIP_CHECK(hipMalloc(&p, bytes ? bytes : 8));
  return p;
}

void grid_mi355x_device_free(void *ptr) {
  if (ptr) HIP_CHECK(hipFree(ptr));
}

void grid_mi355x_memcpy(void *dst, const void *src, const size_t bytes, const int kind) {
  hipMemcpyKind k;
  switch (kind) {
    case 1: k = hipMemcpyHostToDevice; break;
    case 2: k = hipMemcpyDeviceToHost; break;
    case 3: k = hipMemcpyDeviceToDevice; break;
    default: fatal("unknown copy kind %d", kind);
  }
  const hipStream_t s = stream();
  HIP_CHECK(hipMemcpyAsync(dst, src, bytes, k, s));
  sync_stream(s);
}

void grid_mi355x_set_stream(void *hip_stream) {
  S.user_stream = (hipStream_t)hip_stream;
  S.stream_set = hip_stream != nullptr;
}

void *grid_mi355x_get_stream(void) { return (void *)stream(); }

void grid_mi355x_synchronize(void) { sync_stream(stream()); }

void grid_mi355x_set_timing(const bool enable) {
  if (!enable) resolve_timings();
  S.timing = enable;
}

void grid_mi355x_point_launches(const int kernel_id, long *fp64, long *fp32) {
  if (kernel_id < 0 || kernel_id >= kNumKernelIds) fatal("kernel id %d out of range", kernel_id);
  if (fp64) *fp64 = S.launches[kernel_id][0];
  if (fp32) *fp32 = S.launches[kernel_id][1];
}

void grid_mi355x_reset_timing(void) {
  resolve_timings();
  for (int i = 0; i < kNumKernelIds; i++) {
    S.kernel_ms[i] = 0.0;
    S.kernel_n[i] = 0;
    S.launches[i][0] = S.launches[i][1] = 0;
  }
}

double grid_mi355x_kernel_time_ms(const int kernel_id, long *launches) {
  if (kernel_id < 0 || kernel_id >= kNumKernelIds) fatal("unknown kernel id %d", kernel_id);
  resolve_timings();
  if (launches) *launches = S.kernel_n[kernel_id];
  return S.kernel_ms[kernel_id];
}

long long grid_mi355x_task_points(const grid_task_list *task_list) {
  TaskListImpl *tl = checked(task_list);
  if (tl->empty) return 0;
  if (tl->points >= 0) return tl->points;
  const hipStream_t s = stream();
  std::vector<double *> g(tl->nlevels, nullptr);
  const Level *levels = set_levels(tl, g.data(), s);
  PointParams P{};
  P.tasks = (const Task *)tl->d_tasks.p;
  P.ntasks = tl->ntasks;
  P.groups = (const Group *)tl->d_groups.p;
  P.levels = levels;
  P.task_out = (double *)tl->d_task_out.p;
  P.count_only = 1;
  const int ld[4] = {0, 0, 0, 0};
  P.plan = make_plan(tl, ld, false);
  if (tl->n_skip > 0)
    HIP_CHECK(launch_fill((double *)tl->d_task_out.p +
                              (long long)(tl->n_ortho + tl->n_general) * kTaskOutStride,
                          (long long)tl->n_skip * kTaskOutStride, 0.0, s));
  launch_ranges(tl, P, s, kCount);
  std::vector<double> out((size_t)tl->ntasks * kTaskOutStride);
  HIP_CHECK(hipMemcpyAsync(out.data(), tl->d_task_out.p, sizeof(double) * out.size(),
                           hipMemcpyDeviceToHost, s));
  sync_stream(s);
  // classed levels from the geometry classes, the rest from the kernel counts
  long long total = tl->class_points;
  tl->points_by_lp.assign(kLp0Slots, 0);
  for (const OrthoLevel &o : tl->olev)
    if (o.end > o.begin && o.classed)
      for (int t = o.begin; t < o.end; t++)
        tl->points_by_lp[tl->lp0s[t]] += tl->class_npts[t];
  auto add = [&](long long a, long long b) {
    for (long long t = a; t < b; t++) {
      const long long n = (long long)out[(size_t)t * kTaskOutStride];
      total += n;
      tl->points_by_lp[tl->lp0s[t]] += n;
    }
  };
  for (const OrthoLevel &o : tl->olev)
    if (o.end > o.begin && !o.classed) add(o.begin, o.end);
  add(tl->n_ortho, tl->n_ortho + tl->n_general);
  tl->points = total;
  return total;
}

void grid_mi355x_task_points_by_lp(const grid_task_list *task_list, long long *points,
                                   const int n) {
  TaskListImpl *tl = checked(task_list);
  for (int v = 0; v < n; v++) points[v] = 0;
  if (tl->empty) return;
  grid_mi355x_task_points(task_list);
  for (int v = 0; v < n && v < (int)tl->points_by_lp.size(); v++) points[v] = tl->points_by_lp[v];
}

void grid_mi355x_task_counts(const grid_task_list *task_list, long *ortho, long *general,
                             long *skipped) {
  const TaskListImpl *tl = checked(task_list);
  if (ortho) *ortho = tl->n_ortho;
  if (general) *general = tl->n_general;
  if (skipped) *skipped = tl->n_skip;
}


// ---- batched task lists (no reference counterpart) ----
void grid_mi355x_create_batch(const int nlists, const grid_task_list *const *lists,
                              grid_mi355x_batch **batch) {
  if (nlists <= 0) fatal("a batch needs at least one task list");
  BatchImpl *b = new BatchImpl();
  b->nlists = nlists;
  ListInputs all;
  b->level_off.assign(1, 0);
  b->atom_off.assign(1, 0);
  b->pab_off.assign(1, 0);
  int kind_off = 0, block_off = 0;
  bool any_ortho = false;
  for (int i = 0; i < nlists; i++) {
    const TaskListImpl *tl = checked(lists[i]);
    const ListInputs &in = tl->inputs;
    const int lo = b->level_off.back(), ao = b->atom_off.back();
    const long long po = b->pab_off.back();
    if (po + tl->block_end >= (1LL << 31)) fatal("batch: pab buffers above 2^31 doubles");
    for (int v : in.block_offsets) all.block_offsets.push_back((int)(po + v));
    all.atom_positions.insert(all.atom_positions.end(), in.atom_positions.begin(),
                              in.atom_positions.end());
    for (int k : in.atom_kinds) all.atom_kinds.push_back(k + kind_off);
    all.basis_sets.insert(all.basis_sets.end(), in.basis_sets.begin(), in.basis_sets.end());
    for (int t = 0; t < in.ntasks; t++) {
      all.level.push_back(in.level[t] + lo);
      all.iatom.push_back(in.iatom[t] + ao);
      all.jatom.push_back(in.jatom[t] + ao);
      all.block_num.push_back(in.block_num[t] + block_off);
    }
    for (const auto &pr : {std::make_pair(&all.iset, &in.iset), std::make_pair(&all.jset, &in.jset),
                     std::make_pair(&all.ipgf, &in.ipgf), std::make_pair(&all.jpgf, &in.jpgf),
                     std::make_pair(&all.border_mask, &in.border_mask),
                     std::make_pair(&all.npts_global, &in.npts_global),
                     std::make_pair(&all.npts_local, &in.npts_local),
                     std::make_pair(&all.shift_local, &in.shift_local),
                     std::make_pair(&all.border_width, &in.border_width)})
      pr.first->insert(pr.first->end(), pr.second->begin(), pr.second->end());
    for (const auto &pr : {std::make_pair(&all.radius, &in.radius), std::make_pair(&all.rab, &in.rab),
                     std::make_pair(&all.dh, &in.dh), std::make_pair(&all.dh_inv, &in.dh_inv)})
      pr.first->insert(pr.first->end(), pr.second->begin(), pr.second->end());
    all.level_ortho.insert(all.level_ortho.end(), in.level_ortho.begin(), in.level_ortho.end());
    for (char c : in.level_ortho) any_ortho = any_ortho || c;
    all.ntasks += in.ntasks;
    kind_off += in.nkinds;
    block_off += in.nblocks;
    b->level_off.push_back(lo + in.nlevels);
    b->atom_off.push_back(ao + in.natoms);
    b->pab_off.push_back(po + std::max<long long>(tl->block_end, 1));
  }
  const int nlevels = b->level_off.back(), natoms = b->atom_off.back();
  b->merged = new TaskListImpl();
  TaskListImpl *tl = b->merged;
  build_task_list(tl, any_ortho, all.ntasks, nlevels, natoms, kind_off, block_off,
                  all.block_offsets.data(), all.atom_positions.data(), all.atom_kinds.data(),
                  all.basis_sets.data(), all.level.data(), all.iatom.data(), all.jatom.data(),
                  all.iset.data(), all.jset.data(), all.ipgf.data(), all.jpgf.data(),
                  all.border_mask.data(), all.block_num.data(), all.radius.data(),
                  all.rab.data(), all.npts_global.data(), all.npts_local.data(),
                  all.shift_local.data(), all.border_width.data(), all.dh.data(),
                  all.dh_inv.data(), all.level_ortho.data());
  tl->merged = true;
  // buffers
  b->grid_off.assign(1, 0);
  for (int l = 0; l < nlevels; l++) b->grid_off.push_back(b->grid_off.back() + level_points(tl, l));
  b->pab.get(sizeof(double) * (size_t)b->pab_off.back());
  b->hab.get(sizeof(double) * (size_t)b->pab_off.back());
  b->rho.get(sizeof(double) * (size_t)std::max(b->grid_off.back(), 1LL));
  b->pot.get(sizeof(double) * (size_t)std::max(b->grid_off.back(), 1LL));
  b->forces.get(sizeof(double) * 3 * (size_t)std::max(natoms, 1));
  b->virial.get(sizeof(double) * 9 * (size_t)nlists);
  const hipStream_t s = stream();
  HIP_CHECK(launch_fill((double *)b->pab.p, b->pab_off.back(), 0.0, s));
  HIP_CHECK(launch_fill((double *)b->pot.p, std::max(b->grid_off.back(), 1LL), 0.0, s));
  for (int l = 0; l < nlevels; l++) {
    b->rho_lv.push_back((double *)b->rho.p + b->grid_off[l]);
    b->pot_lv.push_back((double *)b->pot.p + b->grid_off[l]);
  }
  // tasks of each list (CSR) for the per-list virial
  std::vector<int> list_of_level(std::max(nlevels, 1), 0), vptr(nlists + 1, 0), vent;
  for (int i = 0; i < nlists; i++)
    for (int l = b->level_off[i]; l < b->level_off[i + 1]; l++) list_of_level[l] = i;
  if (!tl->empty) {
    for (int t = 0; t < tl->ntasks; t++) vptr[list_of_level[tl->task_level[t]] + 1]++;
    for (int i = 0; i < nlists; i++) vptr[i + 1] += vptr[i];
    vent.resize(tl->ntasks);
    std::vector<int> fill(vptr.begin(), vptr.end() - 1);
    for (int t = 0; t < tl->ntasks; t++) vent[fill[list_of_level[tl->task_level[t]]]++] = t;
  }
  upload(b->vptr, vptr);
  upload(b->vent, vent);
  sync_stream(s);
  *batch = (grid_mi355x_batch *)b;
}

void grid_mi355x_free_batch(grid_mi355x_batch *batch) {
  if (!batch) return;
  BatchImpl *b = checked_batch(batch);
  sync_stream(stream());
  delete b;
}

void grid_mi355x_batch_buffers(const grid_mi355x_batch *batch, const int list, double **pab,
                               double **hab, long *length, double **forces, double **virial) {
  const BatchImpl *b = checked_batch(batch);
  if (list < 0 || list >= b->nlists) fatal("batch: list %d out of range", list);
  if (pab) *pab = (double *)b->pab.p + b->pab_off[list];
  if (hab) *hab = (double *)b->hab.p + b->pab_off[list];
  if (length) *length = (long)(b->pab_off[list + 1] - b->pab_off[list]);
  if (forces) *forces = (double *)b->forces.p + 3 * (size_t)b->atom_off[list];
  if (virial) *virial = (double *)b->virial.p + 9 * (size_t)list;
}

double *grid_mi355x_batch_grid(const grid_mi355x_batch *batch, const int list, const int level,
                               const int potential) {
  const BatchImpl *b = checked_batch(batch);
  if (list < 0 || list >= b->nlists) fatal("batch: list %d out of range", list);
  const int l = b->level_off[list] + level;
  if (level < 0 || l >= b->level_off[list + 1]) fatal("batch: level %d out of range", level);
  return potential ? b->pot_lv[l] : b->rho_lv[l];
}

void grid_mi355x_batch_run(grid_mi355x_batch *batch, const int what, const enum grid_func func,
                           const bool compute_tau, const bool forces) {
  BatchImpl *b = checked_batch(batch);
  if (what < 1 || what > 3) fatal("batch: what must be 1, 2 or 3");
  if (b->merged->empty) return;
  int ld[4];
  if (!func_ldiffs(func, ld)) fatal("unknown grid_func %d", (int)func);
  const hipStream_t s = stream();
  const std::array<int, 5> key = {what, (int)func, compute_tau ? 1 : 0, forces ? 1 : 0,
                                  precision()};
  auto it = b->graphs.find(key);
  if (it == b->graphs.end()) {
    // an eager run allocates what the calls allocate on first use and
    // builds the cached launch plans; the capture then holds kernels only.
    // If that run moved a buffer (e.g. the per-task cxyz buffer grows with
    // the l shift of the call), the graphs captured before address freed
    // memory: they are all dropped and captured again on their next use.
    const auto before = captured_buffers(b->merged);
    batch_enqueue(b, what, func, compute_tau, forces, nullptr, s);
    sync_stream(s);
    if (captured_buffers(b->merged) != before) b->drop_graphs();
    BatchImpl::Graph &g = b->graphs[key];
    static_levels(b, g.lev_col, b->rho_lv, nullptr);
    static_levels(b, g.lev_int, b->pot_lv,
                  precision() == GRID_MI355X_FP32_MIXED ? (const float *)b->merged->d_gridf.p
                                                        : nullptr);
    hipGraph_t graph = nullptr;
    S.capturing = true;
    HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    batch_enqueue(b, what, func, compute_tau, forces, &g, s);
    const hipError_t e = hipStreamEndCapture(s, &graph);
    S.capturing = false;
    HIP_CHECK(e);
    HIP_CHECK(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphDestroy(graph));
    it = b->graphs.find(key);
  }
  TimedLaunch tm(14, s);
  HIP_CHECK(hipGraphLaunch(it->second.exec, s));
}

void grid_mi355x_batch_task_points(const grid_mi355x_batch *batch, long long *points) {
  const BatchImpl *b = checked_batch(batch);
  *points = grid_mi355x_task_points((const grid_task_list *)b->merged);
}

void grid_mi355x_batch_set_work_queue(grid_mi355x_batch *batch, const bool enable) {
  BatchImpl *b = checked_batch(batch);
  if (b->merged->work_queue == enable) return;
  sync_stream(stream());
  b->drop_graphs();  // the captured launch shapes change
  b->merged->work_queue = enable;
}

int grid_mi355x_batch_work_queue(const grid_mi355x_batch *batch) {
  return checked_batch(batch)->merged->work_queue ? 1 : 0;
}

// Batch autotuner: times the run (what, func, compute_tau, forces) as
// graph replays with the point kernels launched one workgroup per unit and
// as persistent launches over device work queues, and keeps the faster.
int grid_mi355x_autotune_batch(grid_mi355x_batch *batch, const int what, const enum grid_func func,
                               const bool compute_tau, const bool forces) {
  BatchImpl *b = checked_batch(batch);
  if (b->merged->empty) return 0;
  const hipStream_t s = stream();
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  float best = -1.0f;
  int choice = 0;
  for (int q = 0; q < 2; q++) {
    grid_mi355x_batch_set_work_queue(batch, q == 1);
    grid_mi355x_batch_run(batch, what, func, compute_tau, forces);  // capture
    HIP_CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < 3; r++) grid_mi355x_batch_run(batch, what, func, compute_tau, forces);
    HIP_CHECK(hipEventRecord(e1, s));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (best < 0.0f || ms < best) {
      best = ms;
      choice = q;
    }
  }
  grid_mi355x_batch_set_work_queue(batch, choice == 1);
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  sync_stream(s);
  return choice;
}


// ---- launch-shape autotuner (no reference counterpart) ----
void grid_mi355x_autotune(const grid_task_list *task_list, const enum grid_func func,
                          const double *pab_blocks, double *const *grids) {
  TaskListImpl *tl = checked(task_list);
  if (tl->empty || tl->merged) return;
  const hipStream_t s = stream();
  hipEvent_t a, b;
  HIP_CHECK(hipEventCreate(&a));
  HIP_CHECK(hipEventCreate(&b));
  auto clear_plans = [&]() {
    sync_stream(s);
    for (auto &kv : tl->tile_plans) kv.second.d.release();
    tl->tile_plans.clear();
    for (auto &kv : tl->gtile_plans) kv.second.d.release();
    tl->gtile_plans.clear();
  };
  auto call_ms = [&]() {
    run_collocate(tl, func, pab_blocks, grids, s);  // plans and buffers
    HIP_CHECK(hipEventRecord(a, s));
    for (int r = 0; r < 3; r++) run_collocate(tl, func, pab_blocks, grids, s);
    HIP_CHECK(hipEventRecord(b, s));
    HIP_CHECK(hipEventSynchronize(b));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    return ms / 3.0f;
  };
  // coordinate descent, one tiled level at a time, over (rounds, max_waves)
  for (auto *lv : {&tl->olev, &tl->glev})
    for (OrthoLevel &o : *lv) {
      if (o.end <= o.begin || !o.tiled) continue;
      int best_r = o.rounds, best_w = o.max_waves;
      clear_plans();
      float best = call_ms();
      for (int r : {2, 4, 8})
        for (int w : {16, 12, 8, 6}) {
          if (w > (lv == &tl->glev ? kGenTileWaves : 12)) continue;
          if (r == best_r && w == best_w) continue;
          o.rounds = r;
          o.max_waves = w;
          clear_plans();
          const float ms = call_ms();
          if (ms < best) {
            best = ms;
            best_r = r;
            best_w = w;
          }
        }
      o.rounds = best_r;
      o.max_waves = best_w;
      clear_plans();
    }
  for (int sh = 0; sh <= 4; sh++) {  // as at create: every l shift planned
    if (!tl->tiles.empty()) tile_plan(tl, sh);
    if (!tl->gtiles.empty()) tile_plan(tl, sh, true);
  }
  HIP_CHECK(hipEventDestroy(a));
  HIP_CHECK(hipEventDestroy(b));
}

void grid_mi355x_tuned_shape(const grid_task_list *task_list, const int level, const int general,
                             int *rounds, int *max_waves) {
  const TaskListImpl *tl = checked(task_list);
  if (level < 0 || level >= tl->nlevels) fatal("level %d out of range", level);
  const OrthoLevel &o = general ? tl->glev[level] : tl->olev[level];
  *rounds = o.rounds;
  *max_waves = o.max_waves;
}

}  // extern "C"
