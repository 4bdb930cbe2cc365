// libmpm_hip.so: context management and the C-ABI declared in include/mpm_hip.h.
//
// One context = one simulation domain on one GPU with its own HIP stream. The host side of the
// reference solver (src/mpm_core/backends/warp/solver.py) maps onto the entry points as documented in
// the header; the kernels live in kernels.hpp, binning.hpp, colliders.hpp and substep.hpp.
//
// Substep on the default (binned) path:
//   [every sort_interval substeps, or after any state change]
//     k_cell_keys -> radix passes (k_radix_hist, k_scan_tiles, k_radix_scatter)
//     -> k_brick_bounds, k_brick_cost (cell offsets, cost class), k_brick_rank, k_brick_order (occupied
//     bricks, costliest first) -> k_reorder
//   [between sorts] drift routing (particles that changed cell -> incoming lists of their current cell), by
//                  the kernel after the previous G2P (k_fp64_fixup), else k_drift
//   k_p2g_split    (two waves per occupied brick, one lane per cell, LDS window)
//   k_grid_bricks  (dirty bricks: update, walls, colliders)
//   k_g2p          (gather, advection, collider projection, return map, stress, max speed)
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "mpm_hip.h"
#include "batch.hpp"
#include "emission.hpp"
#include "surface.hpp"

using namespace mpm;

// volume grid of mpm_is_occupied_noalloc: one float per node of a context's brick-rounded padded grid, held by
// each GPU the library runs on until the process ends (MPM_NOALLOC_MAX_NODES * 4 bytes = 32 MiB)
static __device__ __align__(256) float scratch[MPM_NOALLOC_MAX_NODES];

namespace {

thread_local std::string g_error;

struct Failure {
  int32_t code;
};

[[noreturn]] void fail(int32_t code, const std::string& msg) {
  g_error = msg;
  throw Failure{code};
}

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    fail(MPM_ERR_MEMORY, std::string(what) + ": out of device memory");
  fail(MPM_ERR_RUNTIME, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_CHECK(expr) hip_check((expr), #expr)

template <class T> T* dev_alloc(size_t count) {
  if (count == 0) count = 1;
  void* p = nullptr;
  hip_check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
  return static_cast<T*>(p);
}

void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

// device buffer that grows on demand (contents are not kept)
struct Scratch {
  void* ptr = nullptr;
  size_t bytes = 0;
  void* get(size_t need) {
    if (need > bytes) {
      dev_free(ptr);
      ptr = nullptr;
      bytes = 0;
      ptr = dev_alloc<char>(need);
      bytes = need;
    }
    return ptr;
  }
  void release() {
    dev_free(ptr);
    ptr = nullptr;
    bytes = 0;
  }
};

constexpr int kTimedClasses = 16;  // room for a context's (MPM_K_*) and a batch's (MPM_BK_*) kernel classes
static_assert(MPM_K_COUNT <= kTimedClasses && MPM_BK_COUNT <= kTimedClasses, "kernel classes");

struct Timing {
  uint32_t mask = 0;  // kernel classes bracketed with events (bit per class)
  struct Pending {
    int kernel;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  double total_ms[kTimedClasses] = {};
  int64_t launches[kTimedClasses] = {};
};

// all per-slot particle arrays of one buffer set, in one allocation: x 3, v 3, F 9, Jp 1 float planes,
// the 16-float records (C, stress, mass; 64 bytes per slot), then the mat and pid int planes
constexpr int kFloatPlanes = 16;
constexpr int kRecFloats = 16;
constexpr int kIntPlanes = 2;

struct ParticleSet {
  char* base = nullptr;
  int64_t cap = 0;
  float* x() const { return reinterpret_cast<float*>(base); }
  float* v() const { return x() + 3 * cap; }
  float* F() const { return x() + 6 * cap; }
  float* Jp() const { return x() + 15 * cap; }
  float4* rec() const { return reinterpret_cast<float4*>(x() + kFloatPlanes * cap); }
  int32_t* mat() const { return reinterpret_cast<int32_t*>(x() + (kFloatPlanes + kRecFloats) * cap); }
  int32_t* pid() const { return mat() + cap; }
  Particles view() const {
    return Particles{as_global(x()), as_global(v()), as_global(F()), as_global(Jp()), as_global(mat()),
                     as_global(pid()), as_global(rec()), cap};
  }
  static size_t bytes(int64_t cap) { return (size_t)cap * 4 * (kFloatPlanes + kRecFloats + kIntPlanes); }
};

struct ColliderHost {
  std::vector<TriGeom> tris;
  std::vector<TriNormals> normals;
  ColliderDev dev{};
  float built_pad = -1.0f;
  std::vector<uint32_t> cell_start;  // relative to this collider's candidate list
  std::vector<uint32_t> cell_tris;
  std::vector<float> cell_key;       // per candidate: distance from its cell's centre (lists sorted by it)
  std::vector<float> cell_far;       // per cell: signed lower bound of the distance to the surface
};

}  // namespace

struct mpm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;      // the stream every operation of the context runs on
  hipStream_t own_stream = nullptr;  // the context's own stream (stream is the batch's while in a batch)
  mpm_batch* batch = nullptr;
  bool busy = false;                 // in the middle of a batched frame: its state belongs to the batch
  uint64_t version = 0;              // bumped by every change of the particle state
  mpm_domain_desc dom{};
  std::vector<mpm_material_desc> mats;
  MatDev* d_mats = nullptr;

  // particles: set A is live, set B is the reorder target; emitter ids stay in insertion order
  int64_t n = 0, cap = 0;
  ParticleSet A, B;
  int32_t* emitter = nullptr;
  std::vector<char> present;

  // grid
  GridGeom geom{};
  StepConsts sc{};
  int n_nodes = 0, n_bricks = 0, n_cells = 0, key_bits = 1;
  float4 *accum = nullptr, *vel = nullptr;
  uint32_t* dirty = nullptr;
  uint8_t* live = nullptr;

  // binning
  uint32_t *keys0 = nullptr, *keys1 = nullptr, *keys2 = nullptr, *vals1 = nullptr, *vals2 = nullptr;
  uint32_t *offsets = nullptr, *active = nullptr;  // offsets: per cell; active: occupied bricks, costliest first
  uint32_t* n_active = nullptr;  // their count
  uint32_t* dirty_list = nullptr;  // batched grid update work list
  int32_t* slot = nullptr;         // per brick: its position in `active`, or -1
  float4* stage = nullptr;         // P2G windows of the occupied bricks (stage_cap x 512 nodes)
  int64_t stage_cap = 0;
  uint32_t* in_cnt = nullptr;      // drift routing: per cell of an occupied brick (stage_cap x 64)
  uint32_t* in_lst = nullptr;      // its incoming particles (stage_cap x 64 x kInPerCell)
  uint8_t* fwd = nullptr;          // per particle: routed to its current cell (cap)
  uint32_t* fix_list = nullptr;    // particles deferred to the FP64 pass (cap entries)
  uint32_t* n_fix = nullptr;       // their count
  bool n_fix_cleared = false;      // the grid update since the last G2P cleared n_fix
  Scratch hist, scan_sums;
  bool sorted = false;
  int64_t steps_since_sort = 0;
  uint32_t* last_perm = nullptr;  // the sorted values of the last sort (points into vals1/vals2)

  // options
  int p2g_mode = 0;  // MPM_P2G_BINNED, MPM_P2G_ATOMIC or MPM_P2G_FIXED
  unsigned long long* fx = nullptr;  // fixed-point accumulator (MPM_P2G_FIXED only)
  double fx_scale = 0.0;
  int64_t sort_interval = 24;
  int64_t p2g_blocks = 0;
  int p2g_waves = 2;             // waves per occupied brick in the binned P2G (1 or 2)
  int p2g_flush = 0;             // MPM_P2G_FLUSH_STAGED or MPM_P2G_FLUSH_ATOMIC
  float4* flush_stage() const { return p2g_flush == MPM_P2G_FLUSH_STAGED ? stage : nullptr; }

  // colliders
  std::vector<ColliderHost> colliders;
  ColliderDev* d_cols = nullptr;
  std::vector<ColliderDev> col_records;  // the records; d_cols is their copy for the batch
  bool cols_stale = false;               // d_cols older than col_records
  TriGeom* d_tris = nullptr;
  TriNormals* d_normals = nullptr;
  uint32_t *d_cell_start = nullptr, *d_cell_tris = nullptr;
  float *d_cell_key = nullptr, *d_cell_far = nullptr;

  // The reduced maximum speed, in pinned host memory the reduction writes to: h_vmax[0] the speed, h_vmax[1]
  // (as uint32) the sequence number of the request it answers, written after it behind a system fence.
  float* h_vmax = nullptr;
  float* h_vmax_dev = nullptr;    // its device address
  uint32_t vmax_seq = 0;          // the last request's sequence number
  float* d_vmax = nullptr;        // device copy of the published speed
  bool vmax_pending = false;      // a reduction into h_vmax of the current state is on the stream
  float* d_block_max = nullptr;   // per-workgroup maxima of the last G2P (or speed pass)
  int64_t block_max_cap = 0;
  int64_t block_max_count = 0;
  bool vmax_valid = false;
  Scratch scratch;
  Scratch emit_scratch;  // lattice emission (mpm_lattice_inside): mesh, column tables, crossings, points
  Timing timing;
  double frame_dt = 1.0;
  bool capturing = false;  // inside a graph capture: no events, no allocations, no host waits
  bool drift_ahead = false;  // the last G2P's follow-up kernel routed the drift for the next substep

  Particles parts() const { return A.view(); }
  FixedAccum fixed() const {
    return FixedAccum{p2g_mode == MPM_P2G_FIXED ? fx : nullptr, fx_scale, fx_scale > 0.0 ? 1.0 / fx_scale : 0.0};
  }
  ColliderSet collider_set() const {
    ColliderSet cs{};
    cs.count = (int)colliders.size();
    cs.cols = as_global(d_cols);
    cs.tris = as_global(d_tris);
    cs.normals = as_global(d_normals);
    cs.cell_start = as_global(d_cell_start);
    cs.cell_tris = as_global(d_cell_tris);
    cs.cell_key = as_global(d_cell_key);
    cs.cell_far = as_global(d_cell_far);
    cs.thickness = (float)(2.0 * dom.dx_f64);
    cs.band = (float)dom.dx_f64;
    return cs;
  }
};

// A set of contexts advanced together (batch.hpp). While a context belongs to a batch, its operations
// run on the batch's stream, so host-side calls on it (poses, state reads) are ordered with the batch's
// frames without further synchronisation.
struct mpm_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<mpm_ctx*> scenes;
  // per scene, host side. `step` is an exact copy of the device step table between runs (every run
  // uploads it, replays substeps, and reads it back); a frame starts by rewriting a scene's entries.
  std::vector<SceneDev> rec;
  std::vector<SceneStep> step;
  std::vector<float> speed;        // speed maximum the next frame's CFL step uses
  std::vector<uint64_t> version;   // context version `speed` belongs to
  std::vector<char> speed_ok;
  std::vector<double> frame_dt;    // of the frame in flight
  std::vector<double> dt;          // its substep length
  std::vector<char> empty_frame;   // a frame of zero substeps was started (reported by the next run)

  // device tables and pinned staging for their transfers
  int64_t table_cap = 0;
  SceneDev* d_scenes = nullptr;
  SceneStep* d_steps = nullptr;
  ColliderFrame* d_frames = nullptr;
  uint32_t* d_pre = nullptr;  // work-unit prefixes (kPrefixKinds x (scenes + 1))
  float* d_speed = nullptr;
  uint32_t* d_speed_slots = nullptr;  // kSpeedSlots per scene: a substep's G2P speed maxima
  uint32_t* d_tickets = nullptr;      // kTicketKinds counters, zero between launches
  SceneDev* h_scenes = nullptr;
  SceneStep* h_steps = nullptr;
  float* h_speed = nullptr;

  // sizes the captured graphs depend on: the radix histogram (256 x all radix tiles at full capacity)
  // and the largest radix pass count
  int64_t hist_tiles = -1;
  int max_passes = 1;
  Scratch hist, scan_sums;

  hipGraphExec_t g_sort = nullptr, g_plain = nullptr;

  // options: persistent grid sizes (0: automatic), graphs, sort cadence
  int64_t p2g_blocks = 0;
  int64_t g2p_blocks = 0;
  int64_t grid_blocks = 0;
  int64_t g2p_tile = 256;
  int64_t p2g_waves = 1;  // waves per occupied brick in P2G (the mix's many bricks fill the machine)
  int64_t use_graph = 1;
  bool capturing = false;  // batch_enqueue is recording into a graph
  Timing timing;           // per MPM_BK_* class; while on, substeps are launched directly
  bool fp64_pass = false;  // some scene defers particles to the FP64 pass (part of the captured sequence)
  int64_t sort_interval = 16;  // 0: each context's own (the mix's fast scenes drift more per substep)
  int64_t substeps_launched = 0;
};

namespace {

struct DeviceGuard {
  int prev = 0;
  explicit DeviceGuard(int dev) {
    (void)hipGetDevice(&prev);
    hip_check(hipSetDevice(dev), "hipSetDevice");
  }
  ~DeviceGuard() { (void)hipSetDevice(prev); }
};

// ---- the fixed volume buffer of mpm_is_occupied_noalloc ------------------------------------------
// Host code never passes `scratch` itself to a kernel: in host code the name stands for a host-side shadow.
// Each device's own address comes from hipGetSymbolAddress, taken once per device when the first context is
// created there, and every use goes through that address.
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
float* g_noalloc_volume[kMaxDevices] = {};

void noalloc_volume_setup(int device) {  // with `device` current
  if (device < 0 || device >= kMaxDevices) fail(MPM_ERR_VALUE, "device ordinal beyond the library's table");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_volume[device]) return;
  void* p = nullptr;
  hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (!p) fail(MPM_ERR_RUNTIME, "hipGetSymbolAddress(scratch) returned null");
  g_noalloc_volume[device] = static_cast<float*>(p);
}

float* noalloc_volume(int device) {
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  float* p = (device >= 0 && device < kMaxDevices) ? g_noalloc_volume[device] : nullptr;
  if (!p) fail(MPM_ERR_RUNTIME, "the fixed volume buffer was not set up on this device");
  return p;
}

// ---- launch timing -------------------------------------------------------------------------
hipEvent_t pool_event(Timing& t) {
  if (!t.pool.empty()) {
    hipEvent_t e = t.pool.back();
    t.pool.pop_back();
    return e;
  }
  hipEvent_t e;
  HIP_CHECK(hipEventCreate(&e));
  return e;
}

// events around the launches of one kernel class, when the class is timed and nothing is being captured
struct Timed {
  Timing* t = nullptr;
  hipStream_t s = nullptr;
  int kernel;
  hipEvent_t a = nullptr;
  Timed(Timing& timing, hipStream_t stream, int k, bool capturing) : kernel(k) {
    if (((timing.mask >> k) & 1u) && !capturing) {
      t = &timing;
      s = stream;
      a = pool_event(timing);
      HIP_CHECK(hipEventRecord(a, s));
    }
  }
  Timed(mpm_ctx* c, int k) : Timed(c->timing, c->stream, k, c->capturing) {}
  ~Timed() noexcept(false) {
    if (a) {
      hipEvent_t b = pool_event(*t);
      HIP_CHECK(hipEventRecord(b, s));
      t->pending.push_back({kernel, a, b});
    }
  }
};

// One kernel launch of class k, timed (when the class is on) by events the dispatch records itself
// (hipExtLaunchKernelGGL): no marker packets on the stream between the kernel and its neighbours, which
// would add a few microseconds to every bracketed launch.
template <typename... Args, typename F = void (*)(Args...)>
void launch_timed(Timing& timing, hipStream_t stream, bool capturing, int k, F kernel, dim3 grid, dim3 block,
                  Args... args) {
  if (((timing.mask >> k) & 1u) && !capturing) {
    hipEvent_t a = pool_event(timing), b = pool_event(timing);
    hipExtLaunchKernelGGL(kernel, grid, block, 0, stream, a, b, 0, args...);
    timing.pending.push_back({k, a, b});
  } else {
    hipLaunchKernelGGL(kernel, grid, block, 0, stream, args...);
  }
}

void collect_timing(Timing& t, hipStream_t stream) {
  if (t.pending.empty()) return;
  HIP_CHECK(hipStreamSynchronize(stream));
  for (auto& p : t.pending) {
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
    t.total_ms[p.kernel] += ms;
    t.launches[p.kernel] += 1;
    t.pool.push_back(p.a);
    t.pool.push_back(p.b);
  }
  t.pending.clear();
}

void collect_timing(mpm_ctx* c) { collect_timing(c->timing, c->stream); }

void release_timing(Timing& t) {
  for (auto& p : t.pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : t.pool) (void)hipEventDestroy(e);
  t.pending.clear();
  t.pool.clear();
}

void check_launch(const char* what) { hip_check(hipGetLastError(), what); }

// per-workgroup speed maxima, one entry per 256 particle slots
float* block_max(mpm_ctx* c) {
  const int64_t need = std::max<int64_t>(1, (c->cap + kBlock - 1) / kBlock);
  if (need > c->block_max_cap) {
    dev_free(c->d_block_max);
    c->d_block_max = nullptr;
    c->d_block_max = dev_alloc<float>(need);
    c->block_max_cap = need;
  }
  return c->d_block_max;
}

// ---- particle storage ------------------------------------------------------------------------
void free_particles(mpm_ctx* c) {
  dev_free(c->A.base);
  dev_free(c->B.base);
  for (void* p : {(void*)c->emitter, (void*)c->keys0, (void*)c->keys1, (void*)c->keys2, (void*)c->vals1,
                  (void*)c->vals2, (void*)c->fix_list, (void*)c->fwd})
    dev_free(p);
  c->fix_list = nullptr;
  c->fwd = nullptr;
  c->A = ParticleSet{};
  c->B = ParticleSet{};
  c->emitter = nullptr;
  c->keys0 = c->keys1 = c->keys2 = c->vals1 = c->vals2 = nullptr;
  c->last_perm = nullptr;
  c->cap = 0;
}

// grow capacity to hold `need` particles, keeping the first c->n
void reserve(mpm_ctx* c, int64_t need) {
  if (need <= c->cap) return;
  int64_t cap = std::max<int64_t>(need, c->cap + c->cap / 2);
  cap = (cap + 255) & ~int64_t(255);
  ParticleSet A, B;
  A.cap = B.cap = cap;
  A.base = dev_alloc<char>(ParticleSet::bytes(cap));
  B.base = dev_alloc<char>(ParticleSet::bytes(cap));
  int32_t* em = dev_alloc<int32_t>(cap);
  uint32_t* k0 = dev_alloc<uint32_t>(cap);
  uint32_t* k1 = dev_alloc<uint32_t>(cap);
  uint32_t* k2 = dev_alloc<uint32_t>(cap);
  uint32_t* v1 = dev_alloc<uint32_t>(cap);
  uint32_t* v2 = dev_alloc<uint32_t>(cap);
  uint32_t* fl = dev_alloc<uint32_t>(cap);
  uint8_t* fw = dev_alloc<uint8_t>(cap);
  if (c->n) {
    for (int k = 0; k < kFloatPlanes; ++k)
      HIP_CHECK(hipMemcpyAsync(A.x() + k * cap, c->A.x() + k * c->cap, c->n * 4, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(A.rec(), c->A.rec(), c->n * kRecFloats * 4, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(A.mat(), c->A.mat(), c->n * 4, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(A.pid(), c->A.pid(), c->n * 4, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(em, c->emitter, c->n * 4, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  const int64_t n = c->n;
  free_particles(c);
  c->A = A;
  c->B = B;
  c->emitter = em;
  c->keys0 = k0;
  c->keys1 = k1;
  c->keys2 = k2;
  c->vals1 = v1;
  c->vals2 = v2;
  c->fix_list = fl;
  c->fwd = fw;
  c->cap = cap;
  // one staged window and one set of incoming lists per occupied brick: at most one brick per particle
  const int64_t stage_need = std::min<int64_t>(c->n_bricks, cap);
  if (stage_need > c->stage_cap) {
    for (void* p : {(void*)c->stage, (void*)c->in_cnt, (void*)c->in_lst}) dev_free(p);
    c->stage = nullptr;
    c->in_cnt = c->in_lst = nullptr;
    c->stage = dev_alloc<float4>((size_t)stage_need * kStageNodes);
    c->in_cnt = dev_alloc<uint32_t>((size_t)stage_need * 64);
    c->in_lst = dev_alloc<uint32_t>((size_t)stage_need * 64 * kInPerCell);
    HIP_CHECK(hipMemsetAsync(c->in_cnt, 0, (size_t)stage_need * 64 * sizeof(uint32_t), c->stream));
    c->stage_cap = stage_need;
  }
  c->n = n;
}

// upload `n` host rows of `comps` floats into planes starting at slot p0
void upload_planes(mpm_ctx* c, float* planes, const float* host, int comps, int64_t n, int64_t p0) {
  if (n == 0) return;
  float* stage = static_cast<float*>(c->scratch.get(n * comps * sizeof(float)));
  HIP_CHECK(hipMemcpyAsync(stage, host, n * comps * sizeof(float), hipMemcpyHostToDevice, c->stream));
  k_unpack<<<grid_for(n), kBlock, 0, c->stream>>>(stage, comps, (int)n, c->cap, p0, planes);
  check_launch("k_unpack");
}

// download planes into host rows in insertion order (through pid)
void download_planes(mpm_ctx* c, const float* planes, float* host, int comps) {
  if (c->n == 0) return;
  float* stage = static_cast<float*>(c->scratch.get(c->n * comps * sizeof(float)));
  k_pack<<<grid_for(c->n), kBlock, 0, c->stream>>>(planes, c->cap, c->A.pid(), comps, (int)c->n, stage);
  check_launch("k_pack");
  HIP_CHECK(hipMemcpyAsync(host, stage, c->n * comps * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

// C of every particle into host rows in insertion order
void download_C(mpm_ctx* c, float* host) {
  if (c->n == 0) return;
  float* stage = static_cast<float*>(c->scratch.get(c->n * 9 * sizeof(float)));
  k_pack_C<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->parts(), (int)c->n, stage);
  check_launch("k_pack_C");
  HIP_CHECK(hipMemcpyAsync(host, stage, c->n * 9 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

void check_materials(mpm_ctx* c, const int32_t* mat, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (mat[i] < 0 || mat[i] >= (int32_t)c->mats.size()) fail(MPM_ERR_VALUE, "material id out of range");
}

// empty P2G accumulators (float and fixed point) and no pending bricks, as between two substeps
void clear_grid(mpm_ctx* c) {
  HIP_CHECK(hipMemsetAsync(c->accum, 0, (size_t)c->n_nodes * sizeof(float4), c->stream));
  HIP_CHECK(hipMemsetAsync(c->dirty, 0, (size_t)c->n_bricks * 4, c->stream));
  if (c->fx) HIP_CHECK(hipMemsetAsync(c->fx, 0, (size_t)c->n_nodes * 4 * sizeof(unsigned long long), c->stream));
}

void state_changed(mpm_ctx* c) {
  c->version += 1;
  c->sorted = false;
  c->vmax_valid = false;
  c->vmax_pending = false;
}

// ---- device-wide exclusive scan --------------------------------------------------------------
// The scan's buffer for `tiles` tiles: word 0 is k_scan_tiles' ticket, the tile totals follow. The ticket
// is cleared on `stream` when the buffer is (re)allocated, and the kernel leaves it at zero. A buffer that
// a graph capture would have to grow is an error (the batch sizes its buffer before it captures).
uint32_t* scan_buffer(Scratch& s, int64_t tiles, hipStream_t stream) {
  const size_t need = (size_t)(tiles + 1) * sizeof(uint32_t);
  if (need > s.bytes) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    HIP_CHECK(hipStreamIsCapturing(stream, &st));
    if (st != hipStreamCaptureStatusNone) fail(MPM_ERR_RUNTIME, "scan buffer grown during a graph capture");
    s.get(need);
    HIP_CHECK(hipMemsetAsync(s.ptr, 0, sizeof(uint32_t), stream));
  }
  return static_cast<uint32_t*>(s.ptr);
}

// the two-part scan of n words (k_scan_tiles): tile-local scans into out; returns the scanned tile totals
const uint32_t* scan_u32_tiles(Scratch& scan_sums, hipStream_t stream, const uint32_t* in, uint32_t* out,
                               int64_t n) {
  const int64_t tiles = std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile);
  if (tiles > kScanTile) fail(MPM_ERR_VALUE, "scan too large");
  uint32_t* buf = scan_buffer(scan_sums, tiles, stream);
  k_scan_tiles<<<(unsigned)tiles, kScanThreads, 0, stream>>>(in, (int)n, out, buf + 1, buf);
  check_launch("k_scan_tiles");
  return buf + 1;
}

void scan_u32(Scratch& scan_sums, hipStream_t stream, const uint32_t* in, uint32_t* out, int64_t n) {
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  const uint32_t* sums = scan_u32_tiles(scan_sums, stream, in, out, n);
  if (tiles > 1) {
    k_scan_add<<<(unsigned)tiles, kScanThreads, 0, stream>>>(out, (int)n, sums);
    check_launch("k_scan_add");
  }
}

// ---- binning ---------------------------------------------------------------------------------
// the occupied-brick list (costliest bricks first) and the brick -> slot map, from the cell offsets
// cell offsets from the sorted keys (brick boundaries, then each brick's cells within its own range), the
// bricks' cost classes and the occupied-brick list
// the radix sort's digit histograms, then the brick classes' per-workgroup counts (one buffer, sized once
// for both so that no reallocation waits on the stream between them)
uint32_t* sort_hist(mpm_ctx* c, int64_t tiles) {
  const size_t words = std::max<size_t>((size_t)256 * (size_t)tiles, (size_t)kCostClasses * kBrickGroups);
  return static_cast<uint32_t*>(c->hist.get(words * sizeof(uint32_t)));
}

void bin_offsets_and_bricks(mpm_ctx* c, const uint32_t* sorted_keys, int64_t n) {
  const int64_t groups = (c->n_bricks + kBrickGroup - 1) / kBrickGroup;
  if (groups > kBrickGroups) fail(MPM_ERR_VALUE, "grid too large for the brick ordering (at most 2^18 bricks)");
  uint32_t* hist = sort_hist(c, (n + kSortTile - 1) / kSortTile);
  k_brick_bounds<<<grid_for(c->n_bricks + 1), kBlock, 0, c->stream>>>(sorted_keys, (int)n, (int)c->n_bricks,
                                                                      c->offsets);
  check_launch("k_brick_bounds");
  if (groups == 0) {  // no bricks: nothing occupied
    k_zero_u32<<<1, 64, 0, c->stream>>>(c->n_active, 1);
    check_launch("k_zero_u32");
    return;
  }
  k_brick_cost<<<(unsigned)((c->n_bricks + 3) / 4), 256, 0, c->stream>>>(sorted_keys, (int)c->n_bricks, c->offsets,
                                                                         c->slot);
  check_launch("k_brick_cost");
  k_brick_rank<<<(unsigned)groups, kBrickGroup, 0, c->stream>>>((int)c->n_bricks, c->slot, hist);
  check_launch("k_brick_rank");
  k_brick_order<<<(unsigned)groups, kBrickGroup, 0, c->stream>>>((int)c->n_bricks, hist, c->active, c->n_active,
                                                                 c->slot);
  check_launch("k_brick_order");
}

void sort_particles(mpm_ctx* c) {
  const int64_t n = c->n;
  Timed t(c, MPM_K_BIN);
  if (n == 0) {
    bin_offsets_and_bricks(c, c->keys1, 0);
    c->sorted = true;
    c->steps_since_sort = 0;
    return;
  }
  k_cell_keys<<<grid_for(n), kBlock, 0, c->stream>>>(c->parts(), c->geom, (int)n, c->keys0, c->vals2);
  check_launch("k_cell_keys");
  const int passes = (c->key_bits + 7) / 8;
  const int tiles = (int)((n + kSortTile - 1) / kSortTile);
  uint32_t* hist = sort_hist(c, tiles);
  const uint32_t* kin = c->keys0;
  const uint32_t* vin = c->vals2;
  for (int pass = 0; pass < passes; ++pass) {
    uint32_t* kout = (pass % 2 == 0) ? c->keys1 : c->keys2;
    uint32_t* vout = (pass % 2 == 0) ? c->vals1 : c->vals2;
    k_radix_hist<<<tiles, kSortThreads, 0, c->stream>>>(kin, (int)n, 8 * pass, tiles, hist);
    check_launch("k_radix_hist");
    const uint32_t* sums = scan_u32_tiles(c->scan_sums, c->stream, hist, hist, (int64_t)256 * tiles);
    k_radix_scatter<<<tiles, kSortThreads, 0, c->stream>>>(kin, vin, (int)n, 8 * pass, tiles, hist, sums, kout,
                                                           vout);
    check_launch("k_radix_scatter");
    kin = kout;
    vin = vout;
  }
  bin_offsets_and_bricks(c, kin, n);
  k_reorder<<<grid_for(n), kBlock, 0, c->stream>>>(c->A.view(), c->B.view(), vin, (int)n);
  check_launch("k_reorder");
  std::swap(c->A, c->B);
  c->last_perm = const_cast<uint32_t*>(vin);
  c->sorted = true;
  c->steps_since_sort = 0;
}

// the cell keys of the last sort, in slot order: the output of its last radix pass
const uint32_t* ctx_sorted_keys(const mpm_ctx* c) {
  return ((c->key_bits + 7) / 8 - 1) % 2 == 0 ? c->keys1 : c->keys2;
}

// ---- substep phases --------------------------------------------------------------------------
// The binning the next binned P2G reads: a sort every sort_interval substeps, else (BINNED) the drift
// lists. It reads positions only, so it can be enqueued before the substep's dt is known.
// Drift lists routed ahead (by the kernel after the last G2P) for a substep that then does not read them
// (the state changed, or the batched engine takes the scene over): their counts would add to the next
// routing, which P2G clears only after reading it, so they are cleared here.
void discard_drift_ahead(mpm_ctx* c) {
  if (!c->drift_ahead) return;
  if (c->in_cnt) HIP_CHECK(hipMemsetAsync(c->in_cnt, 0, (size_t)c->stage_cap * 64 * sizeof(uint32_t), c->stream));
  c->drift_ahead = false;
}

DriftLists bin_for_p2g(mpm_ctx* c) {
  const bool sorting = !c->sorted || c->steps_since_sort >= c->sort_interval;
  if (sorting) {
    discard_drift_ahead(c);
    sort_particles(c);
  }
  // between sorts, drifted particles go to their current cell's lane (not in fixed point: the lists'
  // order is that of atomics, and the fixed-point sums must not depend on it); usually the kernel after
  // the last G2P has routed them already
  DriftLists dl{};
  if (!sorting && c->n > 0 && c->p2g_mode == MPM_P2G_BINNED) {
    if (!c->drift_ahead) {
      Timed t(c, MPM_K_OTHER);
      k_drift<<<grid_for((c->n + 1) / 2), kBlock, 0, c->stream>>>(c->parts(), c->geom, (int)c->n,
                                                                  ctx_sorted_keys(c), c->slot, c->in_cnt,
                                                                  c->in_lst, c->fwd);
      check_launch("k_drift");
    }
    dl = DriftLists{c->in_cnt, c->in_lst, c->fwd};
  }
  c->drift_ahead = false;
  return dl;
}

// P2G; `pre` is the binning bin_for_p2g already enqueued for it (null: bin now)
// cfl.vmax: derive dt on the device (the advance's first substep, launched before the host knows dt)
void run_p2g(mpm_ctx* c, float dt, const DriftLists* pre = nullptr, CflStep cfl = CflStep{}) {
  if (c->p2g_mode == MPM_P2G_ATOMIC) {
    if (c->n == 0) return;
    Timed t(c, MPM_K_P2G);
    k_p2g_atomic<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->parts(), c->d_mats, c->geom, c->sc, (int)c->n, dt,
                                                           c->accum);
    check_launch("k_p2g_atomic");
    return;
  }
  const DriftLists dl = pre ? *pre : bin_for_p2g(c);
  if (c->n == 0) return;
  // one wave per workgroup, so a brick starts as soon as any wave slot frees; grid-stride over the
  // device-side brick list (costliest first)
  int64_t blocks = c->p2g_blocks > 0 ? c->p2g_blocks : 16384;
  blocks = std::max<int64_t>(1, std::min<int64_t>(blocks, std::min<int64_t>(c->n_bricks, c->n)));
  const auto go = [&](auto kernel, unsigned threads) {
    launch_timed(c->timing, c->stream, c->capturing, MPM_K_P2G, kernel, dim3((unsigned)blocks), dim3(threads),
                 c->parts(), c->geom, c->sc.p_vol, dt, (const uint32_t*)c->offsets, (const uint32_t*)c->active,
                 (const uint32_t*)c->n_active, c->accum, c->dirty, c->fixed(), c->flush_stage(), dl, cfl);
  };
  if (c->p2g_waves == 4) go(k_p2g_split<4>, 256);
  else if (c->p2g_waves == 2) go(k_p2g_split<2>, 128);
  else go(k_p2g_cells<1>, 64);
  check_launch("k_p2g");
}

// the colliders' frames at frame fraction s (the same function the batched plan runs on the device)
ColliderFrames frames_at(const mpm_ctx* c, float s, float inv_frame_dt) {
  ColliderFrames fr{};
  for (size_t i = 0; i < c->col_records.size() && i < (size_t)kMaxColliders; ++i) {
    collider_frame(c->col_records[i], s, inv_frame_dt, fr.f[i]);
    fr.cols[i] = c->col_records[i];
  }
  return fr;
}

void run_grid(mpm_ctx* c, float dt, float s, float inv_frame_dt) {
  Timed t(c, MPM_K_GRID);
  if (c->p2g_mode == MPM_P2G_ATOMIC) {
    k_grid_dense<<<grid_for(c->n_nodes), kBlock, 0, c->stream>>>(c->geom, c->sc, dt, c->n_nodes, c->accum, c->vel,
                                                                  c->live, c->collider_set(),
                                                                  frames_at(c, s, inv_frame_dt), c->n_fix);
    check_launch("k_grid_dense");
  } else {
    const int64_t waves = (c->n_bricks + kGridBricksPerWave - 1) / kGridBricksPerWave;
    const unsigned blocks = (unsigned)((waves + 3) / 4);
    k_grid_bricks<<<blocks, 256, 0, c->stream>>>(c->geom, c->sc, dt, c->n_bricks, c->dirty, c->live, c->accum,
                                                 c->vel, c->collider_set(), frames_at(c, s, inv_frame_dt),
                                                 c->fixed(), c->flush_stage(), c->slot, c->n_fix);
    check_launch("k_grid_bricks");
  }
  c->n_fix_cleared = true;
}

// device address of the published sequence number (next to the speed in pinned host memory)
uint32_t* speed_seq_dev(mpm_ctx* c) { return reinterpret_cast<uint32_t*>(c->h_vmax_dev + 1); }

// reduce_speed: the caller wants this G2P's speed maximum next (the last substep of an advance); with the
// FP64 pass on, that pass reduces it into h_vmax on the way and the request is marked as enqueued
void run_g2p(mpm_ctx* c, float dt, float s, float inv_frame_dt, bool reduce_speed = false) {
  if (c->n == 0) return;
  const bool fix = c->sc.precision != kPrecisionFp32;
  if (fix && !c->n_fix_cleared) {  // a G2P without the grid update before it (a debug hook)
    k_zero_u32<<<1, 64, 0, c->stream>>>(c->n_fix, 1);
    check_launch("k_zero_u32");
  }
  c->n_fix_cleared = false;
  launch_timed(c->timing, c->stream, c->capturing, MPM_K_G2P, k_g2p, dim3((unsigned)grid_for(c->n)), dim3(kBlock),
               c->parts(), (const MatDev*)c->d_mats, c->geom, c->sc, (int)c->n, dt, (const float4*)c->vel,
               block_max(c), c->collider_set(), frames_at(c, s, inv_frame_dt), FixList{c->fix_list, c->n_fix});
  check_launch("k_g2p");
  const bool fused = fix && reduce_speed;
  // Whether the next substep will route drift (no sort before it, state unchanged): then the kernel after
  // this G2P routes it, while the positions G2P wrote are in the cache, and the substep needs no k_drift.
  const bool drift_next = c->p2g_mode == MPM_P2G_BINNED && c->sorted && c->steps_since_sort + 1 < c->sort_interval;
  if (fix || drift_next) {
    // a persistent grid: the deferred particles are usually few
    Timed t(c, MPM_K_OTHER);
    if (fused) ++c->vmax_seq;
    const DriftAhead da = drift_next ? DriftAhead{c->geom, (int)c->n, ctx_sorted_keys(c), c->slot, c->in_cnt,
                                                  c->in_lst, c->fwd}
                                     : DriftAhead{};
    const int64_t blocks = drift_next ? std::min<int64_t>(grid_for((c->n + 1) / 2), 2048) : std::min<int64_t>(grid_for(c->n), 256);
    k_fp64_fixup<<<(unsigned)blocks, kBlock, 0, c->stream>>>(
        c->parts(), c->d_mats, c->fix_list, fix ? c->n_fix : nullptr, block_max(c), (int)grid_for(c->n),
        fused ? c->h_vmax_dev : nullptr, speed_seq_dev(c), c->vmax_seq, c->d_vmax, da);
    check_launch("k_fp64_fixup");
    c->drift_ahead = drift_next;
  }
  c->block_max_count = grid_for(c->n);
  c->vmax_valid = true;
  c->vmax_pending = fused;  // (a reduction enqueued before this G2P is of the old state)
}

// p2g_done: this substep's P2G is already on the stream (the advance's first, launched early)
void substep(mpm_ctx* c, double dt, double s, double frame_dt, const DriftLists* pre = nullptr,
             bool reduce_speed = false, bool p2g_done = false) {
  const float dtf = (float)dt, sf = (float)s, inv_fdt = (float)(1.0 / frame_dt);
  if (!p2g_done) run_p2g(c, dtf, pre);
  run_grid(c, dtf, sf, inv_fdt);
  run_g2p(c, dtf, sf, inv_fdt, reduce_speed);
  c->steps_since_sort += 1;
  c->version += 1;
}

// max |v| over the particles (kernels.py:266-269): the last G2P left per-workgroup maxima; after any
// state change a speed pass recomputes them
// the speed reduction without a host wait: writes max |v| to `out` (device memory) on the context's stream
// seq_out: publish the maximum to the host with that sequence number (see reduce_block_maxima)
void reduce_speed_async(mpm_ctx* c, float* out, uint32_t* seq_out = nullptr, uint32_t seq = 0) {
  if (!c->vmax_valid) {
    k_speed_partials<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->parts(), (int)c->n, block_max(c));
    check_launch("k_speed_partials");
    c->block_max_count = grid_for(c->n);
    c->vmax_valid = true;
  }
  k_reduce_max<<<1, kBlock, 0, c->stream>>>(c->d_block_max, (int)c->block_max_count, out, seq_out, seq,
                                            seq_out ? c->d_vmax : nullptr);
  check_launch("k_reduce_max");
}

// Enqueue the reduction of the current state's maximum speed into pinned host memory, right behind the
// work that produced the state, so that the next max_speed only waits for it (no launch, no copy, on the
// path between two advances).
void request_speed(mpm_ctx* c) {
  if (c->n == 0 || c->vmax_pending) return;
  {
    Timed t(c, MPM_K_OTHER);
    ++c->vmax_seq;
    reduce_speed_async(c, c->h_vmax_dev, speed_seq_dev(c), c->vmax_seq);
  }
  c->vmax_pending = true;
}

// Waits for the requested reduction only (work enqueued behind it keeps running) by polling the sequence
// number the reduction publishes with it: the step between two advances is on the critical path, and
// neither a blocking wait (its wake-up latency) nor an event behind the kernel (a barrier packet between
// it and the next kernel) belongs there. Should the stream finish without the number appearing, the wait
// fails rather than spinning on.
float max_speed(mpm_ctx* c) {
  request_speed(c);
  if (c->n > 0) {
    // (the stream is queried only after a second of spinning: a query puts a marker on the stream, which
    // delays the next kernel by several microseconds)
    const volatile uint32_t* seq = reinterpret_cast<const volatile uint32_t*>(c->h_vmax + 1);
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t spins = 1; *seq != c->vmax_seq; ++spins) {
      if (spins % 65536 == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1)) {
        const hipError_t e = hipStreamQuery(c->stream);
        if (e == hipSuccess && *seq != c->vmax_seq) fail(MPM_ERR_RUNTIME, "speed reduction did not publish");
        if (e != hipSuccess && e != hipErrorNotReady) HIP_CHECK(e);
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
  }
  return static_cast<volatile float*>(c->h_vmax)[0];
}

void load_state(mpm_ctx* c, int64_t n, const float* x, const float* v, const float* Cm, const float* F,
                const float* Jp, const int32_t* mat, const int32_t* emitter) {
  if (n < 0) fail(MPM_ERR_VALUE, "negative particle count");
  if (n > 0 && (!x || !v || !Cm || !F || !Jp || !mat || !emitter)) fail(MPM_ERR_VALUE, "null state array");
  check_materials(c, mat, n);
  if (n > (int64_t)INT32_MAX) fail(MPM_ERR_VALUE, "too many particles");
  clear_grid(c);  // a grid left by a debug phase hook belongs to the old state
  c->n = 0;
  reserve(c, n);
  upload_planes(c, c->A.x(), x, 3, n, 0);
  upload_planes(c, c->A.v(), v, 3, n, 0);
  if (n) {
    float* stage = static_cast<float*>(c->scratch.get(n * 9 * sizeof(float)));
    HIP_CHECK(hipMemcpyAsync(stage, Cm, n * 9 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    k_unpack_C<<<grid_for(n), kBlock, 0, c->stream>>>(stage, (int)n, 0, c->parts());
    check_launch("k_unpack_C");
  }
  upload_planes(c, c->A.F(), F, 9, n, 0);
  upload_planes(c, c->A.Jp(), Jp, 1, n, 0);
  if (n) {
    HIP_CHECK(hipMemcpyAsync(c->A.mat(), mat, n * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(c->emitter, emitter, n * 4, hipMemcpyHostToDevice, c->stream));
    k_iota<<<grid_for(n), kBlock, 0, c->stream>>>(c->A.pid(), (int)n);
    check_launch("k_iota");
    k_stress<<<grid_for(n), kBlock, 0, c->stream>>>(c->parts(), c->d_mats, (int)n, c->sc.precision);
    check_launch("k_stress");
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  c->n = n;
  std::fill(c->present.begin(), c->present.end(), 0);
  for (int64_t i = 0; i < n; ++i) c->present[mat[i]] = 1;
  state_changed(c);
}

void refresh_present(mpm_ctx* c) {
  std::fill(c->present.begin(), c->present.end(), 0);
  if (c->n == 0) return;
  std::vector<int32_t> m(c->n);
  HIP_CHECK(hipMemcpyAsync(m.data(), c->A.mat(), c->n * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  for (int32_t id : m) c->present[id] = 1;
}

int64_t remove_escaped_host(mpm_ctx* c);

// drop particles more than one cell past an open side (solver.py:222-240); keeps insertion order of the
// survivors. On the device: flags by insertion index, their exclusive scan (the survivors' new indices),
// one compaction into the other particle buffer; only the survivor count comes back to the host.
int64_t remove_escaped(mpm_ctx* c) {
  const mpm_domain_desc& d = c->dom;
  bool any_open = false;
  for (int a = 0; a < 3; ++a) any_open = any_open || !d.lo_on[a] || !d.hi_on[a];
  if (!any_open || c->n == 0) return 0;
  const int64_t n = c->n;
  if (n > (int64_t)kScanTile * kScanTile) return remove_escaped_host(c);
  char* base = static_cast<char*>(c->scratch.get(3 * (size_t)n * sizeof(int32_t)));
  int32_t* keep = reinterpret_cast<int32_t*>(base);
  uint32_t* new_pid = reinterpret_cast<uint32_t*>(base + (size_t)n * 4);
  int32_t* em = reinterpret_cast<int32_t*>(base + 2 * (size_t)n * 4);
  k_escape_flags<<<grid_for(n), kBlock, 0, c->stream>>>(c->parts(), (int)n,
                                                        make_float3(d.remove_lo[0], d.remove_lo[1], d.remove_lo[2]),
                                                        make_float3(d.remove_hi[0], d.remove_hi[1], d.remove_hi[2]),
                                                        keep);
  check_launch("k_escape_flags");
  scan_u32(c->scan_sums, c->stream, reinterpret_cast<const uint32_t*>(keep), new_pid, n);
  uint32_t tail[2];
  HIP_CHECK(hipMemcpyAsync(&tail[0], new_pid + n - 1, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipMemcpyAsync(&tail[1], keep + n - 1, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  const int64_t kept = (int64_t)tail[0] + (int64_t)tail[1];
  const int64_t removed = n - kept;
  if (removed == 0) return 0;
  k_compact<<<grid_for(n), kBlock, 0, c->stream>>>(c->A.view(), c->B.view(), keep, new_pid, (int)n, c->emitter, em);
  check_launch("k_compact");
  if (kept) HIP_CHECK(hipMemcpyAsync(c->emitter, em, kept * 4, hipMemcpyDeviceToDevice, c->stream));
  std::swap(c->A, c->B);
  c->n = kept;
  state_changed(c);
  return removed;
}

// the same through a host copy of the state (for counts beyond the device scan's reach)
int64_t remove_escaped_host(mpm_ctx* c) {
  const mpm_domain_desc& d = c->dom;
  const int64_t n = c->n;
  int32_t* keep = static_cast<int32_t*>(c->scratch.get(n * sizeof(int32_t)));
  k_escape_flags<<<grid_for(n), kBlock, 0, c->stream>>>(c->parts(), (int)n,
                                                        make_float3(d.remove_lo[0], d.remove_lo[1], d.remove_lo[2]),
                                                        make_float3(d.remove_hi[0], d.remove_hi[1], d.remove_hi[2]),
                                                        keep);
  check_launch("k_escape_flags");
  std::vector<int32_t> flags(n);
  HIP_CHECK(hipMemcpyAsync(flags.data(), keep, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  int64_t kept = 0;
  for (int32_t f : flags) kept += f;
  const int64_t removed = n - kept;
  if (removed == 0) return 0;
  // compaction through the host copy: removals are rare (only past open sides) and the reference
  // also round-trips the state here (solver.py:237-239)
  std::vector<float> x(3 * n), v(3 * n), C(9 * n), F(9 * n), Jp(n);
  std::vector<int32_t> mat(n), em(n);
  download_planes(c, c->A.x(), x.data(), 3);
  download_planes(c, c->A.v(), v.data(), 3);
  download_C(c, C.data());
  download_planes(c, c->A.F(), F.data(), 9);
  download_planes(c, c->A.Jp(), Jp.data(), 1);
  download_planes(c, reinterpret_cast<float*>(c->A.mat()), reinterpret_cast<float*>(mat.data()), 1);
  HIP_CHECK(hipMemcpyAsync(em.data(), c->emitter, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  int64_t w = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (!flags[i]) continue;
    std::memmove(&x[3 * w], &x[3 * i], 12);
    std::memmove(&v[3 * w], &v[3 * i], 12);
    std::memmove(&C[9 * w], &C[9 * i], 36);
    std::memmove(&F[9 * w], &F[9 * i], 36);
    Jp[w] = Jp[i];
    mat[w] = mat[i];
    em[w] = em[i];
    ++w;
  }
  load_state(c, kept, x.data(), v.data(), C.data(), F.data(), Jp.data(), mat.data(), em.data());
  return removed;
}

// ---- colliders -------------------------------------------------------------------------------
// per-collider records (pose, margin, offsets into the shared tables); small, rewritten on every pose
// update with an async copy on the context's stream
// the collider records (pose endpoints, margins, offsets into the shared tables) on the host; the
// single-scene kernels take them by value, the batch reads a device copy refreshed when stale
void upload_collider_records(mpm_ctx* c) {
  if (c->colliders.empty()) return;
  std::vector<ColliderDev> cols;
  int tri = 0, cell = 0;
  for (auto& h : c->colliders) {
    ColliderDev d = h.dev;
    d.tri_offset = tri;
    d.cell_offset = cell;
    tri += (int)h.tris.size();
    cell += h.cell_start.empty() ? 0 : (int)h.cell_start.size() - 1;
    cols.push_back(d);
  }
  c->col_records = cols;
  c->cols_stale = true;
}

void ensure_collider_records(mpm_ctx* c) {
  if (!c->cols_stale || c->colliders.empty()) return;
  HIP_CHECK(hipMemcpyAsync(c->d_cols, c->col_records.data(), c->col_records.size() * sizeof(ColliderDev),
                           hipMemcpyHostToDevice, c->stream));
  c->cols_stale = false;
}

// meshes, pseudonormals and candidate lists of every collider (after set_colliders or a grid rebuild)
void upload_colliders(mpm_ctx* c) {
  HIP_CHECK(hipStreamSynchronize(c->stream));
  for (void* p : {(void*)c->d_cols, (void*)c->d_tris, (void*)c->d_normals, (void*)c->d_cell_start,
                  (void*)c->d_cell_tris, (void*)c->d_cell_key, (void*)c->d_cell_far})
    dev_free(p);
  c->d_cols = nullptr;
  c->d_tris = nullptr;
  c->d_normals = nullptr;
  c->d_cell_start = nullptr;
  c->d_cell_tris = nullptr;
  c->d_cell_key = nullptr;
  c->d_cell_far = nullptr;
  if (c->colliders.empty()) return;
  std::vector<TriGeom> tris;
  std::vector<TriNormals> normals;
  std::vector<uint32_t> starts, cand;
  std::vector<float> keys, far;
  for (auto& h : c->colliders) {
    tris.insert(tris.end(), h.tris.begin(), h.tris.end());
    normals.insert(normals.end(), h.normals.begin(), h.normals.end());
    const uint32_t base = (uint32_t)cand.size();
    const size_t cells = h.cell_start.empty() ? 0 : h.cell_start.size() - 1;
    for (size_t i = 0; i < cells; ++i) starts.push_back(base + h.cell_start[i]);
    cand.insert(cand.end(), h.cell_tris.begin(), h.cell_tris.end());
    keys.insert(keys.end(), h.cell_key.begin(), h.cell_key.end());
    far.insert(far.end(), h.cell_far.begin(), h.cell_far.end());
  }
  starts.push_back((uint32_t)cand.size());
  far.push_back(0.0f);  // never empty
  c->d_cols = dev_alloc<ColliderDev>(c->colliders.size());
  c->d_tris = dev_alloc<TriGeom>(tris.size());
  c->d_normals = dev_alloc<TriNormals>(normals.size());
  c->d_cell_start = dev_alloc<uint32_t>(starts.size());
  c->d_cell_tris = dev_alloc<uint32_t>(cand.size());
  c->d_cell_key = dev_alloc<float>(keys.size());
  c->d_cell_far = dev_alloc<float>(far.size());
  HIP_CHECK(hipMemcpy(c->d_cell_far, far.data(), far.size() * sizeof(float), hipMemcpyHostToDevice));
  if (!tris.empty()) {
    HIP_CHECK(hipMemcpy(c->d_tris, tris.data(), tris.size() * sizeof(TriGeom), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->d_normals, normals.data(), normals.size() * sizeof(TriNormals), hipMemcpyHostToDevice));
  }
  HIP_CHECK(hipMemcpy(c->d_cell_start, starts.data(), starts.size() * 4, hipMemcpyHostToDevice));
  if (!cand.empty()) {
    HIP_CHECK(hipMemcpy(c->d_cell_tris, cand.data(), cand.size() * 4, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(c->d_cell_key, keys.data(), keys.size() * 4, hipMemcpyHostToDevice));
  }
  upload_collider_records(c);
}

// candidate-triangle voxel grid of one collider over its AABB grown by `pad` (local units)
void build_candidates(mpm_ctx* c, ColliderHost& h, float pad) {
  ColliderDev& d = h.dev;
  h.cell_start.clear();
  h.cell_tris.clear();
  h.cell_key.clear();
  h.cell_far.clear();
  if (d.n_tris == 0) {
    d.g_res[0] = d.g_res[1] = d.g_res[2] = 0;
    h.built_pad = pad;
    return;
  }
  float ext[3], emax = 0.0f;
  for (int a = 0; a < 3; ++a) {
    ext[a] = (d.aabb_hi[a] - d.aabb_lo[a]) + 2.0f * pad;
    emax = std::max(emax, ext[a]);
  }
  // Voxel edge 0.8 sqrt(mean triangle area), within [emax/63.5, emax/16] (at most 64 cells per axis
  // after rounding): about six triangles lie within
  // the candidate radius (twice the half diagonal) of a surface point, so a query near the surface ends
  // after one or two rounds whatever the mesh's resolution.
  double area = 0.0;
  for (const TriGeom& t : h.tris) {
    const double e1[3] = {t.v[3] - t.v[0], t.v[4] - t.v[1], t.v[5] - t.v[2]};
    const double e2[3] = {t.v[6] - t.v[0], t.v[7] - t.v[1], t.v[8] - t.v[2]};
    const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                 cz = e1[0] * e2[1] - e1[1] * e2[0];
    area += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
  }
  const float target = (float)(0.8 * std::sqrt(area / (double)h.tris.size()));
  const float hvox = std::max(std::min(std::max(target, emax / 63.5f), emax / 16.0f), 1e-12f);
  int64_t cells = 1;
  for (int a = 0; a < 3; ++a) {
    d.g_res[a] = std::min(64, std::max(1, (int)std::ceil(ext[a] / hvox)));
    d.g_lo[a] = d.aabb_lo[a] - pad;
    cells *= d.g_res[a];
  }
  d.g_inv_h = 1.0f / hvox;
  // cells with a candidate list: centre within pad (the contact reach of any pose the grid is built for)
  // plus two half diagonals of the surface
  const float band = pad + 2.0f * 0.8660254f * hvox;
  TriGeom* dtris = dev_alloc<TriGeom>(h.tris.size());
  TriNormals* dnorm = dev_alloc<TriNormals>(h.normals.size());
  float* dfar = dev_alloc<float>(cells);
  uint32_t* counts = dev_alloc<uint32_t>(cells + 1);
  uint32_t* starts = dev_alloc<uint32_t>(cells + 1);
  ColliderDev local = d;
  local.tri_offset = 0;
  HIP_CHECK(hipMemcpyAsync(dtris, h.tris.data(), h.tris.size() * sizeof(TriGeom), hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(hipMemcpyAsync(dnorm, h.normals.data(), h.normals.size() * sizeof(TriNormals), hipMemcpyHostToDevice,
                           c->stream));
  k_voxel_candidates<<<grid_for(cells), kBlock, 0, c->stream>>>(local, dtris, dnorm, (int)cells, 0, counts, dfar,
                                                                nullptr, nullptr, nullptr, band);
  check_launch("k_voxel_candidates");
  std::vector<uint32_t> cnt(cells);
  h.cell_far.resize(cells);
  HIP_CHECK(hipMemcpyAsync(cnt.data(), counts, cells * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipMemcpyAsync(h.cell_far.data(), dfar, cells * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  h.cell_start.assign(cells + 1, 0);
  for (int64_t i = 0; i < cells; ++i) h.cell_start[i + 1] = h.cell_start[i] + cnt[i];
  const uint32_t total = h.cell_start[cells];
  uint32_t* out = dev_alloc<uint32_t>(std::max<uint32_t>(total, 1));
  float* out_key = dev_alloc<float>(std::max<uint32_t>(total, 1));
  HIP_CHECK(hipMemcpyAsync(starts, h.cell_start.data(), (cells + 1) * 4, hipMemcpyHostToDevice, c->stream));
  k_voxel_candidates<<<grid_for(cells), kBlock, 0, c->stream>>>(local, dtris, dnorm, (int)cells, 1, nullptr, nullptr,
                                                                starts, out, out_key, band);
  check_launch("k_voxel_candidates");
  k_sort_candidates<<<(unsigned)std::min<int64_t>(cells, 4096), 256, 0, c->stream>>>((int)cells, starts, out, out_key);
  check_launch("k_sort_candidates");
  h.cell_tris.resize(total);
  h.cell_key.resize(total);
  if (total) {
    HIP_CHECK(hipMemcpyAsync(h.cell_tris.data(), out, total * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipMemcpyAsync(h.cell_key.data(), out_key, total * 4, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_CHECK(hipStreamSynchronize(c->stream));
  dev_free(dtris);
  dev_free(dnorm);
  dev_free(dfar);
  dev_free(counts);
  dev_free(starts);
  dev_free(out);
  dev_free(out_key);
  h.built_pad = pad;
}

// triangle corners and pseudonormals of one mesh (double precision, stored as float)
void mesh_tables(const float* verts, int nv, const int32_t* tris, int nt, ColliderHost& h) {
  std::vector<double> fn(3 * (size_t)nt, 0.0), vn(3 * (size_t)nv, 0.0);
  std::vector<std::pair<uint64_t, int>> edges;
  edges.reserve(3 * (size_t)nt);
  for (int t = 0; t < nt; ++t) {
    const int32_t* tri = tris + 3 * t;
    double P[3][3];
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 3; ++i) P[k][i] = verts[3 * tri[k] + i];
    const double e1[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
    const double e2[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (len > 0)
      for (double& x : n) x /= len;
    for (int i = 0; i < 3; ++i) fn[3 * t + i] = n[i];
    for (int k = 0; k < 3; ++k) {
      const double* p0 = P[k];
      const double* p1 = P[(k + 1) % 3];
      const double* p2 = P[(k + 2) % 3];
      const double u[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
      const double w[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
      const double lu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      const double lw = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
      double ang = 0.0;
      if (lu > 0 && lw > 0)
        ang = std::acos(std::min(1.0, std::max(-1.0, (u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) / (lu * lw))));
      for (int i = 0; i < 3; ++i) vn[3 * tri[k] + i] += ang * n[i];
      const int a = tri[k], b = tri[(k + 1) % 3];
      edges.push_back({(uint64_t(std::min(a, b)) << 32) | uint64_t(std::max(a, b)), 3 * t + k});
    }
  }
  std::vector<double> en(9 * (size_t)nt, 0.0);
  std::sort(edges.begin(), edges.end());
  for (size_t i = 0; i < edges.size();) {
    size_t j = i;
    double s[3] = {0, 0, 0};
    while (j < edges.size() && edges[j].first == edges[i].first) {
      const int t = edges[j].second / 3;
      for (int d = 0; d < 3; ++d) s[d] += fn[3 * t + d];
      ++j;
    }
    for (size_t k = i; k < j; ++k)
      for (int d = 0; d < 3; ++d) en[3 * edges[k].second + d] = s[d];
    i = j;
  }
  h.tris.resize(nt);
  h.normals.resize(nt);
  for (int t = 0; t < nt; ++t) {
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 3; ++i) h.tris[t].v[3 * k + i] = verts[3 * tris[3 * t + k] + i];
    float* o = h.normals[t].n;
    for (int i = 0; i < 3; ++i) o[i] = (float)fn[3 * t + i];
    for (int e = 0; e < 3; ++e)
      for (int i = 0; i < 3; ++i) o[3 + 3 * e + i] = (float)en[3 * (3 * t + e) + i];
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 3; ++i) o[12 + 3 * k + i] = (float)vn[3 * tris[3 * t + k] + i];
  }
}

// ---- batches -----------------------------------------------------------------------------------
void require_idle(const mpm_ctx* c) {
  if (c->busy) fail(MPM_ERR_VALUE, "the scene is in the middle of a batched frame: run its batch first");
}

void batch_drop_graphs(mpm_batch* b) {
  if (b->g_sort) (void)hipGraphExecDestroy(b->g_sort);
  if (b->g_plain) (void)hipGraphExecDestroy(b->g_plain);
  b->g_sort = b->g_plain = nullptr;
}

void batch_free_tables(mpm_batch* b) {
  for (void* p : {(void*)b->d_scenes, (void*)b->d_steps, (void*)b->d_frames, (void*)b->d_pre, (void*)b->d_speed,
                  (void*)b->d_speed_slots, (void*)b->d_tickets})
    dev_free(p);
  for (void* p : {(void*)b->h_scenes, (void*)b->h_steps, (void*)b->h_speed})
    if (p) (void)hipHostFree(p);
  b->d_scenes = nullptr;
  b->d_steps = nullptr;
  b->d_frames = nullptr;
  b->d_pre = nullptr;
  b->d_speed = nullptr;
  b->d_speed_slots = nullptr;
  b->d_tickets = nullptr;
  b->h_scenes = nullptr;
  b->h_steps = nullptr;
  b->h_speed = nullptr;
  b->table_cap = 0;
}

template <class T> T* host_alloc(size_t count) {
  void* p = nullptr;
  if (hipHostMalloc(&p, count * sizeof(T), hipHostMallocDefault) != hipSuccess || !p) {
    (void)hipGetLastError();
    fail(MPM_ERR_MEMORY, "pinned host allocation failed");
  }
  return static_cast<T*>(p);
}

// device tables for n scenes (contents are uploaded from the host copies at every run)
void batch_reserve_tables(mpm_batch* b, int64_t n) {
  if (n <= b->table_cap) return;
  HIP_CHECK(hipStreamSynchronize(b->stream));
  batch_drop_graphs(b);
  batch_free_tables(b);
  const int64_t cap = std::max<int64_t>(n, 16);
  b->d_scenes = dev_alloc<SceneDev>(cap);
  b->d_steps = dev_alloc<SceneStep>(cap);
  b->d_frames = dev_alloc<ColliderFrame>(cap * kMaxColliders);
  b->d_pre = dev_alloc<uint32_t>((size_t)kPrefixKinds * (cap + 1));
  b->d_speed = dev_alloc<float>(cap);
  b->d_speed_slots = dev_alloc<uint32_t>((size_t)cap * kSpeedSlots);
  b->d_tickets = dev_alloc<uint32_t>(kTicketKinds);
  // the ticket counters start at zero (ordered before every later launch on the batch stream; never
  // inside a capture: tables are sized before one) and every kernel that takes tickets leaves them so
  HIP_CHECK(hipMemsetAsync(b->d_tickets, 0, kTicketKinds * sizeof(uint32_t), b->stream));
  b->h_scenes = host_alloc<SceneDev>(cap);
  b->h_steps = host_alloc<SceneStep>(cap);
  b->h_speed = host_alloc<float>(cap);
  b->table_cap = cap;
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

int radix_passes(const mpm_ctx* c) { return (c->key_bits + 7) / 8; }

// Buffers the captured graphs reference, sized for every scene at full capacity: the radix histogram and
// its scan. The graphs are recaptured when these grow (or the scene set or an option changes).
void batch_layout(mpm_batch* b) {
  int64_t tiles = 0;
  int passes = 1;
  for (mpm_ctx* c : b->scenes) {
    tiles += ceil_div(c->cap, kSortTile);
    passes = std::max(passes, radix_passes(c));
  }
  tiles = std::max<int64_t>(tiles, 1);
  if (256 * tiles > (int64_t)kScanTile * kScanTile || 256 * tiles > (int64_t)UINT32_MAX)
    fail(MPM_ERR_VALUE, "batch too large for the radix histogram scan");
  if (tiles <= b->hist_tiles && passes == b->max_passes) return;
  HIP_CHECK(hipStreamSynchronize(b->stream));
  batch_drop_graphs(b);
  // sized before any capture: nothing allocates inside one
  b->hist.get(256 * tiles * sizeof(uint32_t));
  scan_buffer(b->scan_sums, ceil_div(256 * tiles, kScanTile), b->stream);
  b->hist_tiles = tiles;
  b->max_passes = passes;
}

unsigned batch_blocks(int64_t option, int64_t automatic) { return (unsigned)(option > 0 ? option : automatic); }

// one substep of every running scene, as kernel launches on the batch stream (captured or direct)
void batch_enqueue(mpm_batch* b, bool sorting) {
  const int n = (int)b->scenes.size();
  hipStream_t s = b->stream;
  auto timed = [&](int k) { return Timed(b->timing, s, k, b->capturing); };
  // persistent grids: a few workgroups per CU, striding over the running scenes' work
  {
    const Timed t = timed(MPM_BK_PLAN);
    k_batch_plan<<<1, kPlanThreads, 0, s>>>(b->d_scenes, b->d_steps, b->d_frames, n, sorting ? 1 : 0,
                                            (int)b->g2p_tile, b->d_pre, b->d_speed_slots);
    check_launch("k_batch_plan");
  }
  if (sorting) {
    const Timed t = timed(MPM_BK_SORT);
    k_batch_cell_keys<<<2048, kTile, 0, s>>>(b->d_scenes, b->d_steps, b->d_pre, n);
    check_launch("k_batch_cell_keys");
    uint32_t* hist = static_cast<uint32_t*>(b->hist.ptr);
    for (int pass = 0; pass < b->max_passes; ++pass) {
      k_batch_radix_hist<<<1024, kSortThreads, 0, s>>>(b->d_scenes, b->d_pre, n, pass, hist);
      check_launch("k_batch_radix_hist");
      const uint32_t* sums = scan_u32_tiles(b->scan_sums, s, hist, hist, 256 * b->hist_tiles);
      k_batch_radix_scatter<<<1024, kSortThreads, 0, s>>>(b->d_scenes, b->d_pre, n, pass, hist, sums);
      check_launch("k_batch_radix_scatter");
    }
    k_batch_cell_bounds<<<4096, 256, 0, s>>>(b->d_scenes, b->d_pre, n);
    check_launch("k_batch_cell_bounds");
    k_batch_reorder<<<2048, kTile, 0, s>>>(b->d_scenes, b->d_steps, b->d_pre, n);
    check_launch("k_batch_reorder");
  }
  {
    const Timed t = timed(MPM_BK_DRIFT);
    k_batch_drift<<<2048, 256, 0, s>>>(b->d_scenes, b->d_steps, b->d_pre, n, (int)b->g2p_tile, b->d_tickets);
    check_launch("k_batch_drift");
  }
  {
  const Timed t = timed(MPM_BK_P2G);
  if (b->p2g_waves == 1) {
    k_batch_p2g<1><<<batch_blocks(b->p2g_blocks, 8192), 64, 0, s>>>(b->d_scenes, b->d_steps, b->d_pre, n);
  } else if (b->p2g_waves == 2) {
    k_batch_p2g_split<2><<<batch_blocks(b->p2g_blocks, 8192), 128, 0, s>>>(b->d_scenes, b->d_steps, b->d_pre, n);
  } else {
    k_batch_p2g_split<4><<<batch_blocks(b->p2g_blocks, 8192), 256, 0, s>>>(b->d_scenes, b->d_steps, b->d_pre, n);
  }
  check_launch("k_batch_p2g");
  }
  {
    const Timed t = timed(MPM_BK_GRID_SCAN);
    k_batch_grid_scan<<<1024, 256, 0, s>>>(b->d_scenes, b->d_steps, b->d_pre, n, b->d_tickets);
    check_launch("k_batch_grid_scan");
  }
  {
    const Timed t = timed(MPM_BK_GRID);
    k_batch_grid<<<batch_blocks(b->grid_blocks, 2048), 256, 0, s>>>(b->d_scenes, b->d_steps, b->d_pre, n,
                                                                    b->d_frames);
    check_launch("k_batch_grid");
  }
  {
    const Timed t = timed(MPM_BK_G2P);
    k_batch_g2p<<<batch_blocks(b->g2p_blocks, 2048), (unsigned)b->g2p_tile, 0, s>>>(
        b->d_scenes, b->d_steps, b->d_pre, n, b->d_frames, b->d_speed_slots, b->fp64_pass ? 1 : 0, b->d_tickets);
    check_launch("k_batch_g2p");
  }
  if (b->fp64_pass) {
    const Timed t = timed(MPM_BK_FIXUP);
    k_batch_fixup<<<1024, 256, 0, s>>>(b->d_scenes, b->d_steps, b->d_pre, n);
    check_launch("k_batch_fixup");
  }
  {
    const Timed t = timed(MPM_BK_ADVANCE);
    k_batch_advance<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(b->d_steps, n, b->d_speed_slots);
    check_launch("k_batch_advance");
  }
}

hipGraphExec_t batch_capture(mpm_batch* b, bool sorting) {
  hipGraph_t graph = nullptr;
  HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
  b->capturing = true;
  try {
    batch_enqueue(b, sorting);
    b->capturing = false;
  } catch (...) {
    b->capturing = false;
    (void)hipStreamEndCapture(b->stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  HIP_CHECK(hipStreamEndCapture(b->stream, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIP_CHECK(e);
  return exec;
}

SceneDev scene_record(const mpm_batch* b, mpm_ctx* c) {
  ensure_collider_records(c);
  SceneDev d{};
  d.set[0] = c->A.view();
  d.set[1] = c->B.view();
  d.mats = as_global(c->d_mats);
  d.g = c->geom;
  d.sc = c->sc;
  d.cs = c->collider_set();
  d.accum = as_global(c->accum);
  d.vel = as_global(c->vel);
  d.dirty = as_global(c->dirty);
  d.live = as_global(c->live);
  d.keys0 = as_global(c->keys0);
  d.keys1 = as_global(c->keys1);
  d.keys2 = as_global(c->keys2);
  d.vals1 = as_global(c->vals1);
  d.vals2 = as_global(c->vals2);
  d.offsets = as_global(c->offsets);
  d.active = as_global(c->active);
  d.n_active = as_global(c->n_active);
  d.dirty_list = as_global(c->dirty_list);
  d.slot = as_global(c->slot);
  d.stage = as_global(c->stage);
  d.in_cnt = as_global(c->in_cnt);
  d.in_lst = as_global(c->in_lst);
  d.fwd = as_global(c->fwd);
  d.fix_list = as_global(c->fix_list);
  d.n = (int32_t)c->n;
  d.n_bricks = c->n_bricks;
  d.n_cells = c->n_cells;
  d.passes = radix_passes(c);
  d.sort_interval = (int32_t)std::min<int64_t>(b->sort_interval > 0 ? b->sort_interval : c->sort_interval, 1 << 30);
  return d;
}

bool in_frame(const SceneStep& s) { return s.k < s.steps; }

// Start a frame of every scene i with frame_dt[i] > 0 (it must not be in a frame): its CFL substep count
// from its speed maximum (the arithmetic of mpm_advance, solver.py:213-220) and fresh table entries. The
// substeps run in the following runs.
void batch_begin(mpm_batch* b, const double* frame_dt) {
  const int64_t n = (int64_t)b->scenes.size();
  std::vector<int64_t> starting;
  for (int64_t i = 0; i < n; ++i) {
    if (!(frame_dt[i] > 0.0)) continue;
    mpm_ctx* c = b->scenes[i];
    if (in_frame(b->step[i]) || b->empty_frame[i]) fail(MPM_ERR_VALUE, "a scene of the batch is still in a frame");
    if (c->p2g_mode != MPM_P2G_BINNED) fail(MPM_ERR_VALUE, "batched scenes use the binned P2G (option P2G_MODE 0)");
    starting.push_back(i);
  }
  if (starting.empty()) return;
  // speed maxima of scenes whose state changed since the batch last saw them
  bool any = false;
  for (int64_t i : starting) {
    mpm_ctx* c = b->scenes[i];
    if (c->n == 0 || (b->speed_ok[i] && b->version[i] == c->version)) continue;
    batch_reserve_tables(b, n);
    reduce_speed_async(c, b->d_speed + i);
    any = true;
  }
  if (any) {
    HIP_CHECK(hipMemcpyAsync(b->h_speed, b->d_speed, n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    HIP_CHECK(hipStreamSynchronize(b->stream));
    for (int64_t i : starting) {
      mpm_ctx* c = b->scenes[i];
      if (c->n == 0 || (b->speed_ok[i] && b->version[i] == c->version)) continue;
      b->speed[i] = b->h_speed[i];
      b->speed_ok[i] = 1;
      b->version[i] = c->version;
    }
  }
  for (int64_t i : starting) {
    mpm_ctx* c = b->scenes[i];
    const double fdt = frame_dt[i];
    b->frame_dt[i] = fdt;
    if (c->n == 0) {  // nothing to advance (mpm_advance returns at once)
      b->dt[i] = fdt;
      b->empty_frame[i] = 1;
      continue;
    }
    double vmax = (double)b->speed[i];
    if (!std::isfinite(vmax)) vmax = 0.0;  // already blown up: keep stepping so the NaN reaches x
    double wave = 0.0;
    for (size_t m = 0; m < c->mats.size(); ++m)
      if (c->present[m]) wave = std::max(wave, c->mats[m].wave_speed);
    const double cfl_dt = std::min(0.4 * c->dom.dx_f64 / (wave + vmax), fdt);
    const int64_t steps = (int64_t)std::ceil(fdt / cfl_dt - 1e-9);
    if (steps > (int64_t)1 << 30) fail(MPM_ERR_VALUE, "too many substeps in one frame");
    b->dt[i] = fdt / (double)steps;
    c->frame_dt = fdt;
    b->rec[i] = scene_record(b, c);
    SceneStep& st = b->step[i];
    st = SceneStep{};
    st.dt = (float)b->dt[i];
    st.inv_frame_dt = (float)(1.0 / fdt);
    st.steps = (int32_t)steps;
    st.since_sort = (int32_t)std::min<int64_t>(c->steps_since_sort, 1 << 30);
    st.sorted = c->sorted ? 1 : 0;
    discard_drift_ahead(c);  // the batch routes drift itself (its stream orders after the context's)
    c->busy = true;
  }
}

// hand a scene back to its context after its frame's last substep (or when it leaves the batch)
void batch_release_scene(mpm_batch* b, int64_t i) {
  mpm_ctx* c = b->scenes[i];
  SceneStep& st = b->step[i];
  if (!c->busy) return;
  if (st.cur) std::swap(c->A, c->B);
  st.cur = 0;
  c->sorted = st.sorted != 0;
  c->steps_since_sort = st.since_sort;
  c->last_perm = const_cast<uint32_t*>((radix_passes(c) - 1) % 2 == 0 ? c->vals1 : c->vals2);
  c->vmax_valid = false;  // the batch keeps the speed maximum; the context's partial maxima are stale
  c->vmax_pending = false;
  c->version += 1;
  c->busy = false;
  b->rec[i] = scene_record(b, c);
}

void batch_finish_scene(mpm_batch* b, int64_t i, mpm_advance_stats* out) {
  mpm_ctx* c = b->scenes[i];
  const SceneStep st = b->step[i];
  batch_release_scene(b, i);
  uint32_t bits = st.speed_bits;
  float v;
  std::memcpy(&v, &bits, 4);
  b->speed[i] = v;
  const int64_t removed = remove_escaped(c);
  if (removed) {
    refresh_present(c);
    b->rec[i] = scene_record(b, c);
  }
  // the last G2P's speed maximum covers the particles just removed too: after a removal the next frame's CFL
  // step comes from the survivors (solver.py:176-185 reduces over the particles that remain), so batch_begin
  // reduces it afresh
  b->speed_ok[i] = removed ? 0 : 1;
  b->version[i] = c->version;
  *out = mpm_advance_stats{st.steps, b->dt[i], removed};
}

// Replay substeps: max_launches > 0 launches that many (at most until every scene's frame is done),
// 0 runs until every frame in flight is done, < 0 until the first frame in flight is done. Scenes whose
// frame ended get finished[i] = 1 and out[i] (as mpm_advance reports it).
void batch_run(mpm_batch* b, int64_t max_launches, int64_t* launched, int32_t* finished, mpm_advance_stats* out) {
  const int64_t n = (int64_t)b->scenes.size();
  *launched = 0;
  for (int64_t i = 0; i < n; ++i) {
    finished[i] = 0;
    if (b->empty_frame[i]) {
      b->empty_frame[i] = 0;
      finished[i] = 1;
      out[i] = mpm_advance_stats{0, b->frame_dt[i], 0};
    }
  }
  int64_t most = 0, least = INT64_MAX;
  for (int64_t i = 0; i < n; ++i)
    if (in_frame(b->step[i])) {
      const int64_t rem = b->step[i].steps - b->step[i].k;
      most = std::max(most, rem);
      least = std::min(least, rem);
    }
  if (most == 0) return;
  const int64_t L = max_launches > 0 ? std::min(max_launches, most) : (max_launches == 0 ? most : least);
  batch_reserve_tables(b, n);
  batch_layout(b);
  bool fp64_pass = false;
  for (const SceneDev& r : b->rec) fp64_pass = fp64_pass || r.sc.precision != kPrecisionFp32;
  if (fp64_pass != b->fp64_pass) {
    HIP_CHECK(hipStreamSynchronize(b->stream));
    batch_drop_graphs(b);
    b->fp64_pass = fp64_pass;
  }
  std::copy(b->rec.begin(), b->rec.end(), b->h_scenes);
  std::copy(b->step.begin(), b->step.end(), b->h_steps);
  HIP_CHECK(hipMemcpyAsync(b->d_scenes, b->h_scenes, n * sizeof(SceneDev), hipMemcpyHostToDevice, b->stream));
  HIP_CHECK(hipMemcpyAsync(b->d_steps, b->h_steps, n * sizeof(SceneStep), hipMemcpyHostToDevice, b->stream));
  // graphs unless kernels are being timed (events cannot bracket a graph's kernels)
  const bool graphs = b->use_graph && b->timing.mask == 0;
  if (graphs && !b->g_sort) {
    HIP_CHECK(hipStreamSynchronize(b->stream));
    b->g_sort = batch_capture(b, true);
    b->g_plain = batch_capture(b, false);
  }
  // the host replays the device's bookkeeping (k_batch_plan, k_batch_advance) to pick the
  // graph of each substep, and checks it against the device afterwards
  std::vector<SceneStep> sim = b->step;
  std::vector<char> was_running(n);
  for (int64_t i = 0; i < n; ++i) was_running[i] = in_frame(sim[i]);
  for (int64_t l = 0; l < L; ++l) {
    bool sorting = false;
    for (int64_t i = 0; i < n; ++i) {
      SceneStep& t = sim[i];
      if (!in_frame(t)) continue;
      if (!t.sorted || t.since_sort >= b->rec[i].sort_interval) {
        sorting = true;
        t.cur = 1 - t.cur;
        t.sorted = 1;
        t.since_sort = 0;
      }
      t.since_sort += 1;
      t.k += 1;
    }
    if (graphs) HIP_CHECK(hipGraphLaunch(sorting ? b->g_sort : b->g_plain, b->stream));
    else batch_enqueue(b, sorting);
  }
  b->substeps_launched += L;
  *launched = L;
  HIP_CHECK(hipMemcpyAsync(b->h_steps, b->d_steps, n * sizeof(SceneStep), hipMemcpyDeviceToHost, b->stream));
  HIP_CHECK(hipStreamSynchronize(b->stream));
  for (int64_t i = 0; i < n; ++i) {
    const SceneStep& d = b->h_steps[i];
    const SceneStep& h = sim[i];
    if (d.k != h.k || d.sorted != h.sorted || d.since_sort != h.since_sort || d.cur != h.cur)
      fail(MPM_ERR_RUNTIME, "batch bookkeeping diverged from the device");
    b->step[i] = d;
  }
  for (int64_t i = 0; i < n; ++i)
    if (was_running[i] && !in_frame(b->step[i])) {
      batch_finish_scene(b, i, &out[i]);
      finished[i] = 1;
    }
}

// one frame of every scene with frame_dt > 0, all run to completion (scenes already in a frame finish
// theirs too)
void batch_frame(mpm_batch* b, const double* frame_dt, mpm_advance_stats* out) {
  const int64_t n = (int64_t)b->scenes.size();
  if (n == 0) return;
  batch_begin(b, frame_dt);
  std::vector<int32_t> fin(n);
  std::vector<mpm_advance_stats> st(n);
  for (int64_t i = 0; i < n; ++i) out[i] = mpm_advance_stats{0, frame_dt[i], 0};
  int64_t launched = 0;
  batch_run(b, 0, &launched, fin.data(), st.data());
  for (int64_t i = 0; i < n; ++i)
    if (fin[i]) out[i] = st[i];
}

void attach_scene(mpm_batch* b, mpm_ctx* c) {
  HIP_CHECK(hipStreamSynchronize(c->own_stream));
  c->stream = b->stream;
  c->batch = b;
}

void detach_scene(mpm_ctx* c) {
  if (!c->batch) return;
  (void)hipStreamSynchronize(c->stream);
  c->stream = c->own_stream;
  c->batch = nullptr;
}

// New membership; scenes that stay keep their place in a frame in flight; scenes that leave in the middle
// of a frame keep the state they reached (a partial frame).
void batch_set_scenes(mpm_batch* b, mpm_ctx* const* scenes, int32_t n) {
  for (int32_t i = 0; i < n; ++i) {
    mpm_ctx* c = scenes[i];
    if (!c) fail(MPM_ERR_VALUE, "null scene");
    if (c->device != b->device) fail(MPM_ERR_VALUE, "scene lives on another device than the batch");
    if (c->batch && c->batch != b) fail(MPM_ERR_VALUE, "scene already belongs to another batch");
    for (int32_t j = 0; j < i; ++j)
      if (scenes[j] == c) fail(MPM_ERR_VALUE, "scene listed twice");
  }
  HIP_CHECK(hipStreamSynchronize(b->stream));
  const size_t old_n = b->scenes.size();
  std::vector<int64_t> from(n, -1);
  std::vector<char> stays(old_n, 0);
  for (int32_t i = 0; i < n; ++i)
    for (size_t j = 0; j < old_n; ++j)
      if (b->scenes[j] == scenes[i]) {
        from[i] = (int64_t)j;
        stays[j] = 1;
      }
  for (size_t j = 0; j < old_n; ++j)
    if (!stays[j]) {
      batch_release_scene(b, (int64_t)j);
      detach_scene(b->scenes[j]);
    }
  std::vector<SceneDev> rec(n);
  std::vector<SceneStep> step(n);
  std::vector<float> speed(n, 0.0f);
  std::vector<uint64_t> version(n, 0);
  std::vector<char> ok(n, 0), empty(n, 0);
  std::vector<double> fdt(n, 0.0), dt(n, 0.0);
  for (int32_t i = 0; i < n; ++i) {
    mpm_ctx* c = scenes[i];
    if (from[i] >= 0) {
      const int64_t j = from[i];
      rec[i] = b->rec[j];
      step[i] = b->step[j];
      speed[i] = b->speed[j];
      version[i] = b->version[j];
      ok[i] = b->speed_ok[j];
      empty[i] = b->empty_frame[j];
      fdt[i] = b->frame_dt[j];
      dt[i] = b->dt[j];
    } else {
      attach_scene(b, c);
      rec[i] = scene_record(b, c);
      step[i] = SceneStep{};
    }
  }
  b->scenes.assign(scenes, scenes + n);
  b->rec = std::move(rec);
  b->step = std::move(step);
  b->speed = std::move(speed);
  b->version = std::move(version);
  b->speed_ok = std::move(ok);
  b->empty_frame = std::move(empty);
  b->frame_dt = std::move(fdt);
  b->dt = std::move(dt);
  batch_drop_graphs(b);
  b->hist_tiles = -1;
}

// drops one scene from its batch (never throws: used by mpm_destroy)
void batch_remove_scene(mpm_batch* b, mpm_ctx* c) {
  (void)hipStreamSynchronize(b->stream);
  for (size_t j = 0; j < b->scenes.size(); ++j)
    if (b->scenes[j] == c) {
      c->busy = false;  // the context goes away; its partial frame does not matter
      b->scenes.erase(b->scenes.begin() + j);
      b->rec.erase(b->rec.begin() + j);
      b->step.erase(b->step.begin() + j);
      b->speed.erase(b->speed.begin() + j);
      b->version.erase(b->version.begin() + j);
      b->speed_ok.erase(b->speed_ok.begin() + j);
      b->empty_frame.erase(b->empty_frame.begin() + j);
      b->frame_dt.erase(b->frame_dt.begin() + j);
      b->dt.erase(b->dt.begin() + j);
      break;
    }
  batch_drop_graphs(b);
  b->hist_tiles = -1;
  detach_scene(c);
}

template <class Fn> int32_t run_on(int device, Fn&& fn) {
  try {
    if (device >= 0) {
      DeviceGuard g(device);
      fn();
    } else {
      fn();
    }
    return MPM_OK;
  } catch (const Failure& f) {
    return f.code;
  } catch (const std::bad_alloc&) {
    g_error = "host out of memory";
    return MPM_ERR_MEMORY;
  } catch (const std::exception& e) {
    g_error = e.what();
    return MPM_ERR_RUNTIME;
  }
}

template <class Fn> int32_t run(mpm_ctx* c, Fn&& fn) { return run_on(c ? c->device : -1, std::forward<Fn>(fn)); }

}  // namespace

// the layouts the ctypes binding mirrors (mpm_plugin_amd/_native.py; checked there by tests/test_native_abi.py)
static_assert(sizeof(mpm_domain_desc) == 184, "mpm_domain_desc layout");
static_assert(sizeof(mpm_material_desc) == 48, "mpm_material_desc layout");
static_assert(sizeof(mpm_collider_pose) == 132, "mpm_collider_pose layout");
static_assert(sizeof(mpm_advance_stats) == 24, "mpm_advance_stats layout");
static_assert(sizeof(mpm_surface_desc) == 40, "mpm_surface_desc layout");

extern "C" {

int32_t mpm_abi_version(void) { return MPM_ABI_VERSION; }

const char* mpm_last_error(void) { return g_error.c_str(); }

int32_t mpm_create(int32_t device, const mpm_domain_desc* domain, const mpm_material_desc* materials,
                   int32_t n_materials, mpm_ctx** out) {
  if (!domain || !out || n_materials < 0 || (n_materials > 0 && !materials)) {
    g_error = "mpm_create: null argument";
    return MPM_ERR_VALUE;
  }
  *out = nullptr;
  mpm_ctx* c = new (std::nothrow) mpm_ctx();
  if (!c) {
    g_error = "host out of memory";
    return MPM_ERR_MEMORY;
  }
  c->device = device;
  const int32_t rc = run(c, [&] {
    int count = 0;
    HIP_CHECK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) fail(MPM_ERR_VALUE, "device ordinal out of range");
    for (int i = 0; i < n_materials; ++i)
      if (materials[i].model < 0 || materials[i].model > MPM_MODEL_SNOW)
        fail(MPM_ERR_NOT_IMPLEMENTED, "material model is not supported by the HIP backend");
    for (int a = 0; a < 3; ++a)
      if (domain->dims[a] < 3) fail(MPM_ERR_VALUE, "grid dimensions must be at least 3 nodes");
    noalloc_volume_setup(device);
    HIP_CHECK(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
    c->stream = c->own_stream;
    c->dom = *domain;
    c->mats.assign(materials, materials + n_materials);
    c->present.assign(std::max(1, n_materials), 0);
    std::vector<MatDev> md(std::max(1, n_materials));
    for (int i = 0; i < n_materials; ++i) {
      const mpm_material_desc& m = materials[i];
      md[i] = MatDev{m.model, m.mu, m.lam, m.yield_stress, m.alpha, m.hardening, m.max_hardening,
                     m.crit_c, m.crit_s, m.p_mass};
    }
    c->d_mats = dev_alloc<MatDev>(md.size());
    HIP_CHECK(hipMemcpy(c->d_mats, md.data(), md.size() * sizeof(MatDev), hipMemcpyHostToDevice));
    GridGeom& g = c->geom;
    for (int a = 0; a < 3; ++a) {
      g.dims[a] = domain->dims[a];
      g.bdims[a] = (domain->dims[a] + 3) / 4;
      g.origin[a] = domain->origin[a];
    }
    g.dx = domain->dx;
    g.inv_dx = domain->inv_dx;
    StepConsts& s = c->sc;
    s.p_vol = domain->p_vol;
    for (int a = 0; a < 3; ++a) {
      s.gravity[a] = domain->gravity[a];
      s.wall_lo_node[a] = domain->wall_lo_node[a];
      s.wall_hi_node[a] = domain->wall_hi_node[a];
      s.lo_on[a] = domain->lo_on[a];
      s.hi_on[a] = domain->hi_on[a];
      s.wall_lo[a] = domain->wall_lo[a];
      s.wall_hi[a] = domain->wall_hi[a];
      s.clamp_lo[a] = domain->clamp_lo[a];
      s.clamp_hi[a] = domain->clamp_hi[a];
    }
    s.wall_friction = domain->wall_friction;
    s.precision = kPrecisionFallback;
    const int64_t bricks = (int64_t)g.bdims[0] * g.bdims[1] * g.bdims[2];
    if (bricks * 64 > ((int64_t)1 << 30)) fail(MPM_ERR_VALUE, "grid too large");
    c->n_bricks = (int)bricks;
    c->n_nodes = (int)(bricks * 64);
    c->n_cells = (int)(bricks * 64);
    c->key_bits = 1;
    while (((int64_t)1 << c->key_bits) < c->n_cells) ++c->key_bits;
    c->accum = dev_alloc<float4>(c->n_nodes);
    c->vel = dev_alloc<float4>(c->n_nodes);
    c->dirty = dev_alloc<uint32_t>(bricks);
    c->live = dev_alloc<uint8_t>(bricks);
    c->offsets = dev_alloc<uint32_t>((size_t)c->n_cells + 1);
    c->active = dev_alloc<uint32_t>(bricks);
    c->dirty_list = dev_alloc<uint32_t>(bricks);
    c->slot = dev_alloc<int32_t>(bricks);
    HIP_CHECK(hipMemsetAsync(c->slot, 0xff, bricks * 4, c->stream));
    c->n_active = dev_alloc<uint32_t>(1);
    c->n_fix = dev_alloc<uint32_t>(1);
    HIP_CHECK(hipMemsetAsync(c->accum, 0, (size_t)c->n_nodes * sizeof(float4), c->stream));
    HIP_CHECK(hipMemsetAsync(c->vel, 0, (size_t)c->n_nodes * sizeof(float4), c->stream));
    HIP_CHECK(hipMemsetAsync(c->dirty, 0, bricks * 4, c->stream));
    HIP_CHECK(hipMemsetAsync(c->live, 0, bricks, c->stream));
    HIP_CHECK(hipMemsetAsync(c->n_active, 0, 4, c->stream));
    // coherent: the reduction's writes reach host memory as it makes them, not at the end of its kernel
    HIP_CHECK(hipHostMalloc((void**)&c->h_vmax, 2 * sizeof(float), hipHostMallocCoherent));
    c->h_vmax[0] = 0.0f;
    reinterpret_cast<uint32_t*>(c->h_vmax)[1] = 0u;
    HIP_CHECK(hipHostGetDevicePointer((void**)&c->h_vmax_dev, c->h_vmax, 0));
    c->d_vmax = dev_alloc<float>(1);
    reserve(c, 256);
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
  if (rc != MPM_OK) {
    mpm_destroy(c);
    return rc;
  }
  *out = c;
  return MPM_OK;
}

int32_t mpm_destroy(mpm_ctx* c) {
  if (!c) return MPM_OK;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(c->device);  // best effort: destroy never throws
  if (c->batch) batch_remove_scene(c->batch, c);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_particles(c);
  for (void* p : {(void*)c->accum, (void*)c->vel, (void*)c->dirty, (void*)c->live, (void*)c->offsets,
                  (void*)c->active, (void*)c->dirty_list, (void*)c->slot, (void*)c->stage, (void*)c->in_cnt, (void*)c->in_lst, (void*)c->fx, (void*)c->n_active, (void*)c->n_fix, (void*)c->d_vmax, (void*)c->d_mats, (void*)c->d_block_max, (void*)c->d_cols,
                  (void*)c->d_tris, (void*)c->d_normals, (void*)c->d_cell_start, (void*)c->d_cell_tris,
                  (void*)c->d_cell_key, (void*)c->d_cell_far})
    dev_free(p);
  c->scratch.release();
  c->emit_scratch.release();
  c->hist.release();
  c->scan_sums.release();
  release_timing(c->timing);
  if (c->h_vmax) (void)hipHostFree(c->h_vmax);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  (void)hipSetDevice(prev);
  delete c;
  return MPM_OK;
}

int32_t mpm_add_particles(mpm_ctx* c, int64_t n, const float* x, const float* v, const int32_t* mat,
                          const int32_t* emitter) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    if (n < 0) fail(MPM_ERR_VALUE, "negative particle count");
    if (n == 0) return;
    if (!x || !v || !mat || !emitter) fail(MPM_ERR_VALUE, "null particle array");
    check_materials(c, mat, n);
    if (c->n + n > (int64_t)INT32_MAX) fail(MPM_ERR_VALUE, "too many particles");
    reserve(c, c->n + n);
    const int64_t p0 = c->n;
    upload_planes(c, c->A.x(), x, 3, n, p0);
    upload_planes(c, c->A.v(), v, 3, n, p0);
    HIP_CHECK(hipMemcpyAsync(c->A.mat() + p0, mat, n * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(c->emitter + p0, emitter, n * 4, hipMemcpyHostToDevice, c->stream));
    k_init_new<<<grid_for(n), kBlock, 0, c->stream>>>(c->parts(), c->d_mats, p0, (int)n);
    check_launch("k_init_new");
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->n = p0 + n;
    for (int64_t i = 0; i < n; ++i) c->present[mat[i]] = 1;
    state_changed(c);
  });
}

int32_t mpm_load_state(mpm_ctx* c, int64_t n, const float* x, const float* v, const float* Cm, const float* F,
                       const float* Jp, const int32_t* mat, const int32_t* emitter) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c); load_state(c, n, x, v, Cm, F, Jp, mat, emitter); });
}

int64_t mpm_particle_count(const mpm_ctx* c) { return c ? c->n : 0; }

int32_t mpm_get_state(mpm_ctx* c, float* x, float* v, float* Cm, float* F, float* Jp, int32_t* mat,
                      int32_t* emitter) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    if (x) download_planes(c, c->A.x(), x, 3);
    if (v) download_planes(c, c->A.v(), v, 3);
    if (Cm) download_C(c, Cm);
    if (F) download_planes(c, c->A.F(), F, 9);
    if (Jp) download_planes(c, c->A.Jp(), Jp, 1);
    if (mat) download_planes(c, reinterpret_cast<const float*>(c->A.mat()), reinterpret_cast<float*>(mat), 1);
    if (emitter && c->n) {
      HIP_CHECK(hipMemcpyAsync(emitter, c->emitter, c->n * 4, hipMemcpyDeviceToHost, c->stream));
      HIP_CHECK(hipStreamSynchronize(c->stream));
    }
  });
}

int32_t mpm_set_colliders(mpm_ctx* c, int32_t n, const int32_t* n_verts, const int32_t* n_tris, const float* verts,
                          const int32_t* tris, const int32_t* closed, const float* friction) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    if (n < 0 || n > kMaxColliders) fail(MPM_ERR_VALUE, "the HIP backend takes at most 8 colliders per domain");
    if (n > 0 && (!n_verts || !n_tris || !closed || !friction)) fail(MPM_ERR_VALUE, "null collider array");
    std::vector<ColliderHost> hs(n);
    int64_t vo = 0, to = 0;
    for (int i = 0; i < n; ++i) {
      const int nv = n_verts[i], nt = n_tris[i];
      if (nv < 0 || nt < 0) fail(MPM_ERR_VALUE, "negative mesh size");
      const float* V = verts + 3 * vo;
      const int32_t* T = tris + 3 * to;
      for (int64_t k = 0; k < 3 * (int64_t)nt; ++k)
        if (T[k] < 0 || T[k] >= nv) fail(MPM_ERR_VALUE, "collider triangle index out of range");
      ColliderHost& h = hs[i];
      mesh_tables(V, nv, T, nt, h);
      ColliderDev& d = h.dev;
      std::memset(&d, 0, sizeof d);
      for (int a = 0; a < 3; ++a) {
        float lo = nv ? V[a] : 0.0f, hi = lo;
        for (int k = 1; k < nv; ++k) {
          lo = std::min(lo, V[3 * k + a]);
          hi = std::max(hi, V[3 * k + a]);
        }
        d.aabb_lo[a] = lo;
        d.aabb_hi[a] = hi;
      }
      d.closed = closed[i] ? 1 : 0;
      d.friction = friction[i];
      d.n_tris = nt;
      d.q0[3] = d.q1[3] = 1.0f;  // identity until the first pose upload
      d.S0[0] = d.S0[4] = d.S0[8] = d.S1[0] = d.S1[4] = d.S1[8] = 1.0f;
      d.margin = (float)(3.0 * c->dom.dx_f64);
      vo += nv;
      to += nt;
    }
    for (auto& h : hs) build_candidates(c, h, h.dev.margin * 1.25f);
    c->colliders = std::move(hs);
    upload_colliders(c);
  });
}

// ---- collider poses from world matrices (solver.py:154-172, 258-260; geometry.py:30-49) -----------------
// M = [A | t]: translation t, rotation R and symmetric stretch S with A = R S (the polar decomposition; a
// mirrored A keeps R a proper rotation and its reflection goes into S through the sign of the smallest
// singular value), from an SVD of A: cyclic Jacobi on A^T A in FP64 for V and the singular values, then
// U from A V (Gram-Schmidt, third column the cross product, so U is proper and the third singular value
// carries the sign of det A). R = U V^T, S = V diag(sig) V^T; R becomes a quaternion by the reference's
// branch-free conversion. The margin's scale is the smallest unsigned singular value.
struct PoseParts {
  double t[3], q[4], S[9], smin;
};

namespace {

void sym_eigen3(double B[3][3], double V[3][3], double lam[3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 64; ++sweep) {
    const double off = B[0][1] * B[0][1] + B[0][2] * B[0][2] + B[1][2] * B[1][2];
    const double diag = B[0][0] * B[0][0] + B[1][1] * B[1][1] + B[2][2] * B[2][2];
    if (off <= 1e-36 * diag || off == 0.0) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (B[p][q] == 0.0) continue;
        const double theta = (B[q][q] - B[p][p]) / (2.0 * B[p][q]);
        const double tt = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double cs = 1.0 / std::sqrt(tt * tt + 1.0), sn = tt * cs;
        for (int k = 0; k < 3; ++k) {  // B <- B J (columns p, q)
          const double bp = B[k][p], bq = B[k][q];
          B[k][p] = cs * bp - sn * bq;
          B[k][q] = sn * bp + cs * bq;
        }
        for (int k = 0; k < 3; ++k) {  // B <- J^T B (rows p, q)
          const double bp = B[p][k], bq = B[q][k];
          B[p][k] = cs * bp - sn * bq;
          B[q][k] = sn * bp + cs * bq;
        }
        for (int k = 0; k < 3; ++k) {  // V <- V J
          const double vp = V[k][p], vq = V[k][q];
          V[k][p] = cs * vp - sn * vq;
          V[k][q] = sn * vp + cs * vq;
        }
      }
  }
  for (int i = 0; i < 3; ++i) lam[i] = B[i][i];
}

// m: row-major 4x4 world matrix
PoseParts decompose_pose(const float* m) {
  PoseParts out{};
  double A[3][3];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) A[i][j] = (double)m[4 * i + j];
    out.t[i] = (double)m[4 * i + 3];
  }
  double B[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) B[i][j] = A[0][i] * A[0][j] + A[1][i] * A[1][j] + A[2][i] * A[2][j];
  double V[3][3], lam[3];
  sym_eigen3(B, V, lam);
  int order[3] = {0, 1, 2};  // descending eigenvalues
  std::sort(order, order + 3, [&](int a, int b) { return lam[a] > lam[b]; });
  double v[3][3];  // v[i]: the i-th right singular vector
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) v[i][k] = V[k][order[i]];
  const auto cross = [](const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  const auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  double c3[3];
  cross(v[0], v[1], c3);
  if (dot(c3, v[2]) < 0.0)
    for (int k = 0; k < 3; ++k) v[2][k] = -v[2][k];  // det V = +1
  double Av[3][3];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) Av[i][k] = A[k][0] * v[i][0] + A[k][1] * v[i][1] + A[k][2] * v[i][2];
  double u[3][3];
  const auto unit_or = [&](double* w, const double* fallback) {
    const double nrm = std::sqrt(dot(w, w));
    if (nrm > 0.0) {
      for (int k = 0; k < 3; ++k) w[k] /= nrm;
    } else {
      for (int k = 0; k < 3; ++k) w[k] = fallback[k];
    }
  };
  for (int k = 0; k < 3; ++k) u[0][k] = Av[0][k];
  unit_or(u[0], v[0]);
  const double p01 = dot(u[0], Av[1]);
  for (int k = 0; k < 3; ++k) u[1][k] = Av[1][k] - p01 * u[0][k];
  {
    // a rank-1 A leaves u1 free: any unit vector orthogonal to u0
    double e[3] = {0.0, 0.0, 0.0};
    e[std::fabs(u[0][0]) < 0.9 ? 0 : 1] = 1.0;
    double w[3];
    cross(u[0], e, w);
    unit_or(w, e);
    unit_or(u[1], w);
  }
  cross(u[0], u[1], u[2]);
  const double sig[3] = {dot(u[0], Av[0]), dot(u[1], Av[1]), dot(u[2], Av[2])};
  out.smin = std::sqrt(std::max(lam[order[2]], 0.0));
  double R[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      R[i][j] = u[0][i] * v[0][j] + u[1][i] * v[1][j] + u[2][i] * v[2][j];
      out.S[3 * i + j] = sig[0] * v[0][i] * v[0][j] + sig[1] * v[1][i] * v[1][j] + sig[2] * v[2][i] * v[2][j];
    }
  const double d0 = R[0][0], d1 = R[1][1], d2 = R[2][2];
  double q[4] = {std::copysign(0.5 * std::sqrt(std::max(0.0, 1.0 + d0 - d1 - d2)), R[2][1] - R[1][2]),
                 std::copysign(0.5 * std::sqrt(std::max(0.0, 1.0 - d0 + d1 - d2)), R[0][2] - R[2][0]),
                 std::copysign(0.5 * std::sqrt(std::max(0.0, 1.0 - d0 - d1 + d2)), R[1][0] - R[0][1]),
                 0.5 * std::sqrt(std::max(0.0, 1.0 + d0 + d1 + d2))};
  const double qn = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int k = 0; k < 4; ++k) out.q[k] = q[k] / qn;
  return out;
}

// the decomposed start/end poses of one collider and its AABB margin (solver.py:160-172, 258-260)
mpm_collider_pose pose_from_matrices(const float* start, const float* end, double dx) {
  const PoseParts a = decompose_pose(start);
  PoseParts b = decompose_pose(end);
  if (a.q[0] * b.q[0] + a.q[1] * b.q[1] + a.q[2] * b.q[2] + a.q[3] * b.q[3] < 0.0)
    for (double& x : b.q) x = -x;  // slerp the short way round
  mpm_collider_pose p{};
  for (int k = 0; k < 3; ++k) {
    p.t0[k] = (float)a.t[k];
    p.t1[k] = (float)b.t[k];
  }
  for (int k = 0; k < 4; ++k) {
    p.q0[k] = (float)a.q[k];
    p.q1[k] = (float)b.q[k];
  }
  for (int k = 0; k < 9; ++k) {
    p.S0[k] = (float)a.S[k];
    p.S1[k] = (float)b.S[k];
  }
  const double inv_scale = std::max(1.0 / std::max(a.smin, 1e-12), 1.0 / std::max(b.smin, 1e-12));
  p.margin = (float)(3.0 * dx * inv_scale);
  return p;
}

}  // namespace

int32_t mpm_set_collider_poses(mpm_ctx* c, int32_t n, const mpm_collider_pose* poses) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    if (n != (int32_t)c->colliders.size())
      fail(MPM_ERR_VALUE, "expected " + std::to_string(c->colliders.size()) + " collider transforms");
    if (n > 0 && !poses) fail(MPM_ERR_VALUE, "null pose array");
    bool rebuilt = false;
    for (int i = 0; i < n; ++i) {
      ColliderHost& h = c->colliders[i];
      ColliderDev& d = h.dev;
      const mpm_collider_pose& p = poses[i];
      std::memcpy(d.t0, p.t0, sizeof d.t0);
      std::memcpy(d.t1, p.t1, sizeof d.t1);
      std::memcpy(d.q0, p.q0, sizeof d.q0);
      std::memcpy(d.q1, p.q1, sizeof d.q1);
      std::memcpy(d.S0, p.S0, sizeof d.S0);
      std::memcpy(d.S1, p.S1, sizeof d.S1);
      d.margin = p.margin;
      if (d.margin > h.built_pad) {
        build_candidates(c, h, d.margin * 1.25f);
        rebuilt = true;
      }
    }
    if (n == 0) return;
    // the previous substeps may still read the records: order the copy after them on the stream
    if (rebuilt) upload_colliders(c);
    else upload_collider_records(c);
  });
}

int32_t mpm_set_collider_transforms(mpm_ctx* c, int32_t n, const float* start, const float* end) {
  if (!c) return MPM_ERR_VALUE;
  std::vector<mpm_collider_pose> poses;
  const int32_t rc = run(c, [&] {
    if (n != (int32_t)c->colliders.size())
      fail(MPM_ERR_VALUE, "expected " + std::to_string(c->colliders.size()) + " collider transforms");
    if (n > 0 && (!start || !end)) fail(MPM_ERR_VALUE, "null transform array");
    for (int32_t i = 0; i < n; ++i)
      for (int k = 0; k < 16; ++k)
        if (!std::isfinite(start[16 * i + k]) || !std::isfinite(end[16 * i + k]))
          fail(MPM_ERR_VALUE, "non-finite collider transform");
    for (int32_t i = 0; i < n; ++i) poses.push_back(pose_from_matrices(start + 16 * i, end + 16 * i, c->dom.dx_f64));
  });
  if (rc != MPM_OK) return rc;
  return mpm_set_collider_poses(c, n, poses.data());
}

int32_t mpm_decompose_transform(const float* m, double* t, double* q, double* S, double* smin) {
  if (!m || !t || !q || !S || !smin) return MPM_ERR_VALUE;
  const PoseParts p = decompose_pose(m);
  std::memcpy(t, p.t, sizeof p.t);
  std::memcpy(q, p.q, sizeof p.q);
  std::memcpy(S, p.S, sizeof p.S);
  *smin = p.smin;
  return MPM_OK;
}

int32_t mpm_advance(mpm_ctx* c, double frame_dt, mpm_advance_stats* out) {
  if (!c || !out) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    if (c->n == 0) {
      *out = mpm_advance_stats{0, frame_dt, 0};
      return;
    }
    // the substep count divides by frame_dt (solver.py:183-185 raises on zero or non-finite values); a
    // valid frame_dt always gives at least one substep, which consumes the binning enqueued below
    if (frame_dt == 0.0 || !std::isfinite(frame_dt)) fail(MPM_ERR_VALUE, "frame_dt must be finite and nonzero");
    c->frame_dt = frame_dt;
    double wave = 0.0;
    for (size_t i = 0; i < c->mats.size(); ++i)
      if (c->present[i]) wave = std::max(wave, c->mats[i].wave_speed);
    // The first substep's binning needs the positions but not dt, and its P2G only dt, which it derives
    // on the device from the reduced speed exactly as below (cfl_dt): both go on the stream behind the
    // speed reduction, so the GPU works through them while the host waits for the speed and picks dt.
    request_speed(c);
    const DriftLists pre = c->p2g_mode == MPM_P2G_ATOMIC ? DriftLists{} : bin_for_p2g(c);
    const bool early_p2g = c->p2g_mode != MPM_P2G_ATOMIC;
    if (early_p2g) run_p2g(c, 0.0f, &pre, CflStep{c->d_vmax, wave, c->dom.dx_f64, frame_dt});
    double vmax = (double)max_speed(c);
    if (!std::isfinite(vmax)) vmax = 0.0;  // already blown up: keep stepping so the NaN reaches x
    const double cfl_dt = std::min(0.4 * c->dom.dx_f64 / (wave + vmax), frame_dt);
    const int64_t steps = (int64_t)std::ceil(frame_dt / cfl_dt - 1e-9);
    const double dt = frame_dt / (double)steps;
    for (int64_t k = 0; k < steps; ++k)
      substep(c, dt, (double)(k + 1) / (double)steps, frame_dt, k == 0 ? &pre : nullptr, k == steps - 1,
              k == 0 && early_p2g);
    // the removal count needs the positions (a host wait, only with open sides); otherwise the substeps
    // stay in flight and the host prepares the next call meanwhile (every read of the state waits for them)
    const int64_t removed = remove_escaped(c);
    if (removed) refresh_present(c);
    request_speed(c);  // the next advance's CFL step
    *out = mpm_advance_stats{steps, dt, removed};
  });
}

int32_t mpm_substep(mpm_ctx* c, double dt, double s, double frame_dt) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    c->frame_dt = frame_dt;
    substep(c, dt, s, frame_dt);
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

// Lattice points inside a closed mesh (geometry.py:89-155 before the jitter), in the reference's order.
int32_t mpm_lattice_inside(mpm_ctx* c, const double* verts, int64_t n_verts, const int32_t* tris, int64_t n_tris,
                           double spacing, int64_t capacity, double* out, int64_t* count) {
  if (!c || !count) return MPM_ERR_VALUE;
  return run(c, [&] {
    *count = 0;
    if (!(spacing > 0.0) || !std::isfinite(spacing)) fail(MPM_ERR_VALUE, "spacing must be positive and finite");
    if (n_verts < 0 || n_tris < 0 || capacity < 0) fail(MPM_ERR_VALUE, "negative size");
    if (n_verts == 0 || n_tris == 0) return;
    if (!verts || !tris) fail(MPM_ERR_VALUE, "null mesh array");
    if (n_tris > (int64_t)kScanTile * kScanTile || n_verts > (int64_t)INT32_MAX)
      fail(MPM_ERR_VALUE, "mesh too large for the emission kernels");
    for (int64_t i = 0; i < 3 * n_tris; ++i)
      if (tris[i] < 0 || tris[i] >= n_verts) fail(MPM_ERR_VALUE, "triangle index out of range");
    for (int64_t i = 0; i < 3 * n_verts; ++i)
      if (!std::isfinite(verts[i])) fail(MPM_ERR_VALUE, "non-finite vertex coordinate");
    // the lattice as numpy computes it: first = ceil(min / spacing - 0.5), last = floor(max / spacing - 0.5)
    EmitLattice L{};
    for (int a = 0; a < 3; ++a) {
      double lo = verts[a], hi = verts[a];
      for (int64_t v = 1; v < n_verts; ++v) {
        lo = std::min(lo, verts[3 * v + a]);
        hi = std::max(hi, verts[3 * v + a]);
      }
      L.first[a] = (int64_t)std::ceil(lo / spacing - 0.5);
      const int64_t last = (int64_t)std::floor(hi / spacing - 0.5);
      if (last < L.first[a]) return;
      if (last - L.first[a] + 1 > (int64_t)INT32_MAX) fail(MPM_ERR_VALUE, "lattice too large");
      L.n[a] = (int)(last - L.first[a] + 1);
    }
    L.spacing = spacing;
    L.nudge[0] = 1.234567e-4 * spacing;  // the reference's ray nudge (geometry.py:106)
    L.nudge[1] = 2.345678e-4 * spacing;
    const int64_t n_cols = (int64_t)L.n[0] * L.n[1];
    if (n_cols > (int64_t)kScanTile * kScanTile) fail(MPM_ERR_VALUE, "lattice too large for the emission kernels");
    // device tables, carved from one buffer (8-byte aligned pieces)
    const auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_verts = al((size_t)n_verts * 24), b_tris = al((size_t)n_tris * 12), b_tri = al((size_t)n_tris * 4),
                 b_col = al((size_t)n_cols * 4);
    char* base = static_cast<char*>(c->emit_scratch.get(b_verts + b_tris + 2 * b_tri + 5 * b_col));
    double* d_verts = reinterpret_cast<double*>(base);
    int32_t* d_tris = reinterpret_cast<int32_t*>(base + b_verts);
    uint32_t* pairs = reinterpret_cast<uint32_t*>(base + b_verts + b_tris);
    uint32_t* pair_off = reinterpret_cast<uint32_t*>(base + b_verts + b_tris + b_tri);
    uint32_t* col_count = reinterpret_cast<uint32_t*>(base + b_verts + b_tris + 2 * b_tri);
    uint32_t* col_off = col_count + b_col / 4;
    uint32_t* col_fill = col_off + b_col / 4;
    uint32_t* pts_count = col_fill + b_col / 4;
    uint32_t* pts_off = pts_count + b_col / 4;
    HIP_CHECK(hipMemcpyAsync(d_verts, verts, (size_t)n_verts * 24, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(d_tris, tris, (size_t)n_tris * 12, hipMemcpyHostToDevice, c->stream));
    k_emit_tri_columns<<<grid_for(n_tris), kBlock, 0, c->stream>>>(L, d_verts, d_tris, (int)n_tris, pairs);
    check_launch("k_emit_tri_columns");
    // (triangle, column) pairs: their total must fit the 32-bit scan
    std::vector<uint32_t> h_pairs(n_tris);
    HIP_CHECK(hipMemcpyAsync(h_pairs.data(), pairs, (size_t)n_tris * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    uint64_t n_pairs = 0;
    for (uint32_t p : h_pairs) n_pairs += p;
    if (n_pairs == 0) return;
    if (n_pairs >= (1ull << 32)) fail(MPM_ERR_VALUE, "mesh covers too many lattice columns");
    scan_u32(c->scan_sums, c->stream, pairs, pair_off, n_tris);
    HIP_CHECK(hipMemsetAsync(col_count, 0, (size_t)n_cols * 4, c->stream));
    const unsigned pair_blocks = (unsigned)((n_pairs + kBlock - 1) / kBlock);
    k_emit_pair_hits<<<pair_blocks, kBlock, 0, c->stream>>>(L, d_verts, d_tris, (int)n_tris, pair_off,
                                                            (uint32_t)n_pairs, 0, col_count, nullptr, nullptr, nullptr);
    check_launch("k_emit_pair_hits");
    scan_u32(c->scan_sums, c->stream, col_count, col_off, n_cols);
    uint32_t tail[2];
    HIP_CHECK(hipMemcpyAsync(&tail[0], col_off + n_cols - 1, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipMemcpyAsync(&tail[1], col_count + n_cols - 1, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    const uint64_t n_hits = (uint64_t)tail[0] + tail[1];
    if (n_hits == 0) return;
    double* hit_z = dev_alloc<double>(n_hits);
    struct Free {
      void* p;
      ~Free() { dev_free(p); }
    } free_hits{hit_z};
    HIP_CHECK(hipMemsetAsync(col_fill, 0, (size_t)n_cols * 4, c->stream));
    k_emit_pair_hits<<<pair_blocks, kBlock, 0, c->stream>>>(L, d_verts, d_tris, (int)n_tris, pair_off,
                                                            (uint32_t)n_pairs, 1, nullptr, col_off, col_fill, hit_z);
    check_launch("k_emit_pair_hits");
    k_emit_column_points<<<grid_for(n_cols), kBlock, 0, c->stream>>>(L, n_cols, col_count, col_off, hit_z, 0,
                                                                     pts_count, nullptr, nullptr);
    check_launch("k_emit_column_points");
    scan_u32(c->scan_sums, c->stream, pts_count, pts_off, n_cols);
    HIP_CHECK(hipMemcpyAsync(&tail[0], pts_off + n_cols - 1, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipMemcpyAsync(&tail[1], pts_count + n_cols - 1, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    const int64_t n_pts = (int64_t)tail[0] + tail[1];
    *count = n_pts;
    if (n_pts == 0 || n_pts > capacity || !out) return;
    double* d_out = dev_alloc<double>(3 * (size_t)n_pts);
    Free free_out{d_out};
    k_emit_column_points<<<grid_for(n_cols), kBlock, 0, c->stream>>>(L, n_cols, col_count, col_off, hit_z, 1,
                                                                     nullptr, pts_off, d_out);
    check_launch("k_emit_column_points");
    HIP_CHECK(hipMemcpyAsync(out, d_out, 3 * (size_t)n_pts * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_is_occupied(mpm_ctx* c, int64_t m, const float* pts, uint8_t* out) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    if (m <= 0) return;
    if (!pts || !out) fail(MPM_ERR_VALUE, "null query array");
    const size_t vol_bytes = (size_t)c->n_nodes * sizeof(float);
    const size_t pts_bytes = (size_t)m * 3 * sizeof(float);
    char* base = static_cast<char*>(c->scratch.get(vol_bytes + pts_bytes + (size_t)m + 256));
    float* vol = reinterpret_cast<float*>(base);
    float* dpts = reinterpret_cast<float*>(base + vol_bytes);
    uint8_t* dout = reinterpret_cast<uint8_t*>(base + vol_bytes + pts_bytes);
    HIP_CHECK(hipMemsetAsync(vol, 0, vol_bytes, c->stream));
    HIP_CHECK(hipMemcpyAsync(dpts, pts, pts_bytes, hipMemcpyHostToDevice, c->stream));
    if (c->n) {
      k_scatter_volume<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->parts(), c->geom, c->sc.p_vol, (int)c->n, vol);
      check_launch("k_scatter_volume");
    }
    k_query_fill<<<grid_for(m), kBlock, 0, c->stream>>>(c->geom, dpts, (int)m, c->dom.cell_vol, vol, dout);
    check_launch("k_query_fill");
    HIP_CHECK(hipMemcpyAsync(out, dout, (size_t)m, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_is_occupied_noalloc(mpm_ctx* c, int64_t m, const float* points_dev, uint8_t* out_dev) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    if (m < 0 || m > INT32_MAX) fail(MPM_ERR_VALUE, "mpm_is_occupied_noalloc: query count out of range");
    if (m == 0) return;
    if (!points_dev || !out_dev) fail(MPM_ERR_VALUE, "mpm_is_occupied_noalloc: null query array");
    if ((int64_t)c->n_nodes > MPM_NOALLOC_MAX_NODES)
      fail(MPM_ERR_VALUE, "mpm_is_occupied_noalloc: the grid has " + std::to_string(c->n_nodes) +
                              " nodes (bricks of 4^3), more than the " + std::to_string(MPM_NOALLOC_MAX_NODES) +
                              " its fixed volume buffer holds; use mpm_is_occupied");
    float* vol = noalloc_volume(c->device);
    // the buffer keeps the previous call's volumes: clear the nodes this grid uses with a kernel (no memset
    // node, so the sequence can sit in a captured graph)
    k_zero_u32<<<grid_for(c->n_nodes), kBlock, 0, c->stream>>>(reinterpret_cast<uint32_t*>(vol), c->n_nodes);
    check_launch("k_zero_u32");
    if (c->n) {
      k_scatter_volume<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->parts(), c->geom, c->sc.p_vol, (int)c->n, vol);
      check_launch("k_scatter_volume");
    }
    k_query_fill<<<grid_for(m), kBlock, 0, c->stream>>>(c->geom, points_dev, (int)m, c->dom.cell_vol, vol, out_dev);
    check_launch("k_query_fill");
  });
}

int32_t mpm_debug_p2g(mpm_ctx* c, double dt) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    clear_grid(c);
    c->sorted = false;  // binned path: bin from the current positions, as a first substep does
    run_p2g(c, (float)dt);
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_debug_grid_update(mpm_ctx* c, double dt, double s, double frame_dt) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    c->frame_dt = frame_dt;
    run_grid(c, (float)dt, (float)s, (float)(1.0 / frame_dt));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_debug_g2p(mpm_ctx* c, double dt, double s, double frame_dt) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    c->frame_dt = frame_dt;
    run_g2p(c, (float)dt, (float)s, (float)(1.0 / frame_dt));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_get_grid(mpm_ctx* c, int32_t which, float* grid_v, float* grid_m) {
  if (!c || !grid_v || !grid_m || (which != 0 && which != 1)) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    const int64_t total = (int64_t)c->geom.dims[0] * c->geom.dims[1] * c->geom.dims[2];
    float4* src = which == 0 ? c->accum : c->vel;
    // the P2G output is the accumulator alone after the reference-order P2G or an atomic window flush
    const bool resolve = which == 0 && c->p2g_mode != MPM_P2G_ATOMIC &&
                         (c->p2g_mode == MPM_P2G_FIXED || c->flush_stage() != nullptr);
    if (resolve) {
      // the P2G output lives in the fixed-point accumulator, or in the accumulator plus the staged windows
      src = static_cast<float4*>(c->scratch.get((size_t)c->n_nodes * sizeof(float4) + total * 4 * sizeof(float)));
      if (c->p2g_mode == MPM_P2G_FIXED) {
        k_fixed_to_float<<<grid_for(c->n_nodes), kBlock, 0, c->stream>>>(c->fx, c->n_nodes, 1.0 / c->fx_scale, src);
        check_launch("k_fixed_to_float");
      } else {
        k_p2g_resolve<<<grid_for(c->n_nodes), kBlock, 0, c->stream>>>(c->geom, c->accum, c->n_nodes, c->flush_stage(),
                                                                     c->slot, src);
        check_launch("k_p2g_resolve");
      }
    }
    float* stage = resolve ? reinterpret_cast<float*>(src + c->n_nodes)
                           : static_cast<float*>(c->scratch.get(total * 4 * sizeof(float)));
    k_grid_to_ref<<<grid_for(total), kBlock, 0, c->stream>>>(c->geom, src,
                                                             which == 0 ? nullptr : c->live, stage, stage + 3 * total);
    check_launch("k_grid_to_ref");
    HIP_CHECK(hipMemcpyAsync(grid_v, stage, total * 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipMemcpyAsync(grid_m, stage + 3 * total, total * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_set_grid(mpm_ctx* c, const float* grid_v, const float* grid_m) {
  if (!c || !grid_v || !grid_m) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    const int64_t total = (int64_t)c->geom.dims[0] * c->geom.dims[1] * c->geom.dims[2];
    float* stage = static_cast<float*>(c->scratch.get(total * 4 * sizeof(float)));
    HIP_CHECK(hipMemcpyAsync(stage, grid_v, total * 3 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(stage + 3 * total, grid_m, total * sizeof(float), hipMemcpyHostToDevice, c->stream));
    k_grid_from_ref<<<grid_for(total), kBlock, 0, c->stream>>>(c->geom, stage, stage + 3 * total, c->vel);
    check_launch("k_grid_from_ref");
    HIP_CHECK(hipMemsetAsync(c->live, 1, c->n_bricks, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_get_binning(mpm_ctx* c, int32_t* keys, int32_t* perm, int32_t* offsets, int64_t* n_bins) {
  if (!c || !n_bins) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    *n_bins = c->n_cells;
    if (!c->last_perm) {
      if (keys || perm || offsets) fail(MPM_ERR_VALUE, "no binning has run yet");
      return;
    }
    if (keys && c->n) HIP_CHECK(hipMemcpyAsync(keys, c->keys0, c->n * 4, hipMemcpyDeviceToHost, c->stream));
    if (perm && c->n) HIP_CHECK(hipMemcpyAsync(perm, c->last_perm, c->n * 4, hipMemcpyDeviceToHost, c->stream));
    if (offsets)
      HIP_CHECK(hipMemcpyAsync(offsets, c->offsets, ((size_t)c->n_cells + 1) * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_debug_sort(mpm_ctx* c) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    sort_particles(c);
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_set_option(mpm_ctx* c, int32_t option, int64_t value) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    require_idle(c);
    switch (option) {
      case MPM_OPT_P2G_MODE:
        if (value != MPM_P2G_BINNED && value != MPM_P2G_ATOMIC && value != MPM_P2G_FIXED)
          fail(MPM_ERR_VALUE, "p2g mode must be 0, 1 or 2");
        if (value != c->p2g_mode) {
          // the dense path leaves no dirty flags; hand over a clean accumulator either way
          HIP_CHECK(hipMemsetAsync(c->accum, 0, (size_t)c->n_nodes * sizeof(float4), c->stream));
          HIP_CHECK(hipMemsetAsync(c->dirty, 0, (size_t)c->n_bricks * 4, c->stream));
          if (value == MPM_P2G_FIXED && !c->fx) {
            c->fx = dev_alloc<unsigned long long>((size_t)c->n_nodes * 4);
            HIP_CHECK(hipMemsetAsync(c->fx, 0, (size_t)c->n_nodes * 4 * sizeof(unsigned long long), c->stream));
            // one particle's mass maps to 2^32: contributions keep ~1e-10 of it, sums of 2^31 particle
            // masses at speeds up to ~1e3 still fit in int64
            float pm = 0.0f;
            for (const auto& m : c->mats) pm = std::max(pm, m.p_mass);
            c->fx_scale = std::ldexp(1.0, 32) / (double)(pm > 0.0f ? pm : 1.0f);
          }
          if (value != MPM_P2G_FIXED && c->fx) {
            HIP_CHECK(hipStreamSynchronize(c->stream));
            dev_free(c->fx);
            c->fx = nullptr;
          }
          HIP_CHECK(hipStreamSynchronize(c->stream));
        }
        c->p2g_mode = (int)value;
        c->sorted = false;
        break;
      case MPM_OPT_PRECISION:
        if (value < kPrecisionFp32 || value > kPrecisionFp64) fail(MPM_ERR_VALUE, "precision must be 0, 1 or 2");
        if (value != c->sc.precision) {
          c->sc.precision = (int)value;
          // the stored stress belongs to the old precision
          if (c->n) {
            k_stress<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->parts(), c->d_mats, (int)c->n, c->sc.precision);
            check_launch("k_stress");
            HIP_CHECK(hipStreamSynchronize(c->stream));
          }
        }
        break;
      case MPM_OPT_SORT_INTERVAL:
        if (value < 1) fail(MPM_ERR_VALUE, "sort interval must be >= 1");
        c->sort_interval = value;
        break;
      case MPM_OPT_P2G_BLOCKS:
        if (value < 0) fail(MPM_ERR_VALUE, "block count must be >= 0");
        c->p2g_blocks = value;
        break;
      case MPM_OPT_P2G_FLUSH:
        if (value != MPM_P2G_FLUSH_STAGED && value != MPM_P2G_FLUSH_ATOMIC) fail(MPM_ERR_VALUE, "p2g flush must be 0 or 1");
        c->p2g_flush = (int)value;
        break;
      case MPM_OPT_P2G_WAVES:
        if (value != 1 && value != 2 && value != 4) fail(MPM_ERR_VALUE, "waves per brick must be 1, 2 or 4");
        c->p2g_waves = (int)value;
        break;
      default:
        fail(MPM_ERR_VALUE, "unknown option");
    }
  });
}

int64_t mpm_get_option(const mpm_ctx* c, int32_t option) {
  if (!c) return -1;
  switch (option) {
    case MPM_OPT_P2G_MODE: return c->p2g_mode;
    case MPM_OPT_PRECISION: return c->sc.precision;
    case MPM_OPT_SORT_INTERVAL: return c->sort_interval;
    case MPM_OPT_P2G_BLOCKS: return c->p2g_blocks;
    case MPM_OPT_P2G_WAVES: return c->p2g_waves;
    case MPM_OPT_P2G_FLUSH: return c->p2g_flush;
    default: return -1;
  }
}

int32_t mpm_set_timing(mpm_ctx* c, int32_t mask) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    collect_timing(c);
    c->timing.mask = (uint32_t)mask & (uint32_t)MPM_TIMING_ALL;
  });
}

int32_t mpm_get_kernel_time(mpm_ctx* c, int32_t kernel, double* total_ms, int64_t* launches) {
  if (!c || kernel < 0 || kernel >= MPM_K_COUNT || !total_ms || !launches) return MPM_ERR_VALUE;
  return run(c, [&] {
    collect_timing(c);
    *total_ms = c->timing.total_ms[kernel];
    *launches = c->timing.launches[kernel];
  });
}

int32_t mpm_reset_timing(mpm_ctx* c) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    collect_timing(c);
    for (int k = 0; k < MPM_K_COUNT; ++k) {
      c->timing.total_ms[k] = 0.0;
      c->timing.launches[k] = 0;
    }
  });
}

int32_t mpm_synchronize(mpm_ctx* c) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] { HIP_CHECK(hipStreamSynchronize(c->stream)); });
}

// ---- batches ------------------------------------------------------------------------------------
int32_t mpm_batch_create(int32_t device, mpm_batch** out) {
  if (!out) return MPM_ERR_VALUE;
  *out = nullptr;
  mpm_batch* b = new (std::nothrow) mpm_batch();
  if (!b) {
    g_error = "host out of memory";
    return MPM_ERR_MEMORY;
  }
  b->device = device;
  const int32_t rc = run_on(-1, [&] {
    int count = 0;
    HIP_CHECK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) fail(MPM_ERR_VALUE, "device ordinal out of range");
    DeviceGuard g(device);
    HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  });
  if (rc != MPM_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return MPM_OK;
}

int32_t mpm_batch_destroy(mpm_batch* b) {
  if (!b) return MPM_OK;
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (mpm_ctx* c : b->scenes) detach_scene(c);
  batch_drop_graphs(b);
  batch_free_tables(b);
  b->hist.release();
  b->scan_sums.release();
  release_timing(b->timing);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  (void)hipSetDevice(prev);
  delete b;
  return MPM_OK;
}

int32_t mpm_batch_set_scenes(mpm_batch* b, mpm_ctx* const* scenes, int32_t n) {
  if (!b || n < 0 || (n > 0 && !scenes)) return MPM_ERR_VALUE;
  return run_on(b->device, [&] { batch_set_scenes(b, scenes, n); });
}

int32_t mpm_batch_scene_count(const mpm_batch* b) { return b ? (int32_t)b->scenes.size() : 0; }

int32_t mpm_batch_advance(mpm_batch* b, const double* frame_dt, mpm_advance_stats* out) {
  if (!b || (!b->scenes.empty() && (!frame_dt || !out))) return MPM_ERR_VALUE;
  return run_on(b->device, [&] { batch_frame(b, frame_dt, out); });
}

int32_t mpm_batch_begin(mpm_batch* b, const double* frame_dt) {
  if (!b || (!b->scenes.empty() && !frame_dt)) return MPM_ERR_VALUE;
  return run_on(b->device, [&] { batch_begin(b, frame_dt); });
}

int32_t mpm_batch_run(mpm_batch* b, int64_t max_launches, int64_t* launched, int32_t* finished,
                      mpm_advance_stats* out) {
  if (!b || !launched || (!b->scenes.empty() && (!finished || !out))) return MPM_ERR_VALUE;
  return run_on(b->device, [&] { batch_run(b, max_launches, launched, finished, out); });
}

int32_t mpm_batch_progress(const mpm_batch* b, int32_t* substeps_done, int32_t* substeps_total) {
  if (!b || (!b->scenes.empty() && (!substeps_done || !substeps_total))) return MPM_ERR_VALUE;
  for (size_t i = 0; i < b->scenes.size(); ++i) {
    substeps_done[i] = b->step[i].k;
    substeps_total[i] = b->step[i].steps;
  }
  return MPM_OK;
}

int32_t mpm_batch_set_option(mpm_batch* b, int32_t option, int64_t value) {
  if (!b) return MPM_ERR_VALUE;
  return run_on(b->device, [&] {
    int64_t* slot = nullptr;
    switch (option) {
      case MPM_BATCH_OPT_P2G_BLOCKS:
      case MPM_BATCH_OPT_G2P_BLOCKS:
      case MPM_BATCH_OPT_GRID_BLOCKS:
        if (value < 0 || value > (1 << 20)) fail(MPM_ERR_VALUE, "workgroup counts must be in [0, 2^20]");
        slot = option == MPM_BATCH_OPT_P2G_BLOCKS ? &b->p2g_blocks
                                                  : option == MPM_BATCH_OPT_G2P_BLOCKS ? &b->g2p_blocks : &b->grid_blocks;
        break;
      case MPM_BATCH_OPT_GRAPH:
        if (value != 0 && value != 1) fail(MPM_ERR_VALUE, "graph option is 0 or 1");
        slot = &b->use_graph;
        break;
      case MPM_BATCH_OPT_SORT_INTERVAL:
        if (value < 0) fail(MPM_ERR_VALUE, "sort interval must be >= 0");
        slot = &b->sort_interval;
        break;
      case MPM_BATCH_OPT_G2P_TILE:
        if (value != 64 && value != 128 && value != 256) fail(MPM_ERR_VALUE, "G2P tile must be 64, 128 or 256");
        slot = &b->g2p_tile;
        break;
      case MPM_BATCH_OPT_P2G_WAVES:
        if (value != 1 && value != 2 && value != 4) fail(MPM_ERR_VALUE, "waves per brick must be 1, 2 or 4");
        slot = &b->p2g_waves;
        break;
      default:
        fail(MPM_ERR_VALUE, "unknown batch option");
    }
    if (*slot == value) return;
    HIP_CHECK(hipStreamSynchronize(b->stream));
    *slot = value;
    batch_drop_graphs(b);
    // the sort cadence lives in the scene records
    for (size_t i = 0; i < b->scenes.size(); ++i) b->rec[i].sort_interval = scene_record(b, b->scenes[i]).sort_interval;
  });
}

int64_t mpm_batch_get_option(const mpm_batch* b, int32_t option) {
  if (!b) return -1;
  switch (option) {
    case MPM_BATCH_OPT_P2G_BLOCKS:
      return batch_blocks(b->p2g_blocks, 8192);
    case MPM_BATCH_OPT_G2P_BLOCKS:
      return batch_blocks(b->g2p_blocks, 2048);
    case MPM_BATCH_OPT_GRID_BLOCKS:
      return batch_blocks(b->grid_blocks, 2048);
    case MPM_BATCH_OPT_G2P_TILE:
      return b->g2p_tile;
    case MPM_BATCH_OPT_P2G_WAVES:
      return b->p2g_waves;
    case MPM_BATCH_OPT_GRAPH:
      return b->use_graph;
    case MPM_BATCH_OPT_SORT_INTERVAL:
      return b->sort_interval;
    case MPM_BATCH_OPT_SUBSTEPS_LAUNCHED:
      return b->substeps_launched;
    default:
      return -1;
  }
}

int32_t mpm_batch_set_timing(mpm_batch* b, int32_t on) {
  if (!b || (on != 0 && on != 1)) return MPM_ERR_VALUE;
  return run_on(b->device, [&] {
    collect_timing(b->timing, b->stream);
    b->timing.mask = on ? (uint32_t)((1 << MPM_BK_COUNT) - 1) : 0u;
  });
}

int32_t mpm_batch_get_kernel_time(mpm_batch* b, int32_t kernel, double* total_ms, int64_t* launches) {
  if (!b || kernel < 0 || kernel >= MPM_BK_COUNT || !total_ms || !launches) return MPM_ERR_VALUE;
  return run_on(b->device, [&] {
    collect_timing(b->timing, b->stream);
    *total_ms = b->timing.total_ms[kernel];
    *launches = b->timing.launches[kernel];
  });
}

int32_t mpm_batch_reset_timing(mpm_batch* b) {
  if (!b) return MPM_ERR_VALUE;
  return run_on(b->device, [&] {
    collect_timing(b->timing, b->stream);
    for (int k = 0; k < kTimedClasses; ++k) {
      b->timing.total_ms[k] = 0.0;
      b->timing.launches[k] = 0;
    }
  });
}

}  // extern "C"

// ---- surfacing (mpm_surface_*) ----------------------------------------------------------------
struct mpm_surface {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[5] = {};
  hipEvent_t ev_upload[2] = {};  // around mpm_surface_upload's copies and bounds
  bool from_upload = false;      // the last build meshed particles left by mpm_surface_upload
  Scratch tables, parts, density, momentum, crossed, vid, ntri, tri_off, verts, vvel, tris, scan_sums;
  Scratch bounds;
  SurfGrid grid{};
  int64_t n_nodes = 0, n_verts = 0, n_tris = 0;
  int64_t uploaded = -1;  // particles on the device (mpm_surface_upload or a build), -1: none
  bool built = false;
};

namespace {

// nodes per field at most: the edge ids (3 per node) and the launch sizes stay within 32 bits
constexpr int64_t kMaxSurfNodes = int64_t(1) << 30;
constexpr int kSurfBoundsBlocks = 1024;  // partial bounds per upload, reduced on the host
constexpr int kGatherSplitReach = 5;     // largest kernel reach (cells) gathered with two waves per brick

// exclusive scan of n 32-bit counts (out != in) of any length: chunks of what one scan_u32 pass takes,
// each after the first carried by the total of the chunk before
void scan_u32_long(Scratch& sums, hipStream_t stream, const uint32_t* in, uint32_t* out, int64_t n) {
  const int64_t chunk = (int64_t)kScanTile * kScanTile;
  for (int64_t base = 0; base < n; base += chunk) {
    const int64_t len = std::min(chunk, n - base);
    scan_u32(sums, stream, in + base, out + base, len);
    if (base > 0) {
      k_scan_carry<<<grid_for(len), kBlock, 0, stream>>>(out + base, len, in + base - 1, out + base - 1);
      check_launch("k_scan_carry");
    }
  }
}

void destroy_surface(mpm_surface* s) {
  int prev = 0;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (Scratch* b : {&s->tables, &s->parts, &s->density, &s->momentum, &s->crossed, &s->vid, &s->ntri,
                     &s->tri_off, &s->verts, &s->vvel, &s->tris, &s->scan_sums, &s->bounds})
    b->release();
  for (hipEvent_t e : s->ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : s->ev_upload)
    if (e) (void)hipEventDestroy(e);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  (void)hipSetDevice(prev);
  delete s;
}

bool finite_positive(float x) { return std::isfinite(x) && x > 0.0f; }

// the build: particles uploaded from x, v (upload) or left by mpm_surface_upload
void surface_build(mpm_surface* s, const mpm_surface_desc* d, int64_t n, const float* x, const float* v, bool upload,
                   int64_t* n_verts, int64_t* n_tris) {
  *n_verts = *n_tris = 0;
  s->built = false;
  s->n_verts = s->n_tris = 0;
  if (n < 0) fail(MPM_ERR_VALUE, "negative particle count");
  if (n > INT32_MAX) fail(MPM_ERR_VALUE, "too many particles for one surfacing call");
  if (upload && n > 0 && (!x || !v)) fail(MPM_ERR_VALUE, "null particle array");
  int64_t nodes = 1;
  for (int a = 0; a < 3; ++a) {
    if (d->dims[a] < 2) fail(MPM_ERR_VALUE, "the field needs at least 2 nodes per axis");
    if (!std::isfinite(d->origin[a])) fail(MPM_ERR_VALUE, "non-finite field origin");
    nodes *= d->dims[a];
    if (nodes > kMaxSurfNodes) fail(MPM_ERR_VALUE, "field too large for one surfacing call");
  }
  if (!finite_positive(d->cell) || !finite_positive(d->h) || !finite_positive(d->inv_rest) || !std::isfinite(d->iso))
    fail(MPM_ERR_VALUE, "cell, h and inv_rest must be positive and finite, iso finite");
  SurfGrid G{};
  for (int a = 0; a < 3; ++a) {
    G.dims[a] = d->dims[a];
    G.origin[a] = d->origin[a];
  }
  G.cell = d->cell;
  G.inv_cell = 1.0f / d->cell;
  G.h = d->h;
  s->grid = G;
  s->n_nodes = nodes;
  const int64_t cells = (int64_t)(G.dims[0] - 1) * (G.dims[1] - 1) * (G.dims[2] - 1);
  float* density = static_cast<float*>(s->density.get((size_t)nodes * 4));
  float* momentum = static_cast<float*>(s->momentum.get((size_t)nodes * 12));
  uint32_t* crossed = static_cast<uint32_t*>(s->crossed.get((size_t)nodes * 12));
  uint32_t* vid = static_cast<uint32_t*>(s->vid.get((size_t)nodes * 12));
  uint32_t* ntri = static_cast<uint32_t*>(s->ntri.get((size_t)cells * 4));
  uint32_t* tri_off = static_cast<uint32_t*>(s->tri_off.get((size_t)cells * 4));
  // particles: sorted grid-unit positions and velocities (float4 each), then the host rows, keys, ranks
  char* pbase = static_cast<char*>(s->parts.get((size_t)std::max<int64_t>(n, 1) * 64));
  float4* gs = reinterpret_cast<float4*>(pbase);
  float4* vs = gs + n;
  float* px = reinterpret_cast<float*>(pbase + (size_t)n * 32);
  float* pv = px + 3 * n;
  uint32_t* key = reinterpret_cast<uint32_t*>(pbase + (size_t)n * 56);
  uint32_t* rank = key + n;
  // the cell counts and starts borrow the marching cubes' edge arrays (3 nodes each, free until then)
  uint32_t* cell_count = crossed;
  uint32_t* cell_start = vid;
  const McTables* T = static_cast<const McTables*>(s->tables.ptr);
  HIP_CHECK(hipEventRecord(s->ev[0], s->stream));
  if (upload && n > 0) {
    HIP_CHECK(hipMemcpyAsync(px, x, (size_t)n * 12, hipMemcpyHostToDevice, s->stream));
    HIP_CHECK(hipMemcpyAsync(pv, v, (size_t)n * 12, hipMemcpyHostToDevice, s->stream));
  }
  s->uploaded = n;
  HIP_CHECK(hipMemsetAsync(cell_count, 0, (size_t)(nodes + 1) * 4, s->stream));
  HIP_CHECK(hipEventRecord(s->ev[1], s->stream));
  if (n > 0) {
    k_surf_cell_keys<<<grid_for(n), kBlock, 0, s->stream>>>(G, px, (int)n, key, rank, cell_count);
    check_launch("k_surf_cell_keys");
  }
  scan_u32_long(s->scan_sums, s->stream, cell_count, cell_start, nodes + 1);
  if (n > 0) {
    k_surf_sort_particles<<<grid_for(n), kBlock, 0, s->stream>>>(G, px, pv, (int)n, key, rank, cell_start, gs, vs);
    check_launch("k_surf_sort_particles");
  }
  // every node is written (zero where no particle reaches it)
  const int reach = (int)std::min(std::ceil((double)(G.h * G.inv_cell)), (double)(1 << 20));
  const int nb[3] = {(G.dims[0] + 3) / 4, (G.dims[1] + 3) / 4, (G.dims[2] + 3) / 4};
  const int64_t n_bricks = (int64_t)nb[0] * nb[1] * nb[2];
  // two waves per brick up to a reach of kGatherSplitReach cells, one beyond (measured: 1M particles at reach
  // 4, 0.90 against 0.95 ms and 1.35 against 1.6 ms on the mix's final frame; at reach 8, 6.5 against 5.6 ms)
  const unsigned groups = (unsigned)((n_bricks + kGatherBricks - 1) / kGatherBricks);
  if (reach <= kGatherSplitReach)
    k_surf_gather<2><<<groups, 64 * 2 * kGatherBricks, 0, s->stream>>>(G, reach, gs, vs, cell_start, nb[1], nb[2],
                                                                        n_bricks, density, momentum);
  else
    k_surf_gather<1><<<groups, 64 * kGatherBricks, 0, s->stream>>>(G, reach, gs, vs, cell_start, nb[1], nb[2],
                                                                    n_bricks, density, momentum);
  check_launch("k_surf_gather");
  HIP_CHECK(hipEventRecord(s->ev[2], s->stream));
  k_surf_normalize<<<grid_for(nodes), kBlock, 0, s->stream>>>(nodes, density, momentum, d->inv_rest);
  check_launch("k_surf_normalize");
  HIP_CHECK(hipEventRecord(s->ev[3], s->stream));
  // pass 0: crossed edges and triangles per cell, then their offsets
  k_surf_edges<<<grid_for(3 * nodes), kBlock, 0, s->stream>>>(G, density, momentum, d->iso, 0, crossed, nullptr,
                                                               nullptr, nullptr);
  check_launch("k_surf_edges");
  k_surf_cubes<<<grid_for(cells), kBlock, 0, s->stream>>>(G, density, d->iso, T, 0, ntri, nullptr, nullptr,
                                                           nullptr);
  check_launch("k_surf_cubes");
  scan_u32_long(s->scan_sums, s->stream, crossed, vid, 3 * nodes);
  scan_u32_long(s->scan_sums, s->stream, ntri, tri_off, cells);
  uint32_t tail[4];
  HIP_CHECK(hipMemcpyAsync(&tail[0], vid + 3 * nodes - 1, 4, hipMemcpyDeviceToHost, s->stream));
  HIP_CHECK(hipMemcpyAsync(&tail[1], crossed + 3 * nodes - 1, 4, hipMemcpyDeviceToHost, s->stream));
  HIP_CHECK(hipMemcpyAsync(&tail[2], tri_off + cells - 1, 4, hipMemcpyDeviceToHost, s->stream));
  HIP_CHECK(hipMemcpyAsync(&tail[3], ntri + cells - 1, 4, hipMemcpyDeviceToHost, s->stream));
  HIP_CHECK(hipStreamSynchronize(s->stream));
  const int64_t nv = (int64_t)tail[0] + tail[1], nt = (int64_t)tail[2] + tail[3];
  if (nt > INT32_MAX) fail(MPM_ERR_VALUE, "surface has too many triangles for one call");
  if (nv > 0) {
    float* verts = static_cast<float*>(s->verts.get((size_t)nv * 12));
    float* vvel = static_cast<float*>(s->vvel.get((size_t)nv * 12));
    k_surf_edges<<<grid_for(3 * nodes), kBlock, 0, s->stream>>>(G, density, momentum, d->iso, 1, nullptr, vid,
                                                                 verts, vvel);
    check_launch("k_surf_edges");
  }
  if (nt > 0) {
    int32_t* tris = static_cast<int32_t*>(s->tris.get((size_t)nt * 12));
    k_surf_cubes<<<grid_for(cells), kBlock, 0, s->stream>>>(G, density, d->iso, T, 1, nullptr, tri_off, vid,
                                                             tris);
    check_launch("k_surf_cubes");
  }
  HIP_CHECK(hipEventRecord(s->ev[4], s->stream));
  HIP_CHECK(hipStreamSynchronize(s->stream));
  s->n_verts = nv;
  s->n_tris = nt;
  s->built = true;
  *n_verts = nv;
  *n_tris = nt;
}


}  // namespace

extern "C" {

int32_t mpm_surface_create(int32_t device, mpm_surface** out) {
  if (!out) return MPM_ERR_VALUE;
  *out = nullptr;
  mpm_surface* s = new (std::nothrow) mpm_surface();
  if (!s) {
    g_error = "host out of memory";
    return MPM_ERR_MEMORY;
  }
  s->device = device;
  const int32_t rc = run_on(device, [&] {
    HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    for (hipEvent_t& e : s->ev) HIP_CHECK(hipEventCreate(&e));
    for (hipEvent_t& e : s->ev_upload) HIP_CHECK(hipEventCreate(&e));
    const McTables t = build_mc_tables();
    HIP_CHECK(hipMemcpyAsync(s->tables.get(sizeof(McTables)), &t, sizeof(McTables), hipMemcpyHostToDevice,
                             s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
  });
  if (rc != MPM_OK) {
    destroy_surface(s);
    return rc;
  }
  *out = s;
  return MPM_OK;
}

int32_t mpm_surface_destroy(mpm_surface* s) {
  if (s) destroy_surface(s);
  return MPM_OK;
}

int32_t mpm_surface_build(mpm_surface* s, const mpm_surface_desc* d, int64_t n, const float* x, const float* v,
                          int64_t* n_verts, int64_t* n_tris) {
  if (!s || !d || !n_verts || !n_tris) return MPM_ERR_VALUE;
  return run_on(s->device, [&] {
    s->from_upload = false;
    surface_build(s, d, n, x, v, true, n_verts, n_tris);
  });
}

int32_t mpm_surface_upload(mpm_surface* s, int64_t n, const float* x, const float* v, float* lo, float* hi) {
  if (!s || !lo || !hi) return MPM_ERR_VALUE;
  return run_on(s->device, [&] {
    s->uploaded = -1;
    s->built = false;
    if (n <= 0) fail(MPM_ERR_VALUE, "no particles to upload");
    if (n > INT32_MAX) fail(MPM_ERR_VALUE, "too many particles for one surfacing call");
    if (!x || !v) fail(MPM_ERR_VALUE, "null particle array");
    char* pbase = static_cast<char*>(s->parts.get((size_t)n * 64));  // the build's particle layout
    float* px = reinterpret_cast<float*>(pbase + (size_t)n * 32);
    float* pv = px + 3 * n;
    float* partial = static_cast<float*>(s->bounds.get((size_t)kSurfBoundsBlocks * 8 * sizeof(float)));
    HIP_CHECK(hipEventRecord(s->ev_upload[0], s->stream));
    HIP_CHECK(hipMemcpyAsync(px, x, (size_t)n * 12, hipMemcpyHostToDevice, s->stream));
    HIP_CHECK(hipMemcpyAsync(pv, v, (size_t)n * 12, hipMemcpyHostToDevice, s->stream));
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, kSurfBoundsBlocks);
    k_surf_bounds<<<blocks, 256, 0, s->stream>>>(px, (int)n, partial);
    check_launch("k_surf_bounds");
    HIP_CHECK(hipEventRecord(s->ev_upload[1], s->stream));
    std::vector<float> h((size_t)blocks * 8);
    HIP_CHECK(hipMemcpyAsync(h.data(), partial, h.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    bool nan = false;
    for (int a = 0; a < 3; ++a) {
      lo[a] = h[a];
      hi[a] = h[3 + a];
    }
    for (int b = 0; b < blocks; ++b) {
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], h[8 * b + a]);
        hi[a] = std::max(hi[a], h[8 * b + 3 + a]);
      }
      nan = nan || h[8 * b + 6] != 0.0f;
    }
    if (nan)  // as numpy's min and max of a column holding NaN
      for (int a = 0; a < 3; ++a) lo[a] = hi[a] = std::numeric_limits<float>::quiet_NaN();
    s->uploaded = n;
  });
}

int32_t mpm_surface_build_uploaded(mpm_surface* s, const mpm_surface_desc* d, int64_t* n_verts, int64_t* n_tris) {
  if (!s || !d || !n_verts || !n_tris) return MPM_ERR_VALUE;
  return run_on(s->device, [&] {
    if (s->uploaded < 0) fail(MPM_ERR_VALUE, "no particles uploaded");
    surface_build(s, d, s->uploaded, nullptr, nullptr, false, n_verts, n_tris);
    s->from_upload = true;
  });
}

int32_t mpm_surface_fetch(mpm_surface* s, float* verts, int32_t* tris, float* vel) {
  if (!s) return MPM_ERR_VALUE;
  return run_on(s->device, [&] {
    if (!s->built) fail(MPM_ERR_VALUE, "no surface built");
    if (verts && s->n_verts)
      HIP_CHECK(hipMemcpyAsync(verts, s->verts.ptr, (size_t)s->n_verts * 12, hipMemcpyDeviceToHost, s->stream));
    if (vel && s->n_verts)
      HIP_CHECK(hipMemcpyAsync(vel, s->vvel.ptr, (size_t)s->n_verts * 12, hipMemcpyDeviceToHost, s->stream));
    if (tris && s->n_tris)
      HIP_CHECK(hipMemcpyAsync(tris, s->tris.ptr, (size_t)s->n_tris * 12, hipMemcpyDeviceToHost, s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
  });
}

int32_t mpm_surface_get_field(mpm_surface* s, float* density, float* velocity) {
  if (!s) return MPM_ERR_VALUE;
  return run_on(s->device, [&] {
    if (!s->built) fail(MPM_ERR_VALUE, "no surface built");
    if (density)
      HIP_CHECK(hipMemcpyAsync(density, s->density.ptr, (size_t)s->n_nodes * 4, hipMemcpyDeviceToHost, s->stream));
    if (velocity)
      HIP_CHECK(hipMemcpyAsync(velocity, s->momentum.ptr, (size_t)s->n_nodes * 12, hipMemcpyDeviceToHost,
                               s->stream));
    HIP_CHECK(hipStreamSynchronize(s->stream));
  });
}

int32_t mpm_surface_times(const mpm_surface* s, float* ms4) {
  if (!s || !ms4) return MPM_ERR_VALUE;
  return run_on(s->device, [&] {
    if (!s->built) fail(MPM_ERR_VALUE, "no surface built");
    const int from[4] = {0, 1, 2, 3}, to[4] = {1, 2, 3, 4};
    for (int i = 0; i < 4; ++i) HIP_CHECK(hipEventElapsedTime(&ms4[i], s->ev[from[i]], s->ev[to[i]]));
    if (s->from_upload) {  // the copies ran in mpm_surface_upload (with the bounds); the build only cleared
      float up = 0.0f;
      HIP_CHECK(hipEventElapsedTime(&up, s->ev_upload[0], s->ev_upload[1]));
      ms4[0] += up;
    }
  });
}

int32_t mpm_surface_tables(int32_t* count, int8_t* edges) {
  if (!count || !edges) return MPM_ERR_VALUE;
  return run_on(-1, [&] {
    const McTables t = build_mc_tables();
    std::memcpy(count, t.count, sizeof(t.count));
    std::memcpy(edges, t.edge, sizeof(t.edge));
  });
}

int32_t mpm_device_memory(int32_t device, int64_t* free_bytes, int64_t* total_bytes) {
  if (!free_bytes || !total_bytes) return MPM_ERR_VALUE;
  return run_on(device, [&] {
    size_t f = 0, t = 0;
    HIP_CHECK(hipMemGetInfo(&f, &t));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
  });
}

}  // extern "C"
