// libmpm_hip.so: context management and the C-ABI declared in include/mpm_hip.h.
//
// One context = one simulation domain on one GPU with its own HIP stream. The host side of the
// reference solver (src/mpm_core/backends/warp/solver.py) maps onto the entry points as documented in
// the header; the kernels live in kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/mpm_hip.h"
#include "kernels.hpp"

using namespace mpm;

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

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// device scratch that grows on demand
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
  ~Scratch() { dev_free(ptr); }
};

struct Timing {
  bool on = false;
  struct Pending {
    int kernel;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  double total_ms[MPM_K_COUNT] = {};
  int64_t launches[MPM_K_COUNT] = {};
};

}  // namespace

struct mpm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  mpm_domain_desc dom{};
  std::vector<mpm_material_desc> mats;
  MatDev* d_mats = nullptr;

  // particles
  int64_t n = 0, cap = 0;
  float *x = nullptr, *v = nullptr, *C = nullptr, *F = nullptr, *Jp = nullptr;
  int32_t *mat = nullptr, *emitter = nullptr, *pid = nullptr;
  std::vector<char> present;  // material index -> any particle of it exists

  // grid
  GridGeom geom{};
  StepConsts sc{};
  int n_nodes = 0;  // nodes in the bricked allocation
  float4 *accum = nullptr, *vel = nullptr;

  unsigned int* d_vmax = nullptr;
  Scratch scratch, scratch2;
  Timing timing;
  double frame_dt = 1.0;

  Particles parts() const { return Particles{x, v, C, F, Jp, mat, pid, cap}; }
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

struct Timed {
  mpm_ctx* c;
  int kernel;
  hipEvent_t a = nullptr;
  Timed(mpm_ctx* ctx, int k) : c(ctx), kernel(k) {
    if (c->timing.on) {
      a = pool_event(c->timing);
      HIP_CHECK(hipEventRecord(a, c->stream));
    }
  }
  ~Timed() noexcept(false) {
    if (a) {
      hipEvent_t b = pool_event(c->timing);
      HIP_CHECK(hipEventRecord(b, c->stream));
      c->timing.pending.push_back({kernel, a, b});
    }
  }
};

void collect_timing(mpm_ctx* c) {
  if (c->timing.pending.empty()) return;
  HIP_CHECK(hipStreamSynchronize(c->stream));
  for (auto& p : c->timing.pending) {
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
    c->timing.total_ms[p.kernel] += ms;
    c->timing.launches[p.kernel] += 1;
    c->timing.pool.push_back(p.a);
    c->timing.pool.push_back(p.b);
  }
  c->timing.pending.clear();
}

void check_launch(const char* what) { hip_check(hipGetLastError(), what); }

// ---- particle storage ------------------------------------------------------------------------
void free_particles(mpm_ctx* c) {
  for (void* p : {(void*)c->x, (void*)c->v, (void*)c->C, (void*)c->F, (void*)c->Jp, (void*)c->mat,
                  (void*)c->emitter, (void*)c->pid})
    dev_free(p);
  c->x = c->v = c->C = c->F = c->Jp = nullptr;
  c->mat = c->emitter = c->pid = nullptr;
  c->cap = 0;
}

// grow capacity to hold `need` particles, keeping the first c->n
void reserve(mpm_ctx* c, int64_t need) {
  if (need <= c->cap) return;
  int64_t cap = std::max<int64_t>(need, c->cap + c->cap / 2);
  cap = (cap + 255) & ~int64_t(255);
  float* x = dev_alloc<float>(3 * cap);
  float* v = dev_alloc<float>(3 * cap);
  float* C = dev_alloc<float>(9 * cap);
  float* F = dev_alloc<float>(9 * cap);
  float* Jp = dev_alloc<float>(cap);
  int32_t* mat = dev_alloc<int32_t>(cap);
  int32_t* em = dev_alloc<int32_t>(cap);
  int32_t* pid = dev_alloc<int32_t>(cap);
  if (c->n) {
    auto planes = [&](float* dst, const float* src, int comps) {
      for (int k = 0; k < comps; ++k)
        HIP_CHECK(hipMemcpyAsync(dst + k * cap, src + k * c->cap, c->n * sizeof(float), hipMemcpyDeviceToDevice,
                                 c->stream));
    };
    planes(x, c->x, 3);
    planes(v, c->v, 3);
    planes(C, c->C, 9);
    planes(F, c->F, 9);
    planes(Jp, c->Jp, 1);
    HIP_CHECK(hipMemcpyAsync(mat, c->mat, c->n * 4, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(em, c->emitter, c->n * 4, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(pid, c->pid, c->n * 4, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  const int64_t n = c->n;
  free_particles(c);
  c->x = x; c->v = v; c->C = C; c->F = F; c->Jp = Jp;
  c->mat = mat; c->emitter = em; c->pid = pid;
  c->cap = cap;
  c->n = n;
}

// upload `n` host AoS rows of `comps` floats into planes starting at slot p0
void upload_planes(mpm_ctx* c, float* planes, const float* host, int comps, int64_t n, int64_t p0) {
  if (n == 0) return;
  float* stage = static_cast<float*>(c->scratch.get(n * comps * sizeof(float)));
  HIP_CHECK(hipMemcpyAsync(stage, host, n * comps * sizeof(float), hipMemcpyHostToDevice, c->stream));
  k_unpack<<<grid_for(n), kBlock, 0, c->stream>>>(stage, comps, (int)n, c->cap, p0, planes);
  check_launch("k_unpack");
}

void download_planes(mpm_ctx* c, const float* planes, float* host, int comps) {
  if (c->n == 0) return;
  float* stage = static_cast<float*>(c->scratch.get(c->n * comps * sizeof(float)));
  k_pack<<<grid_for(c->n), kBlock, 0, c->stream>>>(planes, c->cap, c->pid, comps, (int)c->n, stage);
  check_launch("k_pack");
  HIP_CHECK(hipMemcpyAsync(host, stage, c->n * comps * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

void check_materials(mpm_ctx* c, const int32_t* mat, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (mat[i] < 0 || mat[i] >= (int32_t)c->mats.size()) fail(MPM_ERR_VALUE, "material id out of range");
}

// ---- substep phases --------------------------------------------------------------------------
void run_p2g(mpm_ctx* c, float dt) {
  if (c->n == 0) return;
  Timed t(c, MPM_K_P2G);
  k_p2g_atomic<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->parts(), c->d_mats, c->geom, c->sc, (int)c->n, dt,
                                                         c->accum);
  check_launch("k_p2g_atomic");
}

void run_grid(mpm_ctx* c, float dt) {
  Timed t(c, MPM_K_GRID);
  k_grid_update_dense<<<grid_for(c->n_nodes), kBlock, 0, c->stream>>>(c->geom, c->sc, dt, c->n_nodes, c->accum,
                                                                        c->vel);
  check_launch("k_grid_update_dense");
}

void run_g2p(mpm_ctx* c, float dt) {
  if (c->n == 0) return;
  Timed t(c, MPM_K_G2P);
  k_g2p<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->parts(), c->d_mats, c->geom, c->sc, (int)c->n, dt, c->vel);
  check_launch("k_g2p");
}

void substep(mpm_ctx* c, double dt, double s, double frame_dt) {
  (void)s;
  (void)frame_dt;
  const float dtf = (float)dt;
  run_p2g(c, dtf);
  run_grid(c, dtf);
  run_g2p(c, dtf);
}

float max_speed(mpm_ctx* c) {
  HIP_CHECK(hipMemsetAsync(c->d_vmax, 0, sizeof(unsigned int), c->stream));
  {
    Timed t(c, MPM_K_OTHER);
    k_max_speed<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->parts(), (int)c->n, c->d_vmax);
    check_launch("k_max_speed");
  }
  unsigned int bits = 0;
  HIP_CHECK(hipMemcpyAsync(&bits, c->d_vmax, sizeof bits, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  float f;
  std::memcpy(&f, &bits, sizeof f);
  return f;
}

void refresh_present(mpm_ctx* c) {
  std::fill(c->present.begin(), c->present.end(), 0);
  if (c->n == 0) return;
  std::vector<int32_t> m(c->n);
  HIP_CHECK(hipMemcpyAsync(m.data(), c->mat, c->n * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  for (int32_t id : m) c->present[id] = 1;
}

void load_state(mpm_ctx* c, int64_t n, const float* x, const float* v, const float* Cm, const float* F,
                const float* Jp, const int32_t* mat, const int32_t* emitter) {
  if (n < 0) fail(MPM_ERR_VALUE, "negative particle count");
  if (n > 0 && (!x || !v || !Cm || !F || !Jp || !mat || !emitter)) fail(MPM_ERR_VALUE, "null state array");
  check_materials(c, mat, n);
  if (n > (int64_t)INT32_MAX) fail(MPM_ERR_VALUE, "too many particles");
  c->n = 0;
  reserve(c, n);
  upload_planes(c, c->x, x, 3, n, 0);
  upload_planes(c, c->v, v, 3, n, 0);
  upload_planes(c, c->C, Cm, 9, n, 0);
  upload_planes(c, c->F, F, 9, n, 0);
  upload_planes(c, c->Jp, Jp, 1, n, 0);
  if (n) {
    HIP_CHECK(hipMemcpyAsync(c->mat, mat, n * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(c->emitter, emitter, n * 4, hipMemcpyHostToDevice, c->stream));
    std::vector<int32_t> ids(n);
    for (int64_t i = 0; i < n; ++i) ids[i] = (int32_t)i;
    HIP_CHECK(hipMemcpyAsync(c->pid, ids.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  c->n = n;
  std::fill(c->present.begin(), c->present.end(), 0);
  for (int64_t i = 0; i < n; ++i) c->present[mat[i]] = 1;
}

// drop particles more than one cell past an open side; keeps insertion order of the survivors
int64_t remove_escaped(mpm_ctx* c) {
  const mpm_domain_desc& d = c->dom;
  bool any_open = false;
  for (int a = 0; a < 3; ++a) any_open = any_open || !d.lo_on[a] || !d.hi_on[a];
  if (!any_open || c->n == 0) return 0;
  const int64_t n = c->n;
  int32_t* keep = static_cast<int32_t*>(c->scratch2.get(n * sizeof(int32_t)));
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
  download_planes(c, c->x, x.data(), 3);
  download_planes(c, c->v, v.data(), 3);
  download_planes(c, c->C, C.data(), 9);
  download_planes(c, c->F, F.data(), 9);
  download_planes(c, c->Jp, Jp.data(), 1);
  download_planes(c, reinterpret_cast<float*>(c->mat), reinterpret_cast<float*>(mat.data()), 1);
  download_planes(c, reinterpret_cast<float*>(c->emitter), reinterpret_cast<float*>(em.data()), 1);
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

template <class Fn> int32_t run(mpm_ctx* c, Fn&& fn) {
  try {
    if (c) {
      DeviceGuard g(c->device);
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
}  // namespace

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
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
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
    const int64_t nodes = (int64_t)g.bdims[0] * g.bdims[1] * g.bdims[2] * 64;
    if (nodes > (int64_t)1 << 30) fail(MPM_ERR_VALUE, "grid too large");
    c->n_nodes = (int)nodes;
    c->accum = dev_alloc<float4>(nodes);
    c->vel = dev_alloc<float4>(nodes);
    HIP_CHECK(hipMemsetAsync(c->accum, 0, nodes * sizeof(float4), c->stream));
    HIP_CHECK(hipMemsetAsync(c->vel, 0, nodes * sizeof(float4), c->stream));
    c->d_vmax = dev_alloc<unsigned int>(1);
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
  {
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(c->device);  // best effort: destroy never throws
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_particles(c);
    dev_free(c->accum);
    dev_free(c->vel);
    dev_free(c->d_mats);
    dev_free(c->d_vmax);
    c->scratch.~Scratch();
    new (&c->scratch) Scratch();
    c->scratch2.~Scratch();
    new (&c->scratch2) Scratch();
    for (auto& p : c->timing.pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : c->timing.pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    (void)hipSetDevice(prev);
  }
  delete c;
  return MPM_OK;
}

int32_t mpm_add_particles(mpm_ctx* c, int64_t n, const float* x, const float* v, const int32_t* mat,
                          const int32_t* emitter) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    if (n < 0) fail(MPM_ERR_VALUE, "negative particle count");
    if (n == 0) return;
    if (!x || !v || !mat || !emitter) fail(MPM_ERR_VALUE, "null particle array");
    check_materials(c, mat, n);
    if (c->n + n > (int64_t)INT32_MAX) fail(MPM_ERR_VALUE, "too many particles");
    reserve(c, c->n + n);
    const int64_t p0 = c->n;
    upload_planes(c, c->x, x, 3, n, p0);
    upload_planes(c, c->v, v, 3, n, p0);
    HIP_CHECK(hipMemcpyAsync(c->mat + p0, mat, n * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(c->emitter + p0, emitter, n * 4, hipMemcpyHostToDevice, c->stream));
    c->n = p0 + n;
    k_init_new<<<grid_for(n), kBlock, 0, c->stream>>>(c->parts(), p0, (int)n);
    check_launch("k_init_new");
    HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; ++i) c->present[mat[i]] = 1;
  });
}

int32_t mpm_load_state(mpm_ctx* c, int64_t n, const float* x, const float* v, const float* Cm, const float* F,
                       const float* Jp, const int32_t* mat, const int32_t* emitter) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] { load_state(c, n, x, v, Cm, F, Jp, mat, emitter); });
}

int64_t mpm_particle_count(const mpm_ctx* c) { return c ? c->n : 0; }

int32_t mpm_get_state(mpm_ctx* c, float* x, float* v, float* Cm, float* F, float* Jp, int32_t* mat,
                      int32_t* emitter) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    if (x) download_planes(c, c->x, x, 3);
    if (v) download_planes(c, c->v, v, 3);
    if (Cm) download_planes(c, c->C, Cm, 9);
    if (F) download_planes(c, c->F, F, 9);
    if (Jp) download_planes(c, c->Jp, Jp, 1);
    if (mat) download_planes(c, reinterpret_cast<const float*>(c->mat), reinterpret_cast<float*>(mat), 1);
    if (emitter) download_planes(c, reinterpret_cast<const float*>(c->emitter), reinterpret_cast<float*>(emitter), 1);
  });
}

int32_t mpm_set_colliders(mpm_ctx* c, int32_t n, const int32_t*, const int32_t*, const float*, const int32_t*,
                          const int32_t*, const float*) {
  if (!c) return MPM_ERR_VALUE;
  if (n == 0) return MPM_OK;
  g_error = "colliders are not available in this build";
  return MPM_ERR_NOT_IMPLEMENTED;
}

int32_t mpm_set_collider_poses(mpm_ctx* c, int32_t n, const mpm_collider_pose*) {
  if (!c) return MPM_ERR_VALUE;
  if (n == 0) return MPM_OK;
  g_error = "expected 0 collider transforms";
  return MPM_ERR_VALUE;
}

int32_t mpm_advance(mpm_ctx* c, double frame_dt, mpm_advance_stats* out) {
  if (!c || !out) return MPM_ERR_VALUE;
  return run(c, [&] {
    if (c->n == 0) {
      *out = mpm_advance_stats{0, frame_dt, 0};
      return;
    }
    c->frame_dt = frame_dt;
    double vmax = (double)max_speed(c);
    if (!std::isfinite(vmax)) vmax = 0.0;  // already blown up: keep stepping so the NaN reaches x
    double wave = 0.0;
    for (size_t i = 0; i < c->mats.size(); ++i)
      if (c->present[i]) wave = std::max(wave, c->mats[i].wave_speed);
    const double cfl_dt = std::min(0.4 * c->dom.dx_f64 / (wave + vmax), frame_dt);
    const int64_t steps = (int64_t)std::ceil(frame_dt / cfl_dt - 1e-9);
    const double dt = frame_dt / (double)steps;
    for (int64_t k = 0; k < steps; ++k) substep(c, dt, (double)(k + 1) / (double)steps, frame_dt);
    const int64_t removed = remove_escaped(c);
    if (removed) refresh_present(c);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    *out = mpm_advance_stats{steps, dt, removed};
  });
}

int32_t mpm_substep(mpm_ctx* c, double dt, double s, double frame_dt) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    c->frame_dt = frame_dt;
    substep(c, dt, s, frame_dt);
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_is_occupied(mpm_ctx* c, int64_t m, const float* pts, uint8_t* out) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
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

int32_t mpm_debug_p2g(mpm_ctx* c, double dt) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    HIP_CHECK(hipMemsetAsync(c->accum, 0, (size_t)c->n_nodes * sizeof(float4), c->stream));
    run_p2g(c, (float)dt);
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_debug_grid_update(mpm_ctx* c, double dt, double s, double frame_dt) {
  if (!c) return MPM_ERR_VALUE;
  (void)s;
  return run(c, [&] {
    c->frame_dt = frame_dt;
    run_grid(c, (float)dt);
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_debug_g2p(mpm_ctx* c, double dt, double s, double frame_dt) {
  if (!c) return MPM_ERR_VALUE;
  (void)s;
  return run(c, [&] {
    c->frame_dt = frame_dt;
    run_g2p(c, (float)dt);
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_get_grid(mpm_ctx* c, int32_t which, float* grid_v, float* grid_m) {
  if (!c || !grid_v || !grid_m || (which != 0 && which != 1)) return MPM_ERR_VALUE;
  return run(c, [&] {
    const int64_t total = (int64_t)c->geom.dims[0] * c->geom.dims[1] * c->geom.dims[2];
    float* stage = static_cast<float*>(c->scratch.get(total * 4 * sizeof(float)));
    k_grid_to_ref<<<grid_for(total), kBlock, 0, c->stream>>>(c->geom, which == 0 ? c->accum : c->vel, stage,
                                                             stage + 3 * total);
    check_launch("k_grid_to_ref");
    HIP_CHECK(hipMemcpyAsync(grid_v, stage, total * 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipMemcpyAsync(grid_m, stage + 3 * total, total * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_set_grid(mpm_ctx* c, const float* grid_v, const float* grid_m) {
  if (!c || !grid_v || !grid_m) return MPM_ERR_VALUE;
  return run(c, [&] {
    const int64_t total = (int64_t)c->geom.dims[0] * c->geom.dims[1] * c->geom.dims[2];
    float* stage = static_cast<float*>(c->scratch.get(total * 4 * sizeof(float)));
    HIP_CHECK(hipMemcpyAsync(stage, grid_v, total * 3 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(stage + 3 * total, grid_m, total * sizeof(float), hipMemcpyHostToDevice, c->stream));
    k_grid_from_ref<<<grid_for(total), kBlock, 0, c->stream>>>(c->geom, stage, stage + 3 * total, c->vel);
    check_launch("k_grid_from_ref");
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int32_t mpm_get_binning(mpm_ctx* c, int32_t*, int32_t*, int32_t*, int64_t*) {
  if (!c) return MPM_ERR_VALUE;
  g_error = "binning is not available in this build";
  return MPM_ERR_NOT_IMPLEMENTED;
}

int32_t mpm_set_timing(mpm_ctx* c, int32_t enable) {
  if (!c) return MPM_ERR_VALUE;
  return run(c, [&] {
    collect_timing(c);
    c->timing.on = enable != 0;
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

}  // extern "C"
