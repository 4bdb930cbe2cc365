// Device kernels of one MLS-MPM substep on gfx950 (wave64).
//
// Data layout in HBM
//   particles  structure of arrays, each component a plane of `cap` floats: x[a * cap + p],
//              C[(3 i + j) * cap + p], so a wave reads every component with one coalesced 256-byte load.
//              Slot p holds the particle whose insertion-order index is pid[p].
//   grid       bricks of 4x4x4 nodes, 64 nodes (1 KiB of float4) contiguous per brick, bricks in
//              (bx, by, bz) row-major order. Two float4 arrays: `accum` receives P2G (momentum xyz,
//              mass) and is all-zero outside P2G; `vel` holds the updated velocity (xyz) and the mass (w)
//              that G2P and the collider pass read. The grid update turns accum into vel and writes the
//              zeros back, so no separate clearing pass is needed.
//
// Kernel semantics follow the reference Warp kernels (src/mpm_core/backends/warp/kernels.py and
// colliders.py); each kernel names the function it restates.
#pragma once

#include "device_math.hpp"

namespace mpm {

struct MatDev {
  int32_t model;
  float mu, lam, yield_stress, alpha, hardening, max_hardening, crit_c, crit_s, p_mass;
};

// Particle slots: x, v, F, Jp as planes of `cap` floats; per slot one 64-byte record `rec` (4 float4)
// holding what P2G reads besides x and v: C (row-major, 9), the Kirchhoff stress of F (xx, yy, zz, xy,
// xz, yz) and the particle mass. One record per lane makes P2G's loads whole cache-line pieces where
// planes would scatter them (its lanes walk different cells, not consecutive slots).
// Device arrays in HBM. The qualifier makes device code address them as global memory even when the
// pointer itself is read from memory (a scene table) rather than passed as a kernel argument, where the
// compiler could not infer it and would emit flat (global-or-LDS) accesses; host code sees plain pointers
// with the same representation.
#if defined(__HIP_DEVICE_COMPILE__)
#define MPM_GLOBAL __attribute__((address_space(1)))
#else
#define MPM_GLOBAL
#endif

// a plain pointer as a global one (a no-op cast in host code)
template <class T> __host__ __device__ inline MPM_GLOBAL T* as_global(T* p) { return (MPM_GLOBAL T*)p; }

struct Particles {
  MPM_GLOBAL float* x;
  MPM_GLOBAL float* v;
  MPM_GLOBAL float* F;
  MPM_GLOBAL float* Jp;
  MPM_GLOBAL int32_t* mat;
  MPM_GLOBAL int32_t* pid;
  MPM_GLOBAL float4* rec;
  int64_t cap;
};

struct GridGeom {
  int dims[3];    // nodes per axis (grid_shape + 6)
  int bdims[3];   // bricks per axis
  float origin[3];
  float dx, inv_dx;
};

struct StepConsts {
  float p_vol;
  float gravity[3];
  int wall_lo_node[3], wall_hi_node[3];
  int lo_on[3], hi_on[3];
  float wall_lo[3], wall_hi[3];
  float clamp_lo[3], clamp_hi[3];
  float wall_friction;
  int precision;  // kPrecisionFp32, kPrecisionFallback or kPrecisionFp64 (SVD and return maps)
};

// Where the 3x3 SVD, the return maps and the stress run in FP64: never, always (a reference-quality mode
// for error measurements), or as a fallback for the particles FP32 cannot resolve:
//   ill-conditioned F, sigma_min / sigma_max < 1e-2: F^T F squares the ratio, and below ~1e-3 the small
//     singular value sinks under the rounding of the large one;
//   collapsed volume, J = Jp det(F) < 0.25 (compressed snow; Jp is 1 for the other models), where the
//     hardened stiffness multiplies the rounding of the clamped singular values.
// Near-degenerate singular values alone do not qualify: U is then ill-determined, but the stress
// U diag(d) U^T is not, because the principal stresses d are as close as the singular values (measured:
// singular values equal to 1e-6 give the same FP32 error as FP64, 3e-7 on the grid momentum).
enum : int { kPrecisionFp32 = 0, kPrecisionFallback = 1, kPrecisionFp64 = 2 };

__device__ __forceinline__ bool needs_fp64(int precision, const float (&sig)[3], float jp) {
  if (precision == kPrecisionFp64) return true;
  if (precision != kPrecisionFallback) return false;
  const float a0 = fabsf(sig[0]), a1 = fabsf(sig[1]), a2 = fabsf(sig[2]);
  return a2 < 1.0e-2f * a0 || jp * a0 * a1 * a2 < 0.25f;
}

// Work list of particles whose material update is redone in FP64 after the FP32 pass. Appends are
// aggregated per wave (every lane of the wave must call, flag or not).
struct FixList {
  uint32_t* list;
  uint32_t* count;
};

__device__ __forceinline__ void wave_append(const FixList& fl, bool flag, uint32_t value) {
  const uint64_t mask = __ballot(flag);
  if (mask == 0) return;
  const int lane = threadIdx.x & 63;
  uint32_t base = 0;
  if (lane == __builtin_ctzll(mask)) base = atomicAdd(fl.count, (uint32_t)__popcll(mask));
  base = __shfl(base, __builtin_ctzll(mask), 64);
  if (flag) {
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    fl.list[base + (uint32_t)__popcll(mask & lt)] = value;
  }
}

__global__ void k_zero_u32(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__device__ __forceinline__ int node_addr(const GridGeom& g, int i, int j, int k) {
  return (((i >> 2) * g.bdims[1] + (j >> 2)) * g.bdims[2] + (k >> 2)) * 64 + ((i & 3) << 4) + ((j & 3) << 2) + (k & 3);
}

__device__ __forceinline__ void node_ijk(const GridGeom& g, int addr, int& i, int& j, int& k) {
  const int brick = addr >> 6, local = addr & 63;
  const int bz = brick % g.bdims[2];
  const int by = (brick / g.bdims[2]) % g.bdims[1];
  const int bx = brick / (g.bdims[2] * g.bdims[1]);
  i = (bx << 2) + (local >> 4);
  j = (by << 2) + ((local >> 2) & 3);
  k = (bz << 2) + (local & 3);
}

__device__ __forceinline__ void load_m3(const float* base, int64_t cap, int64_t p, M3<float>& m) {
#pragma unroll
  for (int e = 0; e < 9; ++e) m.a[e / 3][e % 3] = base[e * cap + p];
}

__device__ __forceinline__ void store_m3(float* base, int64_t cap, int64_t p, const M3<float>& m) {
#pragma unroll
  for (int e = 0; e < 9; ++e) base[e * cap + p] = m.a[e / 3][e % 3];
}

__device__ __forceinline__ int clamp_i(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Stencil base node floor(x / dx - 0.5) and fractional offset (kernels.py:25-27, 128-130). The base is
// clamped into the grid so a non-finite position never indexes outside it; such a particle's weights
// are non-finite and carry the NaN into the grid, where the runner's NaN check finds it.
__device__ __forceinline__ void stencil_base(const GridGeom& g, const float (&x)[3], int (&base)[3], float (&fx)[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float xp = (x[a] - g.origin[a]) * g.inv_dx;
    const float fb = floorf(xp - 0.5f);
    base[a] = clamp_i((int)fminf(fmaxf(fb, -1.0f), (float)g.dims[a]), 0, g.dims[a] - 3);
    fx[a] = xp - fb;
  }
}

__device__ __forceinline__ int brick_of_base(const GridGeom& g, const int (&base)[3]) {
  return ((base[0] >> 2) * g.bdims[1] + (base[1] >> 2)) * g.bdims[2] + (base[2] >> 2);
}

// brick-major cell key: brick * 64 + ((i & 3) * 16 + (j & 3) * 4 + (k & 3))
__device__ __forceinline__ uint32_t cell_key(const GridGeom& g, const int (&base)[3]) {
  return (uint32_t)brick_of_base(g, base) * 64u + (uint32_t)(((base[0] & 3) << 4) | ((base[1] & 3) << 2) | (base[2] & 3));
}

// Kirchhoff stress of one particle (kernels.py:115-124): Hencky for von Mises and Drucker-Prager,
// fixed corotated for snow (with hardening h = min(exp(xi (1 - Jp)), h_max)) and elastic.

// Kirchhoff stress from an SVD F = U diag(sig) V^T, as the symmetric U diag(d) U^T (6 entries
// xx yy zz xy xz yz). Fixed corotated: (F - R) F^T = U (diag(sig) - I) diag(sig) U^T, so
// d_i = 2 mu (sig_i - 1) sig_i + lam J (J - 1); Hencky: d_i = 2 mu eps_i + lam tr(eps).
// Kirchhoff stress from the SVD of F as 6 symmetric entries (xx, yy, zz, xy, xz, yz): U diag(d) U^T with
// d the Hencky (kernels.py:41-60) or fixed-corotated (kernels.py:30-38, snow hardening :120-122) principal
// stress; for fixed-corotated (F - R) F^T = U (S - I) S U^T.
template <class T>
__device__ __forceinline__ void stress_from_svd(const MatDev& m, const M3<T>& U, const T (&sig)[3], T Jp,
                                                T (&tau)[6]) {
  T d[3];
  if (m.model == kVonMises || m.model == kDruckerPrager) {
    const T e0 = safe_log(sig[0]), e1 = safe_log(sig[1]), e2 = safe_log(sig[2]);
    const T tr = e0 + e1 + e2;
    d[0] = T(2) * T(m.mu) * e0 + T(m.lam) * tr;
    d[1] = T(2) * T(m.mu) * e1 + T(m.lam) * tr;
    d[2] = T(2) * T(m.mu) * e2 + T(m.lam) * tr;
  } else {
    T mu = T(m.mu), lam = T(m.lam);
    if (m.model == kSnow) {
      const T h = min_t(exp_t(T(m.hardening) * (T(1) - Jp)), T(m.max_hardening));
      mu *= h;
      lam *= h;
    }
    const T J = sig[0] * sig[1] * sig[2];
    const T vol = lam * J * (J - T(1));
#pragma unroll
    for (int a = 0; a < 3; ++a) d[a] = T(2) * mu * (sig[a] - T(1)) * sig[a] + vol;
  }
  const auto& u = U.a;
  tau[0] = u[0][0] * d[0] * u[0][0] + u[0][1] * d[1] * u[0][1] + u[0][2] * d[2] * u[0][2];
  tau[1] = u[1][0] * d[0] * u[1][0] + u[1][1] * d[1] * u[1][1] + u[1][2] * d[2] * u[1][2];
  tau[2] = u[2][0] * d[0] * u[2][0] + u[2][1] * d[1] * u[2][1] + u[2][2] * d[2] * u[2][2];
  tau[3] = u[0][0] * d[0] * u[1][0] + u[0][1] * d[1] * u[1][1] + u[0][2] * d[2] * u[1][2];
  tau[4] = u[0][0] * d[0] * u[2][0] + u[0][1] * d[1] * u[2][1] + u[0][2] * d[2] * u[2][2];
  tau[5] = u[1][0] * d[0] * u[2][0] + u[1][1] * d[1] * u[2][1] + u[1][2] * d[2] * u[2][2];
}

// stress of a particle state in the requested precision (6 entries, FP32 out)
__device__ __forceinline__ void particle_stress(const MatDev& m, const M3<float>& F, float Jp, int precision,
                                                float (&t)[6]) {
  M3<float> U, V;
  float sig[3];
  svd3(F, U, sig, V);
  if (needs_fp64(precision, sig, m.model == kSnow ? Jp : 1.0f)) {
    M3<double> Fd, Ud, Vd;
#pragma unroll
    for (int e = 0; e < 9; ++e) Fd.a[e / 3][e % 3] = (double)F.a[e / 3][e % 3];
    double sd[3], td[6];
    svd3(Fd, Ud, sd, Vd);
    stress_from_svd(m, Ud, sd, (double)Jp, td);
#pragma unroll
    for (int e = 0; e < 6; ++e) t[e] = (float)td[e];
    return;
  }
  stress_from_svd(m, U, sig, Jp, t);
}

__device__ __forceinline__ void load_C(const Particles& P, int64_t p, M3<float>& C) {
  const float4 a = P.rec[4 * p], b = P.rec[4 * p + 1], c = P.rec[4 * p + 2];
  C.a[0][0] = a.x; C.a[0][1] = a.y; C.a[0][2] = a.z; C.a[1][0] = a.w;
  C.a[1][1] = b.x; C.a[1][2] = b.y; C.a[2][0] = b.z; C.a[2][1] = b.w;
  C.a[2][2] = c.x;
}
__device__ __forceinline__ void store_rec(const Particles& P, int64_t p, const M3<float>& C, const float (&t)[6],
                                          float pm) {
  P.rec[4 * p] = make_float4(C.a[0][0], C.a[0][1], C.a[0][2], C.a[1][0]);
  P.rec[4 * p + 1] = make_float4(C.a[1][1], C.a[1][2], C.a[2][0], C.a[2][1]);
  P.rec[4 * p + 2] = make_float4(C.a[2][2], t[0], t[1], t[2]);
  P.rec[4 * p + 3] = make_float4(t[3], t[4], t[5], pm);
}
// stress and mass of a record whose C stays
__device__ __forceinline__ void store_tau_pm(const Particles& P, int64_t p, const float (&t)[6], float pm) {
  const float c22 = P.rec[4 * p + 2].x;
  P.rec[4 * p + 2] = make_float4(c22, t[0], t[1], t[2]);
  P.rec[4 * p + 3] = make_float4(t[3], t[4], t[5], pm);
}

// stress of every particle from its F (and Jp): after load_state, when tau is not yet known
// stress and particle mass of every record from F, Jp and the material (after a state upload)
__global__ void __launch_bounds__(256) k_stress(Particles P, const MatDev* __restrict__ mats, int n, int precision) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const MatDev m = mats[P.mat[p]];
  M3<float> F;
  load_m3(P.F, P.cap, p, F);
  float t[6];
  particle_stress(m, F, P.Jp[p], precision, t);
  store_tau_pm(P, p, t, m.p_mass);
}

// ---------------------------------------------------------------------------------------------
// P2G, reference data flow: one thread per particle, 27 x 4 FP32 global atomics (kernels.py:92-140).
__global__ void __launch_bounds__(256) k_p2g_atomic(Particles P, const MatDev* __restrict__ mats, GridGeom g,
                                                    StepConsts sc, int n, float dt, float4* __restrict__ accum) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int64_t cap = P.cap;
  const MatDev m = mats[P.mat[p]];
  M3<float> F, Cm;
  load_m3(P.F, cap, p, F);
  load_C(P, p, Cm);
  float t6[6];
  particle_stress(m, F, P.Jp[p], sc.precision, t6);
  M3<float> tau;
  tau.a[0][0] = t6[0]; tau.a[1][1] = t6[1]; tau.a[2][2] = t6[2];
  tau.a[0][1] = tau.a[1][0] = t6[3];
  tau.a[0][2] = tau.a[2][0] = t6[4];
  tau.a[1][2] = tau.a[2][1] = t6[5];
  const float pm = m.p_mass;
  const float k = -dt * sc.p_vol * 4.0f * g.inv_dx * g.inv_dx;
  M3<float> aff;
#pragma unroll
  for (int e = 0; e < 9; ++e) aff.a[e / 3][e % 3] = k * tau.a[e / 3][e % 3] + pm * Cm.a[e / 3][e % 3];
  float fx[3];
  int base[3];
  const float xo[3] = {P.x[p], P.x[cap + p], P.x[2 * cap + p]};
  stencil_base(g, xo, base, fx);
  const float mv[3] = {pm * P.v[p], pm * P.v[cap + p], pm * P.v[2 * cap + p]};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float wi = bspline_w(fx[0], i);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float wj = bspline_w(fx[1], j);
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) {
        const float w = wi * wj * bspline_w(fx[2], kk);
        const float d0 = ((float)i - fx[0]) * g.dx, d1 = ((float)j - fx[1]) * g.dx, d2 = ((float)kk - fx[2]) * g.dx;
        float4* node = accum + node_addr(g, base[0] + i, base[1] + j, base[2] + kk);
        atomicAdd(&node->x, w * (mv[0] + (aff.a[0][0] * d0 + aff.a[0][1] * d1 + aff.a[0][2] * d2)));
        atomicAdd(&node->y, w * (mv[1] + (aff.a[1][0] * d0 + aff.a[1][1] * d1 + aff.a[1][2] * d2)));
        atomicAdd(&node->z, w * (mv[2] + (aff.a[2][0] * d0 + aff.a[2][1] * d1 + aff.a[2][2] * d2)));
        atomicAdd(&node->w, w * pm);
      }
    }
  }
}

// wall_contact (kernels.py:143-152): drop the into-wall component, Coulomb friction on the rest
__device__ __forceinline__ void wall_contact(float (&vel)[3], int axis, float friction) {
  const float vn = fabsf(vel[axis]);
  vel[axis] = 0.0f;
  const float len = sqrtf(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
  if (len > 0.0f) {
    const float f = fmaxf(0.0f, 1.0f - friction * vn / len);
    vel[0] *= f;
    vel[1] *= f;
    vel[2] *= f;
  }
}

// grid_update (kernels.py:155-180) over one node: momentum -> velocity, gravity, walls by node index.
__device__ __forceinline__ float4 update_node(const GridGeom& g, const StepConsts& sc, float dt, int i, int j, int k,
                                              float4 a) {
  const float m = a.w;
  if (!(m > 0.0f)) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float vel[3] = {a.x / m + dt * sc.gravity[0], a.y / m + dt * sc.gravity[1], a.z / m + dt * sc.gravity[2]};
  const int node[3] = {i, j, k};
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    if (sc.lo_on[ax] != 0 && node[ax] <= sc.wall_lo_node[ax] && vel[ax] < 0.0f) wall_contact(vel, ax, sc.wall_friction);
    if (sc.hi_on[ax] != 0 && node[ax] >= sc.wall_hi_node[ax] && vel[ax] > 0.0f) wall_contact(vel, ax, sc.wall_friction);
  }
  return make_float4(vel[0], vel[1], vel[2], m);
}

// ---------------------------------------------------------------------------------------------
// State transfer between host layouts (array of structs in insertion order) and device planes.

// dst[pid[p] * comps + c] = src[c * cap + p]
__global__ void k_pack(const float* __restrict__ src, int64_t cap, const int32_t* __restrict__ pid, int comps, int n,
                       float* __restrict__ dst) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int64_t q = pid ? pid[p] : p;
  for (int c = 0; c < comps; ++c) dst[q * comps + c] = src[c * cap + p];
}

// dst[c * cap + (p0 + p)] = src[p * comps + c]
__global__ void k_unpack(const float* __restrict__ src, int comps, int n, int64_t cap, int64_t p0,
                         float* __restrict__ dst) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int c = 0; c < comps; ++c) dst[c * cap + p0 + p] = src[(int64_t)p * comps + c];
}

// new particles: C = 0, F = I, Jp = 1, tau = 0 (the stress of F = I), pid = insertion index
// fresh particles (solver.py:100-125): C = 0, F = I, Jp = 1, zero stress (F = I is unstressed in every model)
__global__ void k_init_new(Particles P, const MatDev* __restrict__ mats, int64_t p0, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t p = p0 + i;
#pragma unroll
  for (int e = 0; e < 9; ++e) P.F[e * P.cap + p] = (e % 4 == 0) ? 1.0f : 0.0f;
  const float pm = mats[P.mat[p]].p_mass;
  P.rec[4 * p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  P.rec[4 * p + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  P.rec[4 * p + 2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  P.rec[4 * p + 3] = make_float4(0.0f, 0.0f, 0.0f, pm);
  P.Jp[p] = 1.0f;
  P.pid[p] = (int32_t)p;
}

// C of every slot into host rows (insertion order through pid) and back; uploads leave stress and mass
// to k_stress
__global__ void k_pack_C(Particles P, int n, float* __restrict__ dst) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  M3<float> C;
  load_C(P, p, C);
  const int64_t q = P.pid[p];
#pragma unroll
  for (int e = 0; e < 9; ++e) dst[q * 9 + e] = C.a[e / 3][e % 3];
}
__global__ void k_unpack_C(const float* __restrict__ src, int n, int64_t p0, Particles P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  M3<float> C;
#pragma unroll
  for (int e = 0; e < 9; ++e) C.a[e / 3][e % 3] = src[(int64_t)i * 9 + e];
  const float z[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  store_rec(P, p0 + i, C, z, 0.0f);
}

__global__ void k_iota(int32_t* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = i;
}

// bricked grid <-> reference C-order (D0, D1, D2) arrays
// live (optional): per-brick flag; nodes of bricks without a fresh value read as zero
__global__ void k_grid_to_ref(GridGeom g, const float4* __restrict__ src, const uint8_t* __restrict__ live,
                              float* __restrict__ v3, float* __restrict__ m) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = g.dims[0] * g.dims[1] * g.dims[2];
  if (idx >= total) return;
  const int k = idx % g.dims[2], j = (idx / g.dims[2]) % g.dims[1], i = idx / (g.dims[2] * g.dims[1]);
  const int addr = node_addr(g, i, j, k);
  float4 a = src[addr];
  if (live && !live[addr >> 6]) a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  v3[3 * (int64_t)idx] = a.x;
  v3[3 * (int64_t)idx + 1] = a.y;
  v3[3 * (int64_t)idx + 2] = a.z;
  m[idx] = a.w;
}

// fixed-point accumulator (4 int64 per node) to float4 (momentum, mass) for the grid download
__global__ void k_fixed_to_float(const unsigned long long* __restrict__ fx, int n_nodes, double inv_scale,
                                 float4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  const unsigned long long* f = fx + 4 * (int64_t)i;
  out[i] = make_float4((float)((double)(long long)f[0] * inv_scale), (float)((double)(long long)f[1] * inv_scale),
                       (float)((double)(long long)f[2] * inv_scale), (float)((double)(long long)f[3] * inv_scale));
}

__global__ void k_grid_from_ref(GridGeom g, const float* __restrict__ v3, const float* __restrict__ m,
                                float4* __restrict__ dst) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = g.dims[0] * g.dims[1] * g.dims[2];
  if (idx >= total) return;
  const int k = idx % g.dims[2], j = (idx / g.dims[2]) % g.dims[1], i = idx / (g.dims[2] * g.dims[1]);
  dst[node_addr(g, i, j, k)] = make_float4(v3[3 * (int64_t)idx], v3[3 * (int64_t)idx + 1], v3[3 * (int64_t)idx + 2], m[idx]);
}

// ---------------------------------------------------------------------------------------------
// is_occupied (kernels.py:272-298): splat particle volume with B-spline weights, then test the
// node nearest to each query point against half a cell volume.
__global__ void __launch_bounds__(256) k_scatter_volume(Particles P, GridGeom g, float p_vol, int n,
                                                        float* __restrict__ vol) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float fx[3];
  int base[3];
  const float xo[3] = {P.x[p], P.x[P.cap + p], P.x[2 * P.cap + p]};
  stencil_base(g, xo, base, fx);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        atomicAdd(vol + node_addr(g, base[0] + i, base[1] + j, base[2] + k),
                  bspline_w(fx[0], i) * bspline_w(fx[1], j) * bspline_w(fx[2], k) * p_vol);
}

__global__ void k_query_fill(GridGeom g, const float* __restrict__ pts, int m, float cell_vol,
                             const float* __restrict__ vol, uint8_t* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  int c[3];
  bool inside = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    c[a] = (int)roundf((pts[3 * q + a] - g.origin[a]) * g.inv_dx);
    inside = inside && c[a] >= 0 && c[a] < g.dims[a];
  }
  out[q] = (inside && vol[node_addr(g, c[0], c[1], c[2])] >= 0.5f * cell_vol) ? 1 : 0;
}

// ---------------------------------------------------------------------------------------------
// Removal of particles past open sides (solver.py:222-240), in three passes: flag, scan, compact.
__global__ void k_escape_flags(Particles P, int n, float3 lo, float3 hi, int32_t* __restrict__ keep_by_pid) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float x0 = P.x[p], x1 = P.x[P.cap + p], x2 = P.x[2 * P.cap + p];
  const bool out = x0 < lo.x || x1 < lo.y || x2 < lo.z || x0 > hi.x || x1 > hi.y || x2 > hi.z;
  keep_by_pid[P.pid[p]] = out ? 0 : 1;
}

// the survivors of an escape pass, stable in insertion order: the particle with insertion index pid is kept
// when keep_by_pid[pid] and goes to slot new_pid[pid] (the exclusive scan of the flags) of dst with that as
// its new insertion index; its emitter id moves the same way
__global__ void k_compact(Particles src, Particles dst, const int32_t* __restrict__ keep_by_pid,
                          const uint32_t* __restrict__ new_pid, int n, const int32_t* __restrict__ emitter_in,
                          int32_t* __restrict__ emitter_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int pid = src.pid[p];
  if (!keep_by_pid[pid]) return;
  const int64_t q = new_pid[pid];
  const int64_t cs = src.cap, cd = dst.cap;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    dst.x[a * cd + q] = src.x[a * cs + p];
    dst.v[a * cd + q] = src.v[a * cs + p];
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) dst.F[e * cd + q] = src.F[e * cs + p];
#pragma unroll
  for (int e = 0; e < 4; ++e) dst.rec[4 * q + e] = src.rec[4 * (int64_t)p + e];
  dst.Jp[q] = src.Jp[p];
  dst.mat[q] = src.mat[p];
  dst.pid[q] = (int32_t)q;
  emitter_out[q] = emitter_in[pid];
}

}  // namespace mpm
