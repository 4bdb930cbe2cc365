// Small-matrix math for the MLS-MPM kernels on gfx950: quadratic B-spline weights, a 3x3 SVD and the
// four material models. Everything is templated on the scalar so the FP64 fallback path runs the same
// code as the FP32 bulk path.
//
// Semantics follow the reference device functions:
//   bspline_w / stencil_base            src/mpm_core/backends/warp/kernels.py:16-27
//   fixed_corotated_tau                 kernels.py:30-38
//   safe_log3 / exp3 / hencky_tau       kernels.py:41-60
//   project_von_mises                   kernels.py:63-72
//   project_drucker_prager              kernels.py:75-89
//   snow clamp + Jp ratio               kernels.py:258-262
//
// The SVD (Warp's wp.svd3 in the reference) is written here from the published two-stage method
// (McAdams et al. 2011, "Computing the SVD of 3x3 matrices with minimal branching..."): a cyclic Jacobi
// eigen-decomposition of F^T F accumulated into V, column ordering of B = F V by decreasing norm, and
// a Givens QR of B that yields U. The Jacobi angles follow the exact formula (not the paper's
// approximate ones), on F^T F scaled to unit trace; in FP32 they are evaluated with the hardware
// reciprocal and square-root instructions. Convention: U and V are proper rotations, |sig_0| >= |sig_1| >= |sig_2|,
// sig_0, sig_1 >= 0 and sig_2 < 0 exactly when det F < 0.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

enum : int32_t { kElastic = 0, kVonMises = 1, kDruckerPrager = 2, kSnow = 3 };

template <class T> struct M3 { T a[3][3]; };

template <class T> __device__ __forceinline__ T sqrt_t(T v);
template <> __device__ __forceinline__ float sqrt_t<float>(float v) { return __builtin_sqrtf(v); }
template <> __device__ __forceinline__ double sqrt_t<double>(double v) { return __builtin_sqrt(v); }
// Logarithm and exponential of the material models. FP32: the hardware base-2 instructions (v_log_f32,
// v_exp_f32, about 1 ulp) scaled by ln 2; their arguments are clamped strains and stretches, far from
// denormals. FP64: the libm functions.
template <class T> __device__ __forceinline__ T log_t(T v);
template <> __device__ __forceinline__ float log_t<float>(float v) {
  return __builtin_amdgcn_logf(v) * 0.693147180559945309f;
}
template <> __device__ __forceinline__ double log_t<double>(double v) { return log(v); }
template <class T> __device__ __forceinline__ T exp_t(T v);
template <> __device__ __forceinline__ float exp_t<float>(float v) {
  return __builtin_amdgcn_exp2f(v * 1.44269504088896341f);
}
template <> __device__ __forceinline__ double exp_t<double>(double v) { return exp(v); }
// Reciprocal, square root and reciprocal square root for rotation angles, where any value within a few
// ulp serves (the rotation stays orthonormal because c = 1/sqrt(1 + t^2) and s = t c are formed from the
// same t): FP32 uses the hardware instructions (v_rcp_f32, v_sqrt_f32, v_rsq_f32); FP64 is exact.
template <class T> __device__ __forceinline__ T rcp_a(T v);
template <> __device__ __forceinline__ float rcp_a<float>(float v) { return __builtin_amdgcn_rcpf(v); }
template <> __device__ __forceinline__ double rcp_a<double>(double v) { return 1.0 / v; }
template <class T> __device__ __forceinline__ T sqrt_a(T v);
template <> __device__ __forceinline__ float sqrt_a<float>(float v) { return __builtin_amdgcn_sqrtf(v); }
template <> __device__ __forceinline__ double sqrt_a<double>(double v) { return __builtin_sqrt(v); }
template <class T> __device__ __forceinline__ T rsqrt_a(T v);
template <> __device__ __forceinline__ float rsqrt_a<float>(float v) { return __builtin_amdgcn_rsqf(v); }
template <> __device__ __forceinline__ double rsqrt_a<double>(double v) { return 1.0 / __builtin_sqrt(v); }
// Quotients of the return maps: in FP32 a times the hardware reciprocal of b (about 1.5 ulp; the IEEE
// division sequence costs ten instructions), in FP64 the exact quotient.
template <class T> __device__ __forceinline__ T quot(T a, T b);
template <> __device__ __forceinline__ float quot<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <> __device__ __forceinline__ double quot<double>(double a, double b) { return a / b; }
template <class T> __device__ __forceinline__ T abs_t(T v) { return v < T(0) ? -v : v; }
template <class T> __device__ __forceinline__ T max_t(T a, T b) { return a > b ? a : b; }
template <class T> __device__ __forceinline__ T min_t(T a, T b) { return a < b ? a : b; }

template <class T> __device__ __forceinline__ M3<T> mul(const M3<T>& A, const M3<T>& B) {
  M3<T> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.a[i][j] = A.a[i][0] * B.a[0][j] + A.a[i][1] * B.a[1][j] + A.a[i][2] * B.a[2][j];
  return r;
}

// A * B^T
template <class T> __device__ __forceinline__ M3<T> mul_bt(const M3<T>& A, const M3<T>& B) {
  M3<T> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.a[i][j] = A.a[i][0] * B.a[j][0] + A.a[i][1] * B.a[j][1] + A.a[i][2] * B.a[j][2];
  return r;
}

// quadratic B-spline weight of stencil node i (0, 1, 2) at fractional offset fx in [0.5, 1.5)
template <class T> __device__ __forceinline__ T bspline_w(T fx, int i) {
  if (i == 0) return T(0.5) * (T(1.5) - fx) * (T(1.5) - fx);
  if (i == 1) return T(0.75) - (fx - T(1)) * (fx - T(1));
  return T(0.5) * (fx - T(0.5)) * (fx - T(0.5));
}

// Symmetric Jacobi rotation that zeroes S[p][q]; S is updated in place (symmetric entries only) and the
// rotation is accumulated into the columns p, q of V. S is scaled to unit trace, so an off-diagonal entry
// below 1e-30 is zero for every purpose (and keeps the reciprocal's argument normal); then t = 0 and the
// rotation is the identity. Branch-free: every lane of a wave runs the same instructions.
template <class T, int p, int q>
__device__ __forceinline__ void jacobi_rotate(T (&S)[3][3], M3<T>& V) {
  constexpr int r = 3 - p - q;
  const T apq = S[p][q];
  const bool skip = abs_t(apq) < T(1.0e-30);
  const T tau = (S[q][q] - S[p][p]) * rcp_a(T(2) * (skip ? T(1) : apq));
  // tau^2 may overflow to +inf: then t = 1 / inf = 0, the rotation angle of an already diagonal pair
  T t = rcp_a(abs_t(tau) + sqrt_a(T(1) + tau * tau));
  t = skip ? T(0) : (tau >= T(0) ? t : -t);
  const T c = rsqrt_a(T(1) + t * t);
  const T s = t * c;
  S[p][p] -= t * apq;
  S[q][q] += t * apq;
  S[p][q] = S[q][p] = T(0);
  const T srp = S[r][p], srq = S[r][q];
  S[r][p] = S[p][r] = c * srp - s * srq;
  S[r][q] = S[q][r] = s * srp + c * srq;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const T vkp = V.a[k][p], vkq = V.a[k][q];
    V.a[k][p] = c * vkp - s * vkq;
    V.a[k][q] = s * vkp + c * vkq;
  }
}

template <class T> struct SweepCount { static constexpr int value = 5; };
template <> struct SweepCount<double> { static constexpr int value = 10; };

template <class T>
__device__ __forceinline__ void svd3(const M3<T>& F, M3<T>& U, T (&sig)[3], M3<T>& V) {
  T S[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) S[i][j] = F.a[0][i] * F.a[0][j] + F.a[1][i] * F.a[1][j] + F.a[2][i] * F.a[2][j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) V.a[i][j] = i == j ? T(1) : T(0);
  // unit trace (an all-zero or denormal F^T F stays zero: no rotations)
  const T trace = S[0][0] + S[1][1] + S[2][2];
  const T inv_trace = trace > T(1.0e-30) ? rcp_a(trace) : T(0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) S[i][j] *= inv_trace;
  const T tol = sizeof(T) == 4 ? T(1e-8) : T(1e-17);
  for (int sw = 0; sw < SweepCount<T>::value; ++sw) {
    jacobi_rotate<T, 0, 1>(S, V);
    jacobi_rotate<T, 0, 2>(S, V);
    jacobi_rotate<T, 1, 2>(S, V);
    if (abs_t(S[0][1]) + abs_t(S[0][2]) + abs_t(S[1][2]) <= tol) break;
  }
  M3<T> B = mul(F, V);
  T n[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) n[j] = B.a[0][j] * B.a[0][j] + B.a[1][j] * B.a[1][j] + B.a[2][j] * B.a[2][j];
  // order columns by decreasing norm; each swap negates one column so det V stays +1
#define MPM_SWAP_COLS(i, j)                                 \
  if (n[i] < n[j]) {                                        \
    _Pragma("unroll") for (int k = 0; k < 3; ++k) {         \
      const T b = B.a[k][i]; B.a[k][i] = B.a[k][j]; B.a[k][j] = -b; \
      const T v = V.a[k][i]; V.a[k][i] = V.a[k][j]; V.a[k][j] = -v; \
    }                                                       \
    const T t = n[i]; n[i] = n[j]; n[j] = t;                \
  }
  MPM_SWAP_COLS(0, 1)
  MPM_SWAP_COLS(0, 2)
  MPM_SWAP_COLS(1, 2)
#undef MPM_SWAP_COLS
  // Givens QR of B: zero (1,0), (2,0), (2,1); U collects the transposed rotations
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) U.a[i][j] = i == j ? T(1) : T(0);
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    const int p = g == 2 ? 1 : 0;
    const int q = g == 0 ? 1 : 2;
    const T a = B.a[p][p], b = B.a[q][p];
    const T n2 = a * a + b * b;
    const bool rot = n2 > T(1.0e-30);
    const T inv = rsqrt_a(rot ? n2 : T(1));
    const T c = rot ? a * inv : T(1);
    const T s = rot ? b * inv : T(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const T bp = B.a[p][k], bq = B.a[q][k];
      B.a[p][k] = c * bp + s * bq;
      B.a[q][k] = -s * bp + c * bq;
      const T up = U.a[k][p], uq = U.a[k][q];
      U.a[k][p] = c * up + s * uq;
      U.a[k][q] = -s * up + c * uq;
    }
  }
  sig[0] = B.a[0][0];
  sig[1] = B.a[1][1];
  sig[2] = B.a[2][2];
}

// U diag(d) V^T
template <class T>
__device__ __forceinline__ M3<T> usv(const M3<T>& U, const T (&d)[3], const M3<T>& V) {
  M3<T> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.a[i][j] = U.a[i][0] * d[0] * V.a[j][0] + U.a[i][1] * d[1] * V.a[j][1] + U.a[i][2] * d[2] * V.a[j][2];
  return r;
}

template <class T> __device__ __forceinline__ T safe_log(T s) { return log_t(max_t(s, T(1.0e-4))); }

// tau = 2 mu (F - R) F^T + lam J (J - 1) I with R = U V^T, J = prod(sig)
template <class T>
__device__ __forceinline__ M3<T> fixed_corotated_tau(const M3<T>& F, const M3<T>& U, const T (&sig)[3],
                                                     const M3<T>& V, T mu, T lam) {
  const M3<T> R = mul_bt(U, V);
  const T J = sig[0] * sig[1] * sig[2];
  M3<T> D;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) D.a[i][j] = F.a[i][j] - R.a[i][j];
  M3<T> tau = mul_bt(D, F);
  const T two_mu = T(2) * mu;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) tau.a[i][j] *= two_mu;
  const T vol = lam * J * (J - T(1));
  tau.a[0][0] += vol;
  tau.a[1][1] += vol;
  tau.a[2][2] += vol;
  return tau;
}

// tau = U diag(2 mu eps + lam tr(eps)) U^T, eps = log(max(sig, 1e-4))
template <class T>
__device__ __forceinline__ M3<T> hencky_tau(const M3<T>& U, const T (&sig)[3], T mu, T lam) {
  const T e0 = safe_log(sig[0]), e1 = safe_log(sig[1]), e2 = safe_log(sig[2]);
  const T tr = e0 + e1 + e2;
  const T d[3] = {T(2) * mu * e0 + lam * tr, T(2) * mu * e1 + lam * tr, T(2) * mu * e2 + lam * tr};
  return usv(U, d, U);
}

template <class T>
__device__ __forceinline__ void project_von_mises(const T (&sig)[3], T mu, T yield_stress, T (&out)[3]) {
  const T e[3] = {safe_log(sig[0]), safe_log(sig[1]), safe_log(sig[2])};
  const T tr3 = (e[0] + e[1] + e[2]) * T(1.0 / 3.0);
  const T d[3] = {e[0] - tr3, e[1] - tr3, e[2] - tr3};
  const T n = sqrt_t(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const T limit = quot(yield_stress, T(2) * mu);
  if (n > limit) {
    const T k = quot(limit, n);
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = exp_t(tr3 + d[i] * k);
  } else {
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = sig[i];
  }
}

// Klar et al. 2016 return mapping in log-strain space
template <class T>
__device__ __forceinline__ void project_drucker_prager(const T (&sig)[3], T mu, T lam, T alpha, T (&out)[3]) {
  const T e[3] = {safe_log(sig[0]), safe_log(sig[1]), safe_log(sig[2])};
  const T tr = e[0] + e[1] + e[2];
#pragma unroll
  for (int i = 0; i < 3; ++i) out[i] = sig[i];
  if (tr >= T(0)) {
    out[0] = out[1] = out[2] = T(1);  // separation: stress free
    return;
  }
  const T tr3 = tr * T(1.0 / 3.0);
  const T d[3] = {e[0] - tr3, e[1] - tr3, e[2] - tr3};
  const T n = sqrt_t(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (n < T(1.0e-10)) return;
  const T dgamma = n + quot(T(3) * lam + T(2) * mu, T(2) * mu) * tr * alpha;
  if (dgamma <= T(0)) return;
  const T k = quot(dgamma, n);
#pragma unroll
  for (int i = 0; i < 3; ++i) out[i] = exp_t(e[i] - d[i] * k);
}

}  // namespace mpm
