// cpu_ref: CPU restatement of the reference MLS-MPM substep. TEST INFRASTRUCTURE ONLY.
//
// This file is the oracle: only the test-suite, the smoke check and the CPU-baseline leg of the
// benchmark load it, and only to check or time against. The product (the HIP library) never calls it.
//
// Every kernel is restated from the reference's Warp source, templated on the scalar type so the same
// code gives the FP32 oracle (same arithmetic as the device) and the FP64 oracle (error reference):
//   bspline_w / stencil_base        src/mpm_core/backends/warp/kernels.py:16-27
//   fixed_corotated_tau             kernels.py:30-38
//   safe_log3 / exp3 / hencky_tau   kernels.py:41-60
//   project_von_mises               kernels.py:63-72
//   project_drucker_prager          kernels.py:75-89
//   p2g                             kernels.py:92-140
//   wall_contact / grid_update      kernels.py:143-180
//   g2p                             kernels.py:183-263
//   max_speed                       kernels.py:266-269
//   scatter_volume / query_fill     kernels.py:272-298
//   collider_matrix / collider_sdf / collider_velocity / grid_colliders / project_particles
//                                   src/mpm_core/backends/warp/colliders.py:30-149
// Warp's BVH point query (wp.mesh_query_point_sign_normal) is replaced by a brute-force closest point
// over all triangles (Ericson, Real-Time Collision Detection, 5.1.5) with the sign taken from the
// angle-weighted pseudonormal of the closest feature (Baerentzen & Aanaes 2005). Warp's svd3 is
// replaced by a Jacobi eigen-decomposition of F^T F followed by a Givens QR of F V, which gives the
// same sign convention (U, V proper rotations, sigma sorted by magnitude, sigma_3 < 0 iff det F < 0).
//
// Array layouts are the reference's numpy layouts: x/v (N,3), C/F (N,3,3) row major, Jp (N,),
// mat (N,) int32, grid_v (D0,D1,D2,3), grid_m (D0,D1,D2), C-order.
//
// Build: g++ -O3 -fopenmp -shared -fPIC (oracle/build.sh).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <vector>
#include <limits>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum { MODEL_ELASTIC = 0, MODEL_VON_MISES = 1, MODEL_DRUCKER_PRAGER = 2, MODEL_SNOW = 3 };

template <class T> struct M3 { T a[3][3]; };

template <class T> inline M3<T> mzero() { M3<T> m; std::memset(&m, 0, sizeof m); return m; }
template <class T> inline M3<T> meye() { M3<T> m = mzero<T>(); m.a[0][0] = m.a[1][1] = m.a[2][2] = T(1); return m; }
template <class T> inline M3<T> mmul(const M3<T>& A, const M3<T>& B) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = A.a[i][0] * B.a[0][j] + A.a[i][1] * B.a[1][j] + A.a[i][2] * B.a[2][j];
  return r;
}
template <class T> inline M3<T> mtrans(const M3<T>& A) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = A.a[j][i];
  return r;
}
template <class T> inline M3<T> load_m(const float* p) {
  M3<T> m;
  for (int i = 0; i < 9; ++i) m.a[i / 3][i % 3] = T(p[i]);
  return m;
}
template <class T> inline void store_m(float* p, const M3<T>& m) {
  for (int i = 0; i < 9; ++i) p[i] = float(m.a[i / 3][i % 3]);
}

// ---------------------------------------------------------------------------------------------
// 3x3 SVD: F = U diag(sig) V^T.
// Eigen-decomposition of S = F^T F by cyclic Jacobi rotations (exact rotation angles, accumulated in
// V), columns of B = F V ordered by decreasing norm, then a Givens QR of B (B = U R) so U is a proper
// rotation, R_00, R_11 >= 0 and sig = diag(R). R's off-diagonal entries vanish up to rounding because
// the columns of B are orthogonal.
template <class T>
void svd3(const M3<T>& F, M3<T>& U, T sig[3], M3<T>& V, int sweeps) {
  M3<T> S = mmul(mtrans(F), F);
  V = meye<T>();
  for (int sw = 0; sw < sweeps; ++sw) {
    for (int pair = 0; pair < 3; ++pair) {
      const int p = pair == 2 ? 1 : 0;
      const int q = pair == 0 ? 1 : 2;
      const T apq = S.a[p][q];
      if (apq == T(0)) continue;
      const T tau = (S.a[q][q] - S.a[p][p]) / (T(2) * apq);
      const T t = (tau >= T(0) ? T(1) : T(-1)) / (std::fabs(tau) + std::sqrt(T(1) + tau * tau));
      const T c = T(1) / std::sqrt(T(1) + t * t);
      const T s = t * c;
      // S <- G^T S G with G the rotation in the (p,q) plane
      for (int k = 0; k < 3; ++k) {
        const T skp = S.a[k][p], skq = S.a[k][q];
        S.a[k][p] = c * skp - s * skq;
        S.a[k][q] = s * skp + c * skq;
      }
      for (int k = 0; k < 3; ++k) {
        const T spk = S.a[p][k], sqk = S.a[q][k];
        S.a[p][k] = c * spk - s * sqk;
        S.a[q][k] = s * spk + c * sqk;
      }
      for (int k = 0; k < 3; ++k) {
        const T vkp = V.a[k][p], vkq = V.a[k][q];
        V.a[k][p] = c * vkp - s * vkq;
        V.a[k][q] = s * vkp + c * vkq;
      }
    }
  }
  M3<T> B = mmul(F, V);
  T n[3];
  for (int j = 0; j < 3; ++j) n[j] = B.a[0][j] * B.a[0][j] + B.a[1][j] * B.a[1][j] + B.a[2][j] * B.a[2][j];
  // order columns by decreasing norm; a swap negates one column so det V stays +1
  auto swap_cols = [&](int i, int j) {
    for (int k = 0; k < 3; ++k) {
      std::swap(B.a[k][i], B.a[k][j]);
      B.a[k][j] = -B.a[k][j];
      std::swap(V.a[k][i], V.a[k][j]);
      V.a[k][j] = -V.a[k][j];
    }
    std::swap(n[i], n[j]);
  };
  if (n[0] < n[1]) swap_cols(0, 1);
  if (n[0] < n[2]) swap_cols(0, 2);
  if (n[1] < n[2]) swap_cols(1, 2);
  // Givens QR: zero (1,0), (2,0), (2,1); U accumulates the transposed rotations
  U = meye<T>();
  const int rows[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int g = 0; g < 3; ++g) {
    const int p = rows[g][0], q = rows[g][1], col = p;
    const T a = B.a[p][col], b = B.a[q][col];
    const T r = std::sqrt(a * a + b * b);
    T c = T(1), s = T(0);
    if (r > T(0)) { c = a / r; s = b / r; }
    for (int k = 0; k < 3; ++k) {
      const T bp = B.a[p][k], bq = B.a[q][k];
      B.a[p][k] = c * bp + s * bq;
      B.a[q][k] = -s * bp + c * bq;
    }
    for (int k = 0; k < 3; ++k) {
      const T up = U.a[k][p], uq = U.a[k][q];
      U.a[k][p] = c * up + s * uq;
      U.a[k][q] = -s * up + c * uq;
    }
  }
  sig[0] = B.a[0][0];
  sig[1] = B.a[1][1];
  sig[2] = B.a[2][2];
}

template <class T> constexpr int svd_sweeps() { return sizeof(T) == 4 ? 6 : 12; }

template <class T> M3<T> fixed_corotated_tau(const M3<T>& F, T mu, T lam) {
  M3<T> U, V;
  T s[3];
  svd3(F, U, s, V, svd_sweeps<T>());
  const M3<T> R = mmul(U, mtrans(V));
  const T J = s[0] * s[1] * s[2];
  M3<T> FmR;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) FmR.a[i][j] = F.a[i][j] - R.a[i][j];
  M3<T> tau = mmul(FmR, mtrans(F));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) tau.a[i][j] *= T(2) * mu;
  const T vol = lam * J * (J - T(1));
  for (int i = 0; i < 3; ++i) tau.a[i][i] += vol;
  return tau;
}

template <class T> inline T safe_log(T s) { return std::log(std::max(s, T(1.0e-4))); }

template <class T> M3<T> hencky_tau(const M3<T>& F, T mu, T lam) {
  M3<T> U, V;
  T s[3];
  svd3(F, U, s, V, svd_sweeps<T>());
  const T e[3] = {safe_log(s[0]), safe_log(s[1]), safe_log(s[2])};
  const T tr = e[0] + e[1] + e[2];
  T d[3];
  for (int i = 0; i < 3; ++i) d[i] = T(2) * mu * e[i] + lam * tr;
  M3<T> tau;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      tau.a[i][j] = U.a[i][0] * d[0] * U.a[j][0] + U.a[i][1] * d[1] * U.a[j][1] + U.a[i][2] * d[2] * U.a[j][2];
  return tau;
}

template <class T> void project_von_mises(const T sig[3], T mu, T yield_stress, T out[3]) {
  const T e[3] = {safe_log(sig[0]), safe_log(sig[1]), safe_log(sig[2])};
  const T tr = e[0] + e[1] + e[2];
  const T d[3] = {e[0] - tr / T(3), e[1] - tr / T(3), e[2] - tr / T(3)};
  const T n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const T limit = yield_stress / (T(2) * mu);
  if (n > limit) {
    for (int i = 0; i < 3; ++i) out[i] = std::exp(tr / T(3) + d[i] * (limit / n));
  } else {
    for (int i = 0; i < 3; ++i) out[i] = sig[i];
  }
}

template <class T> void project_drucker_prager(const T sig[3], T mu, T lam, T alpha, T out[3]) {
  const T e[3] = {safe_log(sig[0]), safe_log(sig[1]), safe_log(sig[2])};
  const T tr = e[0] + e[1] + e[2];
  if (tr >= T(0)) {
    out[0] = out[1] = out[2] = T(1);
    return;
  }
  const T d[3] = {e[0] - tr / T(3), e[1] - tr / T(3), e[2] - tr / T(3)};
  const T n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (n < T(1.0e-10)) {
    for (int i = 0; i < 3; ++i) out[i] = sig[i];
    return;
  }
  const T dgamma = n + (T(3) * lam + T(2) * mu) / (T(2) * mu) * tr * alpha;
  if (dgamma <= T(0)) {
    for (int i = 0; i < 3; ++i) out[i] = sig[i];
    return;
  }
  for (int i = 0; i < 3; ++i) out[i] = std::exp(e[i] - d[i] * (dgamma / n));
}

template <class T> inline T bspline_w(T fx, int i) {
  if (i == 0) return T(0.5) * (T(1.5) - fx) * (T(1.5) - fx);
  if (i == 1) return T(0.75) - (fx - T(1)) * (fx - T(1));
  return T(0.5) * (fx - T(0.5)) * (fx - T(0.5));
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Parameters, laid out as plain C structs (filled by oracle.py from the reference's host formulas).
extern "C" {

struct OracleDomain {
  int32_t dims[3];
  float origin[3];
  float dx, inv_dx, p_vol;
  float gravity[3];
  int32_t wall_lo_node[3], wall_hi_node[3];
  int32_t lo_on[3], hi_on[3];
  float wall_lo[3], wall_hi[3];
  float clamp_lo[3], clamp_hi[3];
  float wall_friction;
  float cell_vol;  // dx^3 rounded once from double (solver.py:254)
};

struct OracleMaterial {
  int32_t model;
  float mu, lam, yield_stress, alpha, hardening, max_hardening, crit_c, crit_s, p_mass;
};

// one animated collider: mesh in local space, pose endpoints (colliders.py:20-28), flags
struct OracleCollider {
  const float* verts;  // (V,3)
  const int32_t* tris;  // (T,3)
  int32_t n_verts, n_tris;
  int32_t closed;
  float friction;
  float aabb_lo[3], aabb_hi[3];
  float margin;
  float t0[3], t1[3];
  float q0[4], q1[4];  // x, y, z, w
  float S0[9], S1[9];
};

}  // extern "C"

namespace {

template <class T> struct Vec { T v[3]; };

// ---------------------------------------------------------------------------------------------
// Colliders

// rotation matrix of a unit quaternion (x, y, z, w)
template <class T> M3<T> quat_to_matrix(const T q[4]) {
  const T x = q[0], y = q[1], z = q[2], w = q[3];
  M3<T> m;
  m.a[0][0] = T(1) - T(2) * (y * y + z * z); m.a[0][1] = T(2) * (x * y - z * w); m.a[0][2] = T(2) * (x * z + y * w);
  m.a[1][0] = T(2) * (x * y + z * w); m.a[1][1] = T(1) - T(2) * (x * x + z * z); m.a[1][2] = T(2) * (y * z - x * w);
  m.a[2][0] = T(2) * (x * z - y * w); m.a[2][1] = T(2) * (y * z + x * w); m.a[2][2] = T(1) - T(2) * (x * x + y * y);
  return m;
}

// spherical linear interpolation on the unit quaternion sphere (q1 already on q0's hemisphere)
template <class T> void quat_slerp(const T q0[4], const T q1[4], T s, T out[4]) {
  T d = q0[0] * q1[0] + q0[1] * q1[1] + q0[2] * q1[2] + q0[3] * q1[3];
  d = std::min(std::max(d, T(-1)), T(1));
  const T ang = std::acos(d);
  T a, b;
  if (ang < T(1e-6)) {
    a = T(1) - s;
    b = s;
  } else {
    const T si = std::sin(ang);
    a = std::sin((T(1) - s) * ang) / si;
    b = std::sin(s * ang) / si;
  }
  T n = 0;
  for (int i = 0; i < 4; ++i) {
    out[i] = a * q0[i] + b * q1[i];
    n += out[i] * out[i];
  }
  n = std::sqrt(n);
  for (int i = 0; i < 4; ++i) out[i] /= n;
}

// affine pose M(s) = [R(s) S(s) | t(s)] (colliders.py:30-37)
template <class T> struct Affine { M3<T> A; T t[3]; };

template <class T> Affine<T> collider_matrix(const OracleCollider& c, T s) {
  T q0[4], q1[4], q[4];
  for (int i = 0; i < 4; ++i) { q0[i] = T(c.q0[i]); q1[i] = T(c.q1[i]); }
  quat_slerp(q0, q1, s, q);
  M3<T> S;
  for (int i = 0; i < 9; ++i) S.a[i / 3][i % 3] = T(c.S0[i]) + s * (T(c.S1[i]) - T(c.S0[i]));
  Affine<T> M;
  M.A = mmul(quat_to_matrix(q), S);
  for (int i = 0; i < 3; ++i) M.t[i] = T(c.t0[i]) + s * (T(c.t1[i]) - T(c.t0[i]));
  return M;
}

template <class T> M3<T> inverse3(const M3<T>& A) {
  const auto& a = A.a;
  const T c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
  const T c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
  const T c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
  const T det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
  const T id = T(1) / det;
  M3<T> r;
  r.a[0][0] = c00 * id; r.a[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id; r.a[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
  r.a[1][0] = c01 * id; r.a[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id; r.a[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
  r.a[2][0] = c02 * id; r.a[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id; r.a[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
  return r;
}

template <class T> inline void mat_vec(const M3<T>& A, const T v[3], T out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = A.a[i][0] * v[0] + A.a[i][1] * v[1] + A.a[i][2] * v[2];
}

// closest point on triangle abc to p (Ericson 5.1.5). feature: 0 face, 1..3 vertex a/b/c, 4..6 edge ab/bc/ca
template <class T> void closest_on_triangle(const T p[3], const T a[3], const T b[3], const T c[3], T out[3],
                                            int& feature) {
  T ab[3], ac[3], ap[3];
  for (int i = 0; i < 3; ++i) { ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; ap[i] = p[i] - a[i]; }
  auto dot = [](const T* u, const T* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
  const T d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= T(0) && d2 <= T(0)) { for (int i = 0; i < 3; ++i) out[i] = a[i]; feature = 1; return; }
  T bp[3];
  for (int i = 0; i < 3; ++i) bp[i] = p[i] - b[i];
  const T d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= T(0) && d4 <= d3) { for (int i = 0; i < 3; ++i) out[i] = b[i]; feature = 2; return; }
  const T vc = d1 * d4 - d3 * d2;
  if (vc <= T(0) && d1 >= T(0) && d3 <= T(0)) {
    const T v = d1 / (d1 - d3);
    for (int i = 0; i < 3; ++i) out[i] = a[i] + v * ab[i];
    feature = 4; return;
  }
  T cp[3];
  for (int i = 0; i < 3; ++i) cp[i] = p[i] - c[i];
  const T d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= T(0) && d5 <= d6) { for (int i = 0; i < 3; ++i) out[i] = c[i]; feature = 3; return; }
  const T vb = d5 * d2 - d1 * d6;
  if (vb <= T(0) && d2 >= T(0) && d6 <= T(0)) {
    const T w = d2 / (d2 - d6);
    for (int i = 0; i < 3; ++i) out[i] = a[i] + w * ac[i];
    feature = 6; return;
  }
  const T va = d3 * d6 - d5 * d4;
  if (va <= T(0) && (d4 - d3) >= T(0) && (d5 - d6) >= T(0)) {
    const T w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int i = 0; i < 3; ++i) out[i] = b[i] + w * (c[i] - b[i]);
    feature = 5; return;
  }
  const T denom = T(1) / (va + vb + vc);
  const T v = vb * denom, w = vc * denom;
  for (int i = 0; i < 3; ++i) out[i] = a[i] + ab[i] * v + ac[i] * w;
  feature = 0;
}

// Per-collider pseudonormals: face normals, angle-weighted vertex normals, edge normals (sum of the
// two adjacent face normals; one face on a boundary edge).
struct MeshNormals {
  std::vector<double> face, vert, edge;  // face (T,3), vert (V,3), edge (T,3 edges,3)
};

MeshNormals mesh_normals(const OracleCollider& c) {
  MeshNormals mn;
  const int nt = c.n_tris, nv = c.n_verts;
  mn.face.assign(3 * nt, 0.0);
  mn.vert.assign(3 * nv, 0.0);
  mn.edge.assign(9 * nt, 0.0);
  std::vector<std::pair<uint64_t, int>> edges;
  for (int t = 0; t < nt; ++t) {
    const int* tri = c.tris + 3 * t;
    double P[3][3];
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 3; ++i) P[k][i] = c.verts[3 * tri[k] + i];
    double e1[3], e2[3], n[3];
    for (int i = 0; i < 3; ++i) { e1[i] = P[1][i] - P[0][i]; e2[i] = P[2][i] - P[0][i]; }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (len > 0)
      for (int i = 0; i < 3; ++i) n[i] /= len;
    for (int i = 0; i < 3; ++i) mn.face[3 * t + i] = n[i];
    for (int k = 0; k < 3; ++k) {
      const double* p0 = P[k];
      const double* p1 = P[(k + 1) % 3];
      const double* p2 = P[(k + 2) % 3];
      double u[3], w[3];
      for (int i = 0; i < 3; ++i) { u[i] = p1[i] - p0[i]; w[i] = p2[i] - p0[i]; }
      const double lu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      const double lw = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
      double ang = 0;
      if (lu > 0 && lw > 0) {
        double cs = (u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) / (lu * lw);
        cs = std::min(1.0, std::max(-1.0, cs));
        ang = std::acos(cs);
      }
      for (int i = 0; i < 3; ++i) mn.vert[3 * tri[k] + i] += ang * n[i];
      const int va = tri[k], vb = tri[(k + 1) % 3];
      const uint64_t key = (uint64_t(std::min(va, vb)) << 32) | uint64_t(std::max(va, vb));
      edges.push_back({key, 3 * t + k});
    }
  }
  std::sort(edges.begin(), edges.end());
  for (size_t i = 0; i < edges.size();) {
    size_t j = i;
    double s[3] = {0, 0, 0};
    while (j < edges.size() && edges[j].first == edges[i].first) {
      const int t = edges[j].second / 3;
      for (int d = 0; d < 3; ++d) s[d] += mn.face[3 * t + d];
      ++j;
    }
    for (size_t k = i; k < j; ++k)
      for (int d = 0; d < 3; ++d) mn.edge[3 * edges[k].second + d] = s[d];
    i = j;
  }
  return mn;
}

// (outward normal, signed distance); distance 1e30 when outside the early-out box (colliders.py:40-65)
template <class T> void collider_sdf(const OracleCollider& c, const MeshNormals& mn, const Affine<T>& M,
                                     const T p[3], T thickness, T out_n[3], T& out_phi) {
  out_n[0] = 0; out_n[1] = 0; out_n[2] = 1;
  out_phi = T(1.0e30);
  const M3<T> Ainv = inverse3(M.A);
  T rel[3] = {p[0] - M.t[0], p[1] - M.t[1], p[2] - M.t[2]};
  T pl[3];
  mat_vec(Ainv, rel, pl);
  for (int a = 0; a < 3; ++a)
    if (pl[a] < T(c.aabb_lo[a]) - T(c.margin) || pl[a] > T(c.aabb_hi[a]) + T(c.margin)) return;
  if (c.n_tris == 0) return;
  T best = std::numeric_limits<T>::max();
  T best_cp[3] = {0, 0, 0};
  int best_t = -1, best_f = 0;
  for (int t = 0; t < c.n_tris; ++t) {
    const int* tri = c.tris + 3 * t;
    T A[3], B[3], Cc[3], q[3];
    for (int i = 0; i < 3; ++i) {
      A[i] = T(c.verts[3 * tri[0] + i]);
      B[i] = T(c.verts[3 * tri[1] + i]);
      Cc[i] = T(c.verts[3 * tri[2] + i]);
    }
    int f;
    closest_on_triangle(pl, A, B, Cc, q, f);
    const T d2 = (pl[0] - q[0]) * (pl[0] - q[0]) + (pl[1] - q[1]) * (pl[1] - q[1]) + (pl[2] - q[2]) * (pl[2] - q[2]);
    if (d2 < best) {
      best = d2;
      best_t = t;
      best_f = f;
      for (int i = 0; i < 3; ++i) best_cp[i] = q[i];
    }
  }
  const int* tri = c.tris + 3 * best_t;
  double pn[3];
  if (best_f == 0) {
    for (int i = 0; i < 3; ++i) pn[i] = mn.face[3 * best_t + i];
  } else if (best_f <= 3) {
    for (int i = 0; i < 3; ++i) pn[i] = mn.vert[3 * tri[best_f - 1] + i];
  } else {
    for (int i = 0; i < 3; ++i) pn[i] = mn.edge[3 * (3 * best_t + (best_f - 4)) + i];
  }
  const T off_l[3] = {pl[0] - best_cp[0], pl[1] - best_cp[1], pl[2] - best_cp[2]};
  const T sd = off_l[0] * T(pn[0]) + off_l[1] * T(pn[1]) + off_l[2] * T(pn[2]);
  const T sign = sd < T(0) ? T(-1) : T(1);
  T d[3];
  mat_vec(M.A, off_l, d);  // world offset from the closest surface point
  const T dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  T n[3];
  if (dist > T(1.0e-9)) {
    for (int i = 0; i < 3; ++i) n[i] = d[i] * (sign / dist);
  } else {
    // face normal through the inverse transpose
    const T fl[3] = {T(mn.face[3 * best_t]), T(mn.face[3 * best_t + 1]), T(mn.face[3 * best_t + 2])};
    const M3<T> AiT = mtrans(Ainv);
    mat_vec(AiT, fl, n);
    const T l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int i = 0; i < 3; ++i) n[i] /= l;
  }
  const T phi = sign * dist;
  if (c.closed == 0 && sign < T(0) && dist > thickness) return;
  for (int i = 0; i < 3; ++i) out_n[i] = n[i];
  out_phi = phi;
}

template <class T> void collider_velocity(const OracleCollider& c, const Affine<T>& M, T s, const T p[3],
                                          T inv_frame_dt, T out[3]) {
  const M3<T> Ainv = inverse3(M.A);
  const T rel[3] = {p[0] - M.t[0], p[1] - M.t[1], p[2] - M.t[2]};
  T pl[3];
  mat_vec(Ainv, rel, pl);
  const T h = T(1.0e-3);
  const Affine<T> Mp = collider_matrix<T>(c, s + h), Mm = collider_matrix<T>(c, s - h);
  const T k = inv_frame_dt / (T(2) * h);
  for (int i = 0; i < 3; ++i) {
    T acc = 0;
    for (int j = 0; j < 3; ++j) acc += (Mp.A.a[i][j] - Mm.A.a[i][j]) * k * pl[j];
    out[i] = acc + (Mp.t[i] - Mm.t[i]) * k;
  }
}

// ---------------------------------------------------------------------------------------------
// Substep kernels

inline int64_t node_index(const OracleDomain& d, int i, int j, int k) {
  return (int64_t(i) * d.dims[1] + j) * d.dims[2] + k;
}

template <class T>
void p2g(const OracleDomain& d, const OracleMaterial* mats, int64_t n, const float* x, const float* v,
         const float* C, const float* F, const float* Jp, const int32_t* mat, T dt, T* gv, T* gm, int threads) {
  const T origin[3] = {T(d.origin[0]), T(d.origin[1]), T(d.origin[2])};
  const T inv_dx = T(d.inv_dx), dx = T(d.dx), p_vol = T(d.p_vol);
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t p = 0; p < n; ++p) {
    const OracleMaterial& m = mats[mat[p]];
    const M3<T> Fp = load_m<T>(F + 9 * p);
    M3<T> tau;
    if (m.model == MODEL_VON_MISES || m.model == MODEL_DRUCKER_PRAGER) {
      tau = hencky_tau<T>(Fp, T(m.mu), T(m.lam));
    } else if (m.model == MODEL_SNOW) {
      const T h = std::min(T(std::exp(T(m.hardening) * (T(1) - T(Jp[p])))), T(m.max_hardening));
      tau = fixed_corotated_tau<T>(Fp, T(m.mu) * h, T(m.lam) * h);
    } else {
      tau = fixed_corotated_tau<T>(Fp, T(m.mu), T(m.lam));
    }
    const T pm = T(m.p_mass);
    const T k = -dt * p_vol * T(4) * inv_dx * inv_dx;
    M3<T> aff;
    for (int i = 0; i < 9; ++i) aff.a[i / 3][i % 3] = k * tau.a[i / 3][i % 3] + pm * T(C[9 * p + i]);
    T xp[3], fx[3];
    int base[3];
    for (int a = 0; a < 3; ++a) {
      xp[a] = (T(x[3 * p + a]) - origin[a]) * inv_dx;
      base[a] = int(std::floor(xp[a] - T(0.5)));
      fx[a] = xp[a] - T(base[a]);
    }
    const T mv[3] = {pm * T(v[3 * p]), pm * T(v[3 * p + 1]), pm * T(v[3 * p + 2])};
    for (int i = 0; i < 3; ++i) {
      const T wi = bspline_w(fx[0], i);
      for (int j = 0; j < 3; ++j) {
        const T wj = bspline_w(fx[1], j);
        for (int kk = 0; kk < 3; ++kk) {
          const T w = wi * wj * bspline_w(fx[2], kk);
          const T dpos[3] = {(T(i) - fx[0]) * dx, (T(j) - fx[1]) * dx, (T(kk) - fx[2]) * dx};
          const int64_t id = node_index(d, base[0] + i, base[1] + j, base[2] + kk);
          for (int a = 0; a < 3; ++a) {
            const T val = w * (mv[a] + aff.a[a][0] * dpos[0] + aff.a[a][1] * dpos[1] + aff.a[a][2] * dpos[2]);
#pragma omp atomic
            gv[3 * id + a] += val;
          }
          const T wm = w * pm;
#pragma omp atomic
          gm[id] += wm;
        }
      }
    }
  }
}

template <class T> void wall_contact(T vel[3], int axis, T friction) {
  const T vn = std::fabs(vel[axis]);
  vel[axis] = T(0);
  const T len = std::sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
  if (len > T(0)) {
    const T f = std::max(T(0), T(1) - friction * vn / len);
    for (int a = 0; a < 3; ++a) vel[a] *= f;
  }
}

template <class T> void grid_update(const OracleDomain& d, T dt, T* gv, const T* gm, int threads) {
  const int64_t total = int64_t(d.dims[0]) * d.dims[1] * d.dims[2];
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t id = 0; id < total; ++id) {
    const T m = gm[id];
    if (m <= T(0)) continue;
    const int node[3] = {int(id / (int64_t(d.dims[1]) * d.dims[2])), int((id / d.dims[2]) % d.dims[1]),
                         int(id % d.dims[2])};
    T vel[3];
    for (int a = 0; a < 3; ++a) vel[a] = gv[3 * id + a] / m + dt * T(d.gravity[a]);
    for (int a = 0; a < 3; ++a) {
      if (d.lo_on[a] != 0 && node[a] <= d.wall_lo_node[a] && vel[a] < T(0)) wall_contact(vel, a, T(d.wall_friction));
      if (d.hi_on[a] != 0 && node[a] >= d.wall_hi_node[a] && vel[a] > T(0)) wall_contact(vel, a, T(d.wall_friction));
    }
    for (int a = 0; a < 3; ++a) gv[3 * id + a] = vel[a];
  }
}

template <class T>
void grid_colliders(const OracleDomain& d, const OracleCollider* cs, const std::vector<MeshNormals>& mns,
                    int n_col, T s, T inv_frame_dt, T* gv, const T* gm, int threads) {
  const int64_t total = int64_t(d.dims[0]) * d.dims[1] * d.dims[2];
  const T dx = T(d.dx), thickness = T(2.0) * dx, band = dx;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256)
  for (int64_t id = 0; id < total; ++id) {
    if (gm[id] <= T(0)) continue;
    const int node[3] = {int(id / (int64_t(d.dims[1]) * d.dims[2])), int((id / d.dims[2]) % d.dims[1]),
                         int(id % d.dims[2])};
    T pos[3];
    for (int a = 0; a < 3; ++a) pos[a] = T(d.origin[a]) + T(node[a]) * dx;
    T vel[3] = {gv[3 * id], gv[3 * id + 1], gv[3 * id + 2]};
    for (int c = 0; c < n_col; ++c) {
      const Affine<T> M = collider_matrix<T>(cs[c], s);
      T n[3], phi;
      collider_sdf<T>(cs[c], mns[c], M, pos, thickness, n, phi);
      if (phi < band) {
        T vcol[3];
        collider_velocity<T>(cs[c], M, s, pos, inv_frame_dt, vcol);
        T rel[3];
        for (int a = 0; a < 3; ++a) rel[a] = vel[a] - vcol[a];
        const T vn = rel[0] * n[0] + rel[1] * n[1] + rel[2] * n[2];
        if (vn < T(0)) {
          T rt[3];
          for (int a = 0; a < 3; ++a) rt[a] = rel[a] - vn * n[a];
          const T rl = std::sqrt(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
          if (rl > T(0)) {
            const T f = std::max(T(0), T(1) - T(cs[c].friction) * (-vn) / rl);
            for (int a = 0; a < 3; ++a) rt[a] *= f;
          }
          if (phi < T(0)) {
            for (int a = 0; a < 3; ++a) vel[a] = vcol[a] + rt[a];
          } else {
            for (int a = 0; a < 3; ++a) vel[a] = vcol[a] + rt[a] + vn * n[a];
          }
        }
      }
    }
    for (int a = 0; a < 3; ++a) gv[3 * id + a] = vel[a];
  }
}

template <class T>
void g2p(const OracleDomain& d, const OracleMaterial* mats, int64_t n, float* x, float* v, float* C, float* F,
         float* Jp, const int32_t* mat, const T* gv, T dt, int threads) {
  const T origin[3] = {T(d.origin[0]), T(d.origin[1]), T(d.origin[2])};
  const T inv_dx = T(d.inv_dx), dx = T(d.dx);
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t p = 0; p < n; ++p) {
    T xp[3], fx[3], xo[3];
    int base[3];
    for (int a = 0; a < 3; ++a) {
      xo[a] = T(x[3 * p + a]);
      xp[a] = (xo[a] - origin[a]) * inv_dx;
      base[a] = int(std::floor(xp[a] - T(0.5)));
      fx[a] = xp[a] - T(base[a]);
    }
    T nv[3] = {0, 0, 0};
    M3<T> nC = mzero<T>();
    for (int i = 0; i < 3; ++i) {
      const T wi = bspline_w(fx[0], i);
      for (int j = 0; j < 3; ++j) {
        const T wj = bspline_w(fx[1], j);
        for (int kk = 0; kk < 3; ++kk) {
          const T w = wi * wj * bspline_w(fx[2], kk);
          const T dpos[3] = {(T(i) - fx[0]) * dx, (T(j) - fx[1]) * dx, (T(kk) - fx[2]) * dx};
          const int64_t id = node_index(d, base[0] + i, base[1] + j, base[2] + kk);
          const T g[3] = {gv[3 * id], gv[3 * id + 1], gv[3 * id + 2]};
          const T wc = T(4) * inv_dx * inv_dx * w;
          for (int a = 0; a < 3; ++a) {
            nv[a] += w * g[a];
            for (int b = 0; b < 3; ++b) nC.a[a][b] += wc * g[a] * dpos[b];
          }
        }
      }
    }
    for (int a = 0; a < 3; ++a) v[3 * p + a] = float(nv[a]);
    store_m(C + 9 * p, nC);
    T nx[3];
    for (int a = 0; a < 3; ++a) nx[a] = xo[a] + dt * nv[a];
    for (int a = 0; a < 3; ++a) {
      if (d.lo_on[a] != 0 && nx[a] < T(d.wall_lo[a])) nx[a] = T(d.wall_lo[a]);
      if (d.hi_on[a] != 0 && nx[a] > T(d.wall_hi[a])) nx[a] = T(d.wall_hi[a]);
    }
    if (std::isfinite(nx[0]) && std::isfinite(nx[1]) && std::isfinite(nx[2]))
      for (int a = 0; a < 3; ++a) nx[a] = std::min(std::max(nx[a], T(d.clamp_lo[a])), T(d.clamp_hi[a]));
    for (int a = 0; a < 3; ++a) x[3 * p + a] = float(nx[a]);
    M3<T> G = meye<T>();
    for (int i = 0; i < 9; ++i) G.a[i / 3][i % 3] += dt * nC.a[i / 3][i % 3];
    const M3<T> Fn = mmul(G, load_m<T>(F + 9 * p));
    const OracleMaterial& m = mats[mat[p]];
    if (m.model == MODEL_ELASTIC) {
      store_m(F + 9 * p, Fn);
      continue;
    }
    M3<T> U, V;
    T sig[3], ns[3];
    svd3(Fn, U, sig, V, svd_sweeps<T>());
    if (m.model == MODEL_VON_MISES) {
      project_von_mises<T>(sig, T(m.mu), T(m.yield_stress), ns);
    } else if (m.model == MODEL_DRUCKER_PRAGER) {
      project_drucker_prager<T>(sig, T(m.mu), T(m.lam), T(m.alpha), ns);
    } else {
      const T lo = T(1) - T(m.crit_c), hi = T(1) + T(m.crit_s);
      for (int a = 0; a < 3; ++a) ns[a] = std::min(std::max(sig[a], lo), hi);
      Jp[p] = float(T(Jp[p]) * (sig[0] * sig[1] * sig[2]) / (ns[0] * ns[1] * ns[2]));
    }
    M3<T> out;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        out.a[i][j] = U.a[i][0] * ns[0] * V.a[j][0] + U.a[i][1] * ns[1] * V.a[j][1] + U.a[i][2] * ns[2] * V.a[j][2];
    store_m(F + 9 * p, out);
  }
}

template <class T>
void project_particles(const OracleDomain& d, const OracleCollider* cs, const std::vector<MeshNormals>& mns,
                       int n_col, int64_t n, float* x, float* v, T s, T inv_frame_dt, int threads) {
  const T thickness = T(2.0) * T(d.dx);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256)
  for (int64_t p = 0; p < n; ++p) {
    T pos[3] = {T(x[3 * p]), T(x[3 * p + 1]), T(x[3 * p + 2])};
    T vel[3] = {T(v[3 * p]), T(v[3 * p + 1]), T(v[3 * p + 2])};
    for (int c = 0; c < n_col; ++c) {
      const Affine<T> M = collider_matrix<T>(cs[c], s);
      T nrm[3], phi;
      collider_sdf<T>(cs[c], mns[c], M, pos, thickness, nrm, phi);
      if (phi < T(0)) {
        for (int a = 0; a < 3; ++a) pos[a] -= phi * nrm[a];
        T vcol[3];
        collider_velocity<T>(cs[c], M, s, pos, inv_frame_dt, vcol);
        const T vn = (vel[0] - vcol[0]) * nrm[0] + (vel[1] - vcol[1]) * nrm[1] + (vel[2] - vcol[2]) * nrm[2];
        if (vn < T(0))
          for (int a = 0; a < 3; ++a) vel[a] -= vn * nrm[a];
      }
    }
    for (int a = 0; a < 3; ++a) {
      x[3 * p + a] = float(pos[a]);
      v[3 * p + a] = float(vel[a]);
    }
  }
}

template <class T>
void substep(const OracleDomain& d, const OracleMaterial* mats, const OracleCollider* cs, int n_col, int64_t n,
             float* x, float* v, float* C, float* F, float* Jp, const int32_t* mat, T dt, T s, T inv_frame_dt,
             T* gv, T* gm, int threads) {
  const int64_t total = int64_t(d.dims[0]) * d.dims[1] * d.dims[2];
  std::memset(gv, 0, sizeof(T) * 3 * total);
  std::memset(gm, 0, sizeof(T) * total);
  p2g<T>(d, mats, n, x, v, C, F, Jp, mat, dt, gv, gm, threads);
  grid_update<T>(d, dt, gv, gm, threads);
  std::vector<MeshNormals> mns;
  for (int c = 0; c < n_col; ++c) mns.push_back(mesh_normals(cs[c]));
  if (n_col) grid_colliders<T>(d, cs, mns, n_col, s, inv_frame_dt, gv, gm, threads);
  g2p<T>(d, mats, n, x, v, C, F, Jp, mat, gv, dt, threads);
  if (n_col) project_particles<T>(d, cs, mns, n_col, n, x, v, s, inv_frame_dt, threads);
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// C entry points (precision suffix _f32 / _f64: the scalar type the arithmetic runs in)

extern "C" {

void oracle_svd3_f64(const double* F, double* U, double* sig, double* V) {
  M3<double> f, u, vv;
  for (int i = 0; i < 9; ++i) f.a[i / 3][i % 3] = F[i];
  svd3<double>(f, u, sig, vv, svd_sweeps<double>());
  for (int i = 0; i < 9; ++i) { U[i] = u.a[i / 3][i % 3]; V[i] = vv.a[i / 3][i % 3]; }
}

void oracle_svd3_f32(const float* F, float* U, float* sig, float* V) {
  M3<float> f, u, vv;
  for (int i = 0; i < 9; ++i) f.a[i / 3][i % 3] = F[i];
  svd3<float>(f, u, sig, vv, svd_sweeps<float>());
  for (int i = 0; i < 9; ++i) { U[i] = u.a[i / 3][i % 3]; V[i] = vv.a[i / 3][i % 3]; }
}

// Kirchhoff stress of one material model on one F (tests)
void oracle_tau_f64(int32_t model, double mu, double lam, double hardening, double max_hardening, double Jp,
                    const double* F, double* tau) {
  M3<double> f, t;
  for (int i = 0; i < 9; ++i) f.a[i / 3][i % 3] = F[i];
  if (model == MODEL_VON_MISES || model == MODEL_DRUCKER_PRAGER) {
    t = hencky_tau<double>(f, mu, lam);
  } else if (model == MODEL_SNOW) {
    const double h = std::min(std::exp(hardening * (1.0 - Jp)), max_hardening);
    t = fixed_corotated_tau<double>(f, mu * h, lam * h);
  } else {
    t = fixed_corotated_tau<double>(f, mu, lam);
  }
  for (int i = 0; i < 9; ++i) tau[i] = t.a[i / 3][i % 3];
}

// return maps on singular values (tests)
void oracle_project_f64(int32_t model, double mu, double lam, double yield_stress, double alpha, double crit_c,
                        double crit_s, const double* sig, double* out) {
  if (model == MODEL_VON_MISES) {
    project_von_mises<double>(sig, mu, yield_stress, out);
  } else if (model == MODEL_DRUCKER_PRAGER) {
    project_drucker_prager<double>(sig, mu, lam, alpha, out);
  } else if (model == MODEL_SNOW) {
    for (int a = 0; a < 3; ++a) out[a] = std::min(std::max(sig[a], 1.0 - crit_c), 1.0 + crit_s);
  } else {
    for (int a = 0; a < 3; ++a) out[a] = sig[a];
  }
}

double oracle_bspline_w_f64(double fx, int32_t i) { return bspline_w<double>(fx, i); }

void oracle_p2g_f32(const OracleDomain* d, const OracleMaterial* m, int64_t n, const float* x, const float* v,
                    const float* C, const float* F, const float* Jp, const int32_t* mat, double dt, float* gv,
                    float* gm, int32_t threads) {
  p2g<float>(*d, m, n, x, v, C, F, Jp, mat, float(dt), gv, gm, threads);
}
void oracle_p2g_f64(const OracleDomain* d, const OracleMaterial* m, int64_t n, const float* x, const float* v,
                    const float* C, const float* F, const float* Jp, const int32_t* mat, double dt, double* gv,
                    double* gm, int32_t threads) {
  p2g<double>(*d, m, n, x, v, C, F, Jp, mat, dt, gv, gm, threads);
}

void oracle_grid_update_f32(const OracleDomain* d, double dt, float* gv, const float* gm, int32_t threads) {
  grid_update<float>(*d, float(dt), gv, gm, threads);
}
void oracle_grid_update_f64(const OracleDomain* d, double dt, double* gv, const double* gm, int32_t threads) {
  grid_update<double>(*d, dt, gv, gm, threads);
}

void oracle_grid_colliders_f32(const OracleDomain* d, const OracleCollider* cs, int32_t n_col, double s,
                               double inv_frame_dt, float* gv, const float* gm, int32_t threads) {
  std::vector<MeshNormals> mns;
  for (int c = 0; c < n_col; ++c) mns.push_back(mesh_normals(cs[c]));
  grid_colliders<float>(*d, cs, mns, n_col, float(s), float(inv_frame_dt), gv, gm, threads);
}

void oracle_g2p_f32(const OracleDomain* d, const OracleMaterial* m, int64_t n, float* x, float* v, float* C,
                    float* F, float* Jp, const int32_t* mat, const float* gv, double dt, int32_t threads) {
  g2p<float>(*d, m, n, x, v, C, F, Jp, mat, gv, float(dt), threads);
}
void oracle_g2p_f64(const OracleDomain* d, const OracleMaterial* m, int64_t n, float* x, float* v, float* C,
                    float* F, float* Jp, const int32_t* mat, const double* gv, double dt, int32_t threads) {
  g2p<double>(*d, m, n, x, v, C, F, Jp, mat, gv, dt, threads);
}

void oracle_project_particles_f32(const OracleDomain* d, const OracleCollider* cs, int32_t n_col, int64_t n,
                                  float* x, float* v, double s, double inv_frame_dt, int32_t threads) {
  std::vector<MeshNormals> mns;
  for (int c = 0; c < n_col; ++c) mns.push_back(mesh_normals(cs[c]));
  project_particles<float>(*d, cs, mns, n_col, n, x, v, float(s), float(inv_frame_dt), threads);
}

// signed distance + normal of one point against one collider (tests)
void oracle_collider_sdf_f64(const OracleCollider* c, double s, const double* p, double thickness, double* n,
                             double* phi) {
  const MeshNormals mn = mesh_normals(*c);
  const Affine<double> M = collider_matrix<double>(*c, s);
  collider_sdf<double>(*c, mn, M, p, thickness, n, *phi);
}

void oracle_collider_velocity_f64(const OracleCollider* c, double s, const double* p, double inv_frame_dt,
                                  double* out) {
  const Affine<double> M = collider_matrix<double>(*c, s);
  collider_velocity<double>(*c, M, s, p, inv_frame_dt, out);
}

// the smallest signed distance of n points (float32 rows) to the colliders at frame fraction s, as
// project_particles evaluates it (colliders.py:137-140; 1e30 beyond every collider's reach), for tests that
// must tell particles lying on a surface, where rounding decides the projection, from the others
void oracle_collider_phi_f32(const OracleDomain* d, const OracleCollider* cs, int32_t n_col, int64_t n,
                             const float* x, double s, float* phi, int32_t threads) {
  std::vector<MeshNormals> mns;
  for (int c = 0; c < n_col; ++c) mns.push_back(mesh_normals(cs[c]));
  const float thickness = 2.0f * d->dx;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256)
  for (int64_t p = 0; p < n; ++p) {
    const float pos[3] = {x[3 * p], x[3 * p + 1], x[3 * p + 2]};
    float best = 1.0e30f;
    for (int c = 0; c < n_col; ++c) {
      const Affine<float> M = collider_matrix<float>(cs[c], float(s));
      float nrm[3], ph;
      collider_sdf<float>(cs[c], mns[c], M, pos, thickness, nrm, ph);
      best = std::min(best, ph);
    }
    phi[p] = best;
  }
}

// full substep: zero, p2g, grid_update, [grid_colliders], g2p, [project_particles] (solver.py:187-211)
void oracle_substep_f32(const OracleDomain* d, const OracleMaterial* m, const OracleCollider* cs, int32_t n_col,
                        int64_t n, float* x, float* v, float* C, float* F, float* Jp, const int32_t* mat, double dt,
                        double s, double inv_frame_dt, float* gv, float* gm, int32_t threads) {
  substep<float>(*d, m, cs, n_col, n, x, v, C, F, Jp, mat, float(dt), float(s), float(inv_frame_dt), gv, gm, threads);
}
void oracle_substep_f64(const OracleDomain* d, const OracleMaterial* m, const OracleCollider* cs, int32_t n_col,
                        int64_t n, float* x, float* v, float* C, float* F, float* Jp, const int32_t* mat, double dt,
                        double s, double inv_frame_dt, double* gv, double* gm, int32_t threads) {
  substep<double>(*d, m, cs, n_col, n, x, v, C, F, Jp, mat, dt, s, inv_frame_dt, gv, gm, threads);
}

// max |v| (kernels.py:266-269); a NaN or inf speed is returned as is
float oracle_max_speed(int64_t n, const float* v) {
  float best = 0.0f;
  bool bad = false;
  for (int64_t p = 0; p < n; ++p) {
    const float s = std::sqrt(v[3 * p] * v[3 * p] + v[3 * p + 1] * v[3 * p + 1] + v[3 * p + 2] * v[3 * p + 2]);
    if (!std::isfinite(s)) bad = true;
    best = std::max(best, s);
  }
  return bad ? std::numeric_limits<float>::infinity() : best;
}

// occupancy for inflow emission (kernels.py:272-298): volume splat, then nearest-node threshold
void oracle_is_occupied(const OracleDomain* d, int64_t n, const float* x, int64_t m, const float* pts,
                        uint8_t* out) {
  const int64_t total = int64_t(d->dims[0]) * d->dims[1] * d->dims[2];
  std::vector<float> vol(total, 0.0f);
  for (int64_t p = 0; p < n; ++p) {
    float xp[3], fx[3];
    int base[3];
    for (int a = 0; a < 3; ++a) {
      xp[a] = (x[3 * p + a] - d->origin[a]) * d->inv_dx;
      base[a] = int(std::floor(xp[a] - 0.5f));
      fx[a] = xp[a] - float(base[a]);
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          vol[node_index(*d, base[0] + i, base[1] + j, base[2] + k)] +=
              bspline_w(fx[0], i) * bspline_w(fx[1], j) * bspline_w(fx[2], k) * d->p_vol;
  }
  const float cell_vol = d->cell_vol;
  for (int64_t q = 0; q < m; ++q) {
    int g[3];
    for (int a = 0; a < 3; ++a) g[a] = int(std::round((pts[3 * q + a] - d->origin[a]) * d->inv_dx));
    bool inside = true;
    for (int a = 0; a < 3; ++a) inside = inside && g[a] >= 0 && g[a] < d->dims[a];
    out[q] = inside && vol[node_index(*d, g[0], g[1], g[2])] >= 0.5f * cell_vol ? 1 : 0;
  }
}

int32_t oracle_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
