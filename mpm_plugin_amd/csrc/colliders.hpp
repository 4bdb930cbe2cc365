// Animated rigid mesh colliders on gfx950 (colliders.py semantics).
//
// Pose (colliders.py:20-37): M(s) = [R(slerp(q0, q1, s)) (S0 + s (S1 - S0)) | t0 + s (t1 - t0)],
// s the substep's position in the frame. Each workgroup derives M(s), its inverse and the central
// difference velocity matrix dM = (M(s + h) - M(s - h)) / (2 h) * inv_frame_dt (h = 1e-3,
// colliders.py:68-74) once, into LDS, for every collider.
//
// Signed distance (collider_sdf, colliders.py:40-65): the query goes to mesh-local space, is dropped
// outside the mesh AABB grown by `margin`, then finds the closest point on the mesh. The reference
// queries a BVH (wp.mesh_query_point_sign_normal); here each collider carries a local-space voxel
// grid covering the AABB plus the margin, built on the GPU when the mesh is set or the margin grows.
// Voxel v (centre c, half diagonal h) lists every triangle t with dist(c, t) <= dmin(c) + 2 h, where
// dmin(c) is the distance from c to the mesh. For any point p in the voxel the closest triangle
// satisfies dist(c, t) <= dist(p, t) + h <= dist(p, mesh) + h <= dmin(c) + 2 h, so it is in the list
// and the query is exact, not interpolated. Lists are in ascending triangle order and the first
// strictly smaller distance wins, as in a linear scan. The sign is that of (p - closest) against the
// angle-weighted pseudonormal of the closest feature (face, edge or vertex; Baerentzen & Aanaes 2005),
// which is inside-negative for closed meshes and back-side-negative for open ones. Closest point on a
// triangle: Ericson, Real-Time Collision Detection, 5.1.5.
//
// grid contact (grid_colliders, colliders.py:77-117) runs inside the grid update on every node with
// mass; particle projection (project_particles, colliders.py:120-149) runs at the end of G2P.
#pragma once

#include "kernels.hpp"

namespace mpm {

constexpr int kMaxColliders = 8;

struct ColliderDev {
  float t0[3], t1[3], q0[4], q1[4], S0[9], S1[9];
  float aabb_lo[3], aabb_hi[3];
  float margin, friction;
  int closed;
  int tri_offset, n_tris;
  // voxel grid of candidate triangles (local space)
  float g_lo[3];
  float g_inv_h;
  int g_res[3];
  int cell_offset;  // first cell of this collider in cell_start
};

// triangle corners, then pseudonormals: face, edges (ab, bc, ca), vertices (a, b, c)
struct TriGeom {
  float v[9];
};
struct TriNormals {
  float n[21];
};

struct ColliderSet {
  int count;
  MPM_GLOBAL const ColliderDev* cols;
  MPM_GLOBAL const TriGeom* tris;
  MPM_GLOBAL const TriNormals* normals;
  MPM_GLOBAL const uint32_t* cell_start;  // per cell: first candidate; cell_start[cell + 1] ends it
  MPM_GLOBAL const uint32_t* cell_tris;   // candidate triangle indices (collider-local)
  // per candidate: its distance from the cell's centre; each cell's list is in increasing order of it
  // (ties by index), or all zero for a list too long to sort at build time
  MPM_GLOBAL const float* cell_key;
  // per cell: a lower bound of the local distance from any point of the cell to the surface, signed by
  // the side of the cell's centre (0: the surface may come that close)
  MPM_GLOBAL const float* cell_far;
  float thickness;             // open meshes are solid this far behind the surface (2 dx)
  float band;                  // grid contact band in front of the surface (dx)
};

struct ColliderFrame {
  float A[9], t[3];        // M(s)
  float Ai[9], ti[3];      // M(s)^-1
  float dA[9], dt[3];      // velocity matrix
  float smin;              // at most the smallest singular value of A, so |A u| >= smin |u|
};

// every collider's frame of one substep and its record, computed once on the host and passed by value to
// the kernels (they read them through the scalar cache; no workgroup recomputes a frame, and a pose
// change needs no upload)
struct ColliderFrames {
  ColliderFrame f[kMaxColliders];
  ColliderDev cols[kMaxColliders];
};

// the collider set with its records read from the kernel arguments
__device__ __forceinline__ ColliderSet with_records(ColliderSet cs, const ColliderFrames& fr) {
  cs.cols = (MPM_GLOBAL const ColliderDev*)fr.cols;
  return cs;
}

__host__ __device__ inline void quat_rot(const float (&q)[4], float (&R)[9]) {
  const float x = q[0], y = q[1], z = q[2], w = q[3];
  R[0] = 1.0f - 2.0f * (y * y + z * z); R[1] = 2.0f * (x * y - z * w); R[2] = 2.0f * (x * z + y * w);
  R[3] = 2.0f * (x * y + z * w); R[4] = 1.0f - 2.0f * (x * x + z * z); R[5] = 2.0f * (y * z - x * w);
  R[6] = 2.0f * (x * z - y * w); R[7] = 2.0f * (y * z + x * w); R[8] = 1.0f - 2.0f * (x * x + y * y);
}

// pose at s (colliders.py:30-37): slerp of the rotation, lerp of stretch and translation
__host__ __device__ inline void collider_matrix(const ColliderDev& c, float s, float (&A)[9], float (&t)[3]) {
  float d = c.q0[0] * c.q1[0] + c.q0[1] * c.q1[1] + c.q0[2] * c.q1[2] + c.q0[3] * c.q1[3];
  d = fminf(fmaxf(d, -1.0f), 1.0f);
  const float ang = acosf(d);
  float wa, wb;
  if (ang < 1.0e-6f) {
    wa = 1.0f - s;
    wb = s;
  } else {
    const float si = sinf(ang);
    wa = sinf((1.0f - s) * ang) / si;
    wb = sinf(s * ang) / si;
  }
  float q[4], nq = 0.0f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    q[i] = wa * c.q0[i] + wb * c.q1[i];
    nq += q[i] * q[i];
  }
  nq = 1.0f / sqrtf(nq);
#pragma unroll
  for (int i = 0; i < 4; ++i) q[i] *= nq;
  float R[9];
  quat_rot(q, R);
  float S[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) S[i] = c.S0[i] + s * (c.S1[i] - c.S0[i]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A[3 * i + j] = R[3 * i] * S[j] + R[3 * i + 1] * S[3 + j] + R[3 * i + 2] * S[6 + j];
#pragma unroll
  for (int i = 0; i < 3; ++i) t[i] = c.t0[i] + s * (c.t1[i] - c.t0[i]);
}

__host__ __device__ inline void collider_frame(const ColliderDev& c, float s, float inv_frame_dt, ColliderFrame& f) {
  collider_matrix(c, s, f.A, f.t);
  const float* a = f.A;
  const float c00 = a[4] * a[8] - a[5] * a[7];
  const float c01 = a[5] * a[6] - a[3] * a[8];
  const float c02 = a[3] * a[7] - a[4] * a[6];
  const float inv = 1.0f / (a[0] * c00 + a[1] * c01 + a[2] * c02);
  f.Ai[0] = c00 * inv; f.Ai[1] = (a[2] * a[7] - a[1] * a[8]) * inv; f.Ai[2] = (a[1] * a[5] - a[2] * a[4]) * inv;
  f.Ai[3] = c01 * inv; f.Ai[4] = (a[0] * a[8] - a[2] * a[6]) * inv; f.Ai[5] = (a[2] * a[3] - a[0] * a[5]) * inv;
  f.Ai[6] = c02 * inv; f.Ai[7] = (a[1] * a[6] - a[0] * a[7]) * inv; f.Ai[8] = (a[0] * a[4] - a[1] * a[3]) * inv;
#pragma unroll
  for (int i = 0; i < 3; ++i) f.ti[i] = -(f.Ai[3 * i] * f.t[0] + f.Ai[3 * i + 1] * f.t[1] + f.Ai[3 * i + 2] * f.t[2]);
  // A lower bound of the smallest singular value of A = R S (rotation times the interpolated stretch):
  // 1 / |A^-1|_F always; and, since |S u| >= u^T S u >= lambda_min((S + S^T) / 2) for a unit u, the
  // Gershgorin bound of the stretch's symmetric part, exact for uniform and axis-aligned scales. Rounded
  // down a little (R's and A's rounding): a lower bound is all the far-cell test may use.
  float fro = 0.0f;
  for (int i = 0; i < 9; ++i) fro += f.Ai[i] * f.Ai[i];
  float S[9];
  for (int i = 0; i < 9; ++i) S[i] = c.S0[i] + s * (c.S1[i] - c.S0[i]);
  float gersh = 3.0e38f;
  for (int i = 0; i < 3; ++i) {
    float r = 0.0f;
    for (int j = 0; j < 3; ++j)
      if (j != i) r += fabsf(0.5f * (S[3 * i + j] + S[3 * j + i]));
    gersh = fminf(gersh, S[4 * i] - r);
  }
  const float from_inverse = fro > 0.0f ? 1.0f / sqrtf(fro) : 0.0f;
  f.smin = 0.9999f * fmaxf(from_inverse, gersh);
  const float h = 1.0e-3f;
  float Ap[9], tp[3], Am[9], tm[3];
  collider_matrix(c, s + h, Ap, tp);
  collider_matrix(c, s - h, Am, tm);
  const float k = inv_frame_dt / (2.0f * h);
#pragma unroll
  for (int i = 0; i < 9; ++i) f.dA[i] = (Ap[i] - Am[i]) * k;
#pragma unroll
  for (int i = 0; i < 3; ++i) f.dt[i] = (tp[i] - tm[i]) * k;
}

__device__ __forceinline__ void affine_point(const float (&A)[9], const float (&t)[3], const float (&p)[3], float (&o)[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) o[i] = A[3 * i] * p[0] + A[3 * i + 1] * p[1] + A[3 * i + 2] * p[2] + t[i];
}

// closest point on triangle abc to p; feature 0 face, 1..3 vertex a/b/c, 4..6 edge ab/bc/ca
__device__ __forceinline__ void closest_on_triangle(const float (&p)[3], const float* v, float (&q)[3], int& feature) {
  const float ab[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
  const float ac[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
  const float ap[3] = {p[0] - v[0], p[1] - v[1], p[2] - v[2]};
  const float d1 = ab[0] * ap[0] + ab[1] * ap[1] + ab[2] * ap[2];
  const float d2 = ac[0] * ap[0] + ac[1] * ap[1] + ac[2] * ap[2];
  if (d1 <= 0.0f && d2 <= 0.0f) { q[0] = v[0]; q[1] = v[1]; q[2] = v[2]; feature = 1; return; }
  const float bp[3] = {p[0] - v[3], p[1] - v[4], p[2] - v[5]};
  const float d3 = ab[0] * bp[0] + ab[1] * bp[1] + ab[2] * bp[2];
  const float d4 = ac[0] * bp[0] + ac[1] * bp[1] + ac[2] * bp[2];
  if (d3 >= 0.0f && d4 <= d3) { q[0] = v[3]; q[1] = v[4]; q[2] = v[5]; feature = 2; return; }
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    const float w = d1 / (d1 - d3);
#pragma unroll
    for (int i = 0; i < 3; ++i) q[i] = v[i] + w * ab[i];
    feature = 4; return;
  }
  const float cp[3] = {p[0] - v[6], p[1] - v[7], p[2] - v[8]};
  const float d5 = ab[0] * cp[0] + ab[1] * cp[1] + ab[2] * cp[2];
  const float d6 = ac[0] * cp[0] + ac[1] * cp[1] + ac[2] * cp[2];
  if (d6 >= 0.0f && d5 <= d6) { q[0] = v[6]; q[1] = v[7]; q[2] = v[8]; feature = 3; return; }
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    const float w = d2 / (d2 - d6);
#pragma unroll
    for (int i = 0; i < 3; ++i) q[i] = v[i] + w * ac[i];
    feature = 6; return;
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
#pragma unroll
    for (int i = 0; i < 3; ++i) q[i] = v[3 + i] + w * (v[6 + i] - v[3 + i]);
    feature = 5; return;
  }
  const float den = 1.0f / (va + vb + vc);
  const float bv = vb * den, cw = vc * den;
#pragma unroll
  for (int i = 0; i < 3; ++i) q[i] = v[i] + ab[i] * bv + ac[i] * cw;
  feature = 0;
}

__device__ __forceinline__ void scan_corners(const float* v, int t, const float (&pl)[3], float& best, int& best_t,
                                             int& best_f, float (&best_q)[3]) {
  float q[3];
  int f;
  closest_on_triangle(pl, v, q, f);
  const float d2 = (pl[0] - q[0]) * (pl[0] - q[0]) + (pl[1] - q[1]) * (pl[1] - q[1]) + (pl[2] - q[2]) * (pl[2] - q[2]);
  // the smallest (distance, index): the same triangle whatever order the candidates come in
  if (d2 < best || (d2 == best && t < best_t)) {
    best = d2;
    best_t = t;
    best_f = f;
    best_q[0] = q[0]; best_q[1] = q[1]; best_q[2] = q[2];
  }
}

__device__ __forceinline__ void scan_triangle(const ColliderSet& cs, const ColliderDev& c, int t, const float (&pl)[3],
                                              float& best, int& best_t, int& best_f, float (&best_q)[3]) {
  scan_corners(cs.tris[c.tri_offset + t].v, t, pl, best, best_t, best_f, best_q);
}

// (outward normal, signed distance); phi = 1e30 when far (colliders.py:40-65)
// `threshold`: the caller only acts on phi < threshold; a point provably farther (or, for an open mesh,
// also beyond the solid thickness behind it) gets phi = 1e30 without the triangle scan
// kRound: candidates loaded per round (more rounds in flight at once, more registers)
template <int kRound>
__device__ void collider_sdf(const ColliderSet& cs, const ColliderDev& c, const ColliderFrame& f, const float (&p)[3],
                             float (&n)[3], float& phi, float threshold) {
  n[0] = 0.0f; n[1] = 0.0f; n[2] = 1.0f;
  phi = 1.0e30f;
  float pl[3];
  affine_point(f.Ai, f.ti, p, pl);
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (pl[a] < c.aabb_lo[a] - c.margin || pl[a] > c.aabb_hi[a] + c.margin) return;
  if (c.n_tris == 0) return;
  float best = 3.0e38f, q[3] = {0.0f, 0.0f, 0.0f};
  int bt = -1, bf = 0;
  int cell[3];
  bool in_grid = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float u = (pl[a] - c.g_lo[a]) * c.g_inv_h;
    cell[a] = (int)floorf(u);
    in_grid = in_grid && u >= 0.0f && cell[a] < c.g_res[a];
  }
  const int ci = in_grid ? c.cell_offset + (cell[0] * c.g_res[1] + cell[1]) * c.g_res[2] + cell[2] : 0;
  const uint32_t a0 = in_grid ? cs.cell_start[ci] : 0u, a1 = in_grid ? cs.cell_start[ci + 1] : 0u;
  if (in_grid) {
    // the world distance is at least smin x the local one: no contact possible from this cell
    const float far = cs.cell_far[ci];
    if (c.closed ? (far > 0.0f && far * f.smin >= threshold)
                 : (fabsf(far) * f.smin >= fmaxf(threshold, cs.thickness)))
      return;
  }
  if (a1 > a0) {
    // distance of the query from the cell's centre: no candidate t is closer to pl than key(t) - rp
    float rp2 = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float cc = c.g_lo[a] + ((float)cell[a] + 0.5f) / c.g_inv_h;
      rp2 += (pl[a] - cc) * (pl[a] - cc);
    }
    const float rp = sqrtf(rp2);
    // kRound candidates per round, their corners loaded together; the next round's indices and key load
    // while this round's triangles are tested, so a round costs one round trip to memory. The list is
    // sorted by key, so the scan ends at the first round whose key, less rp and a rounding allowance, is
    // farther than the best distance so far: every later candidate is at least that far.
    int t[kRound];
#pragma unroll
    for (int j = 0; j < kRound; ++j) t[j] = (int)cs.cell_tris[a0 + j < a1 ? a0 + j : a1 - 1];
    float key = cs.cell_key[a0];
    for (uint32_t k = a0; k < a1; k += kRound) {
      const float lb = key - rp - 1.0e-5f * (key + rp);
      if (lb > 0.0f && lb * lb > best * 1.0001f) break;
      TriGeom tg[kRound];
#pragma unroll
      for (int j = 0; j < kRound; ++j) tg[j] = cs.tris[c.tri_offset + t[j]];
      const uint32_t kn = k + kRound;
      int tn[kRound];
      float key_n = 0.0f;
      if (kn < a1) {
#pragma unroll
        for (int j = 0; j < kRound; ++j) tn[j] = (int)cs.cell_tris[kn + j < a1 ? kn + j : a1 - 1];
        key_n = cs.cell_key[kn];
      } else {
#pragma unroll
        for (int j = 0; j < kRound; ++j) tn[j] = t[j];
      }
#pragma unroll
      for (int j = 0; j < kRound; ++j)
        if (k + j < a1) scan_corners(tg[j].v, t[j], pl, best, bt, bf, q);
#pragma unroll
      for (int j = 0; j < kRound; ++j) t[j] = tn[j];
      key = key_n;
    }
  } else {
    for (int t = 0; t < c.n_tris; ++t) scan_triangle(cs, c, t, pl, best, bt, bf, q);
  }
  const float* nn = cs.normals[c.tri_offset + bt].n;
  const float* pn = bf == 0 ? nn : (bf <= 3 ? nn + 12 + 3 * (bf - 1) : nn + 3 + 3 * (bf - 4));
  const float ol[3] = {pl[0] - q[0], pl[1] - q[1], pl[2] - q[2]};
  const float sign = (ol[0] * pn[0] + ol[1] * pn[1] + ol[2] * pn[2]) < 0.0f ? -1.0f : 1.0f;
  float d[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) d[i] = f.A[3 * i] * ol[0] + f.A[3 * i + 1] * ol[1] + f.A[3 * i + 2] * ol[2];
  const float dist = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  float nw[3];
  if (dist > 1.0e-9f) {
    const float k = sign / dist;
    nw[0] = d[0] * k; nw[1] = d[1] * k; nw[2] = d[2] * k;
  } else {
    // face normal through the inverse transpose
#pragma unroll
    for (int i = 0; i < 3; ++i) nw[i] = f.Ai[i] * nn[0] + f.Ai[3 + i] * nn[1] + f.Ai[6 + i] * nn[2];
    const float l = 1.0f / sqrtf(nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2]);
    nw[0] *= l; nw[1] *= l; nw[2] *= l;
  }
  if (c.closed == 0 && sign < 0.0f && dist > cs.thickness) return;
  n[0] = nw[0]; n[1] = nw[1]; n[2] = nw[2];
  phi = sign * dist;
}

// velocity of the collider material point at p (colliders.py:68-74)
__device__ __forceinline__ void collider_velocity(const ColliderFrame& f, const float (&p)[3], float (&vc)[3]) {
  float pl[3];
  affine_point(f.Ai, f.ti, p, pl);
  affine_point(f.dA, f.dt, pl, vc);
}

// grid_colliders on one node with mass (colliders.py:95-117)
__device__ void collide_node_frames(const ColliderSet& cs, const ColliderFrame* frames, const GridGeom& g, int i, int j,
                                    int k, float4& node) {
  const float pos[3] = {g.origin[0] + (float)i * g.dx, g.origin[1] + (float)j * g.dx, g.origin[2] + (float)k * g.dx};
  float vel[3] = {node.x, node.y, node.z};
  for (int c = 0; c < cs.count; ++c) {
    const ColliderDev col = cs.cols[c];
    float n[3], phi;
    collider_sdf<4>(cs, col, frames[c], pos, n, phi, cs.band);
    if (phi < cs.band) {
      float vcol[3];
      collider_velocity(frames[c], pos, vcol);
      const float rel[3] = {vel[0] - vcol[0], vel[1] - vcol[1], vel[2] - vcol[2]};
      const float vn = rel[0] * n[0] + rel[1] * n[1] + rel[2] * n[2];
      if (vn < 0.0f) {
        float rt[3] = {rel[0] - vn * n[0], rel[1] - vn * n[1], rel[2] - vn * n[2]};
        const float rl = sqrtf(rt[0] * rt[0] + rt[1] * rt[1] + rt[2] * rt[2]);
        if (rl > 0.0f) {
          const float fr = fmaxf(0.0f, 1.0f - col.friction * (-vn) / rl);
          rt[0] *= fr; rt[1] *= fr; rt[2] *= fr;
        }
        if (phi < 0.0f) {
#pragma unroll
          for (int a = 0; a < 3; ++a) vel[a] = vcol[a] + rt[a];
        } else {
#pragma unroll
          for (int a = 0; a < 3; ++a) vel[a] = vcol[a] + rt[a] + vn * n[a];
        }
      }
    }
  }
  node.x = vel[0];
  node.y = vel[1];
  node.z = vel[2];
}

// project_particles on one particle (colliders.py:134-149)
__device__ void project_particle(const ColliderSet& cs, const ColliderFrame* frames, float (&pos)[3], float (&vel)[3]) {
  for (int c = 0; c < cs.count; ++c) {
    const ColliderDev col = cs.cols[c];
    float n[3], phi;
    // one candidate per round: G2P's register budget; particle queries are rarer than node queries
    collider_sdf<1>(cs, col, frames[c], pos, n, phi, 0.0f);
    if (phi < 0.0f) {
#pragma unroll
      for (int a = 0; a < 3; ++a) pos[a] -= phi * n[a];
      float vcol[3];
      collider_velocity(frames[c], pos, vcol);
      const float vn = (vel[0] - vcol[0]) * n[0] + (vel[1] - vcol[1]) * n[1] + (vel[2] - vcol[2]) * n[2];
      if (vn < 0.0f) {
#pragma unroll
        for (int a = 0; a < 3; ++a) vel[a] -= vn * n[a];
      }
    }
  }
}

// ---- candidate grid build (one thread per voxel, every triangle) -------------------------------

__device__ __forceinline__ float tri_dist2(const TriGeom* tris, int t, const float (&p)[3]) {
  float q[3];
  int f;
  closest_on_triangle(p, tris[t].v, q, f);
  return (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) + (p[2] - q[2]) * (p[2] - q[2]);
}

__device__ __forceinline__ void voxel_centre(const ColliderDev& c, int cell, float (&p)[3], float& half_diag) {
  const int k = cell % c.g_res[2], j = (cell / c.g_res[2]) % c.g_res[1], i = cell / (c.g_res[2] * c.g_res[1]);
  const float h = 1.0f / c.g_inv_h;
  p[0] = c.g_lo[0] + ((float)i + 0.5f) * h;
  p[1] = c.g_lo[1] + ((float)j + 0.5f) * h;
  p[2] = c.g_lo[2] + ((float)k + 0.5f) * h;
  half_diag = 0.8660254f * h;
}

// pass 0: candidate count per voxel; pass 1: candidate indices, ascending, and their distances from the
// centre. Only cells whose centre lies within `band` of the surface get a list; a query in any other cell
// (deep inside, or outside beyond every contact distance) scans all triangles.
// pass 0 also writes the cell's far bound: (distance of the centre - half diagonal), shrunk by the same
// slack, signed by the pseudonormal side of the centre's closest feature; 0 when not positive
__global__ void __launch_bounds__(256) k_voxel_candidates(ColliderDev c, const TriGeom* __restrict__ tris,
                                                          const TriNormals* __restrict__ normals, int n_cells,
                                                          int pass, uint32_t* __restrict__ counts,
                                                          float* __restrict__ far,
                                                          const uint32_t* __restrict__ starts,
                                                          uint32_t* __restrict__ out, float* __restrict__ out_key,
                                                          float band) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= n_cells) return;
  float p[3], hd;
  voxel_centre(c, cell, p, hd);
  const TriGeom* T = tris + c.tri_offset;
  float dmin2 = 3.0e38f, bq[3] = {0.0f, 0.0f, 0.0f};
  int bt = 0, bf = 0;
  for (int t = 0; t < c.n_tris; ++t) {
    float q[3];
    int f;
    closest_on_triangle(p, T[t].v, q, f);
    const float d2 = (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) + (p[2] - q[2]) * (p[2] - q[2]);
    if (d2 < dmin2) {
      dmin2 = d2;
      bt = t;
      bf = f;
      bq[0] = q[0]; bq[1] = q[1]; bq[2] = q[2];
    }
  }
  if (pass == 0) {
    const float lower = (sqrtf(dmin2) - hd) * 0.99999f - 1.0e-6f;
    float signed_far = 0.0f;
    if (lower > 0.0f) {
      const float* nn = normals[bt].n;
      const float* pn = bf == 0 ? nn : (bf <= 3 ? nn + 12 + 3 * (bf - 1) : nn + 3 + 3 * (bf - 4));
      const float side = (p[0] - bq[0]) * pn[0] + (p[1] - bq[1]) * pn[1] + (p[2] - bq[2]) * pn[2];
      signed_far = side < 0.0f ? -lower : lower;
    }
    far[cell] = signed_far;
  }
  // 1e-5 relative slack absorbs FP32 rounding of the distances
  const float lim = (sqrtf(dmin2) + 2.0f * hd) * 1.00001f + 1.0e-6f;
  const float lim2 = lim * lim;
  if (sqrtf(dmin2) > band) {
    if (pass == 0) counts[cell] = 0u;
    return;
  }
  if (pass == 0) {
    uint32_t cnt = 0;
    for (int t = 0; t < c.n_tris; ++t) cnt += tri_dist2(T, t, p) <= lim2 ? 1u : 0u;
    counts[cell] = cnt;
  } else {
    uint32_t w = starts[cell];
    for (int t = 0; t < c.n_tris; ++t) {
      const float d2 = tri_dist2(T, t, p);
      if (d2 <= lim2) {
        out_key[w] = sqrtf(d2);
        out[w++] = (uint32_t)t;
      }
    }
  }
}

// Sorts each cell's candidates by (key, index), one workgroup per cell (grid-stride): a bitonic sort in
// LDS of up to kSortCand entries. A longer list keeps index order with every key zero (a valid lower
// bound that never ends a scan early).
constexpr int kSortCand = 2048;
__global__ void __launch_bounds__(256) k_sort_candidates(int n_cells, const uint32_t* __restrict__ starts,
                                                         uint32_t* __restrict__ idx, float* __restrict__ key) {
  __shared__ float s_key[kSortCand];
  __shared__ uint32_t s_idx[kSortCand];
  for (int cell = blockIdx.x; cell < n_cells; cell += gridDim.x) {
    const uint32_t a0 = starts[cell], n = starts[cell + 1] - a0;
    if (n <= 1) continue;
    if (n > (uint32_t)kSortCand) {
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) key[a0 + i] = 0.0f;
      continue;
    }
    uint32_t P = 2;
    while (P < n) P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
      s_key[i] = i < n ? key[a0 + i] : 3.0e38f;
      s_idx[i] = i < n ? idx[a0 + i] : 0xffffffffu;
    }
    __syncthreads();
    for (uint32_t size = 2; size <= P; size <<= 1) {
      for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
          const uint32_t j = i ^ stride;
          if (j > i) {
            const bool up = (i & size) == 0;
            const float ki = s_key[i], kj = s_key[j];
            const uint32_t ii = s_idx[i], ij = s_idx[j];
            const bool greater = ki > kj || (ki == kj && ii > ij);
            if (greater == up) {
              s_key[i] = kj; s_key[j] = ki;
              s_idx[i] = ij; s_idx[j] = ii;
            }
          }
        }
        __syncthreads();
      }
    }
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      key[a0 + i] = s_key[i];
      idx[a0 + i] = s_idx[i];
    }
    __syncthreads();
  }
}

}  // namespace mpm
