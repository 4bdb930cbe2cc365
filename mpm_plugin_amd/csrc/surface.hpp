// Surfacing on gfx950 (reference src/mpm_surface/field.py:21-142): particles -> SPH-style density and
// velocity on a regular field grid -> iso-surface triangles -> vertex velocities.
//
//   k_surf_gather       per node: the sum over particles within the kernel radius h of the cubic B-spline
//                       weight w(|node - x| / h) (density) and of w v (momentum), as the reference's splat
//                       accumulates them with atomics (field.py:35-55); particles counting-sorted by cell
//                       first (k_surf_cell_keys, k_surf_sort_particles)
//   k_surf_normalize    per node: momentum / density (velocity), density / rest density (field.py:58-63)
//   k_surf_edges        per node and axis: whether the edge to the next node crosses iso (pass 0), and the
//                       crossing point and its trilinearly sampled velocity (pass 1, field.py:66-85)
//   k_surf_cubes        per cell: the number of triangles of its corner configuration (pass 0), and the
//                       triangles, as indices of the crossed edges' vertices (pass 1)
//
// The marching cubes tables (McTables) are built once on the host from the cube's faces, not copied from a
// published table: on every face the crossings of its edges are joined in pairs, each segment directed by
// the face's counter-clockwise walk seen from outside (from where the walk leaves the inside, field > iso,
// to where it enters it); a face with four crossings (inside corners on one diagonal) cuts each inside
// corner off. The face rule depends only on the face's four corners, so two cells sharing a face make the
// same segments there, traversed in opposite directions: the surface is closed and consistently wound.
// The segments of a configuration chain into loops (every crossed edge starts one segment and ends
// another), each fanned into triangles wound so that their normals point away from the inside: a cut-off
// inside corner at the origin gives the triangle (x edge, y edge, z edge), normal +(1, 1, 1).
#pragma once

#include "kernels.hpp"

namespace mpm {

constexpr int kMcMaxTris = 12;

struct McTables {
  int32_t count[256];
  int8_t edge[256][3 * kMcMaxTris];
};

// corner c of a cell sits at (c & 1, (c >> 1) & 1, (c >> 2) & 1); edge e joins kMcEdgeEnds[e][0] to [1]
// along axis kMcEdgeAxis[e] (from the lower corner)
constexpr int kMcEdgeEnds[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3},
                                    {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
constexpr int kMcEdgeAxis[12] = {0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2};

inline int mc_edge_of(int a, int b) {
  for (int e = 0; e < 12; ++e)
    if ((kMcEdgeEnds[e][0] == a && kMcEdgeEnds[e][1] == b) || (kMcEdgeEnds[e][0] == b && kMcEdgeEnds[e][1] == a))
      return e;
  return -1;
}

// the tables for every configuration (bit c of the index: corner c inside)
inline McTables build_mc_tables() {
  // each face's corners counter-clockwise seen from outside the cell (right-hand rule about its outward normal)
  static const int faces[6][4] = {{0, 4, 6, 2},   // x = 0, normal -x
                                  {1, 3, 7, 5},   // x = 1, normal +x
                                  {0, 1, 5, 4},   // y = 0, normal -y
                                  {2, 6, 7, 3},   // y = 1, normal +y
                                  {0, 2, 3, 1},   // z = 0, normal -z
                                  {4, 5, 7, 6}};  // z = 1, normal +z
  // the faces each edge lies on
  int edge_face[12][2];
  for (int e = 0; e < 12; ++e) {
    int k = 0;
    for (int fi = 0; fi < 6; ++fi) {
      int hits = 0;
      for (int i = 0; i < 4; ++i) hits += (faces[fi][i] == kMcEdgeEnds[e][0] || faces[fi][i] == kMcEdgeEnds[e][1]);
      if (hits == 2) edge_face[e][k++] = fi;
    }
  }
  McTables t{};
  for (int cfg = 0; cfg < 256; ++cfg) {
    const auto in = [&](int c) { return ((cfg >> c) & 1) != 0; };
    int next[12];
    bool ambiguous[6] = {};
    for (int e = 0; e < 12; ++e) next[e] = -1;
    for (int fi = 0; fi < 6; ++fi) {
      const int* f = faces[fi];
      // walking the face counter-clockwise from outside, each segment runs from where the walk leaves the
      // inside to where it entered it
      int n_in = 0;
      for (int i = 0; i < 4; ++i) n_in += in(f[i]) ? 1 : 0;
      if (n_in == 0 || n_in == 4) continue;
      ambiguous[fi] = n_in == 2 && in(f[0]) == in(f[2]);
      if (ambiguous[fi]) {
        for (int i = 0; i < 4; ++i) {
          if (!in(f[i])) continue;
          const int before = mc_edge_of(f[(i + 3) % 4], f[i]), after = mc_edge_of(f[i], f[(i + 1) % 4]);
          next[after] = before;
        }
      } else {
        int start = -1, end = -1;
        for (int i = 0; i < 4; ++i) {
          const int a = f[i], b = f[(i + 1) % 4];
          if (!in(a) && in(b)) start = mc_edge_of(a, b);
          if (in(a) && !in(b)) end = mc_edge_of(a, b);
        }
        next[end] = start;
      }
    }
    const auto share_ambiguous_face = [&](int e1, int e2) {
      for (int u : edge_face[e1])
        for (int w : edge_face[e2])
          if (u == w && ambiguous[u]) return true;
      return false;
    };
    // chain into loops and fan each from an apex that shares no four-crossing face with a vertex it is
    // joined to: a chord on such a face could be made by the neighbouring cell as well (four triangles on
    // one edge), a chord through the cell's interior cannot. Every loop of every configuration has one.
    bool used[12] = {};
    int n = 0;
    for (int e0 = 0; e0 < 12; ++e0) {
      if (next[e0] < 0 || used[e0]) continue;
      int loop[12], len = 0;
      for (int e = e0; !used[e]; e = next[e]) {
        used[e] = true;
        loop[len++] = e;
      }
      int apex = 0;
      for (int i = 0; i < len; ++i) {
        bool ok = true;
        for (int j = 0; j < len && ok; ++j)
          if (j != i && j != (i + 1) % len && j != (i + len - 1) % len && share_ambiguous_face(loop[i], loop[j]))
            ok = false;
        if (ok) {
          apex = i;
          break;
        }
      }
      // the loop runs clockwise seen from the low side, so the fan takes it backwards
      for (int k = 1; k + 1 < len; ++k) {
        t.edge[cfg][3 * n] = (int8_t)loop[apex];
        t.edge[cfg][3 * n + 1] = (int8_t)loop[(apex + k + 1) % len];
        t.edge[cfg][3 * n + 2] = (int8_t)loop[(apex + k) % len];
        ++n;
      }
    }
    t.count[cfg] = n;
  }
  return t;
}

struct SurfGrid {
  int dims[3];
  float origin[3];  // FP32, as the reference passes it to its kernels
  float inv_cell, cell, h;
};

__device__ __forceinline__ int64_t surf_node(const SurfGrid& G, int i, int j, int k) {
  return ((int64_t)i * G.dims[1] + j) * G.dims[2] + k;
}

// the unnormalised cubic B-spline of support [0, 1) (field.py:24-32) in straight-line form: both pieces
// computed, the outer one clamped to zero past q = 1
__device__ __forceinline__ float surf_cubic_select(float q) {
  const float t = fmaxf(1.0f - q, 0.0f);
  const float inner = 1.0f - 6.0f * q * q + 6.0f * q * q * q;
  const float outer = 2.0f * t * t * t;
  return q < 0.5f ? inner : outer;
}

// The splat as a gather: particles are counting-sorted by the field cell that holds them (k_surf_cell_keys,
// a scan, k_surf_sort_particles), then one wave per 4x4x4 brick of nodes (k_surf_gather) walks the cells
// within the kernel radius of the brick, row by row (a row of cells is one contiguous run of sorted
// particles), reading each particle once per wave (the index is wave-uniform: scalar loads); each lane sums
// the weights of its own node in registers and stores density and momentum once. No atomics on the field, no clearing of
// it, and a node's sum is independent of how many particles crowd it (a scatter with atomics slows down
// where the material is compressed).

// the node of the cell holding each particle (clamped into the grid; n_nodes for a non-finite particle)
// and its rank within that cell's count
__global__ void __launch_bounds__(256) k_surf_cell_keys(SurfGrid G, const float* __restrict__ x, int n,
                                                        uint32_t* __restrict__ key, uint32_t* __restrict__ rank,
                                                        uint32_t* __restrict__ count) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float g[3] = {(x[3 * p] - G.origin[0]) * G.inv_cell, (x[3 * p + 1] - G.origin[1]) * G.inv_cell,
                      (x[3 * p + 2] - G.origin[2]) * G.inv_cell};
  uint32_t k = (uint32_t)((int64_t)G.dims[0] * G.dims[1] * G.dims[2]);
  if (isfinite(g[0] + g[1] + g[2])) {
    int c[3];
    for (int a = 0; a < 3; ++a) c[a] = (int)fminf(fmaxf(floorf(g[a]), 0.0f), (float)(G.dims[a] - 1));
    k = (uint32_t)surf_node(G, c[0], c[1], c[2]);
  }
  key[p] = k;
  rank[p] = atomicAdd(&count[k], 1u);
}

// the particles in cell order: grid-unit position and velocity, one float4 each
__global__ void __launch_bounds__(256) k_surf_sort_particles(SurfGrid G, const float* __restrict__ x,
                                                             const float* __restrict__ v, int n,
                                                             const uint32_t* __restrict__ key,
                                                             const uint32_t* __restrict__ rank,
                                                             const uint32_t* __restrict__ start,
                                                             float4* __restrict__ gs, float4* __restrict__ vs) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t q = start[key[p]] + rank[p];
  gs[q] = make_float4((x[3 * p] - G.origin[0]) * G.inv_cell, (x[3 * p + 1] - G.origin[1]) * G.inv_cell,
                      (x[3 * p + 2] - G.origin[2]) * G.inv_cell, 0.0f);
  vs[q] = make_float4(v[3 * p], v[3 * p + 1], v[3 * p + 2], 0.0f);
}

constexpr int kGatherBricks = 4;  // neighbouring bricks per workgroup: they read many of the same particles

// brick b = kGatherBricks blockIdx.x + (wave % kGatherBricks): its nodes (4 bi + lane / 16, 4 bj + lane / 4
// % 4, 4 bk + lane % 4), one per lane of each of its kGatherWaves waves; wave half = wave / kGatherBricks
// sums the particles of the cell planes ci = i0 + half, i0 + half + kGatherWaves, ... and the waves' sums
// meet in LDS. At the default radius (4 cells) two shorter waves per brick rather than one long one keep
// the chip busy to the end of the launch, where the bricks inside the material (about 1300 particles each)
// would otherwise run on after the light ones have drained; at larger radii a row of cells holds a few
// particles and one wave per brick is faster. reach = ceil(h / cell).
template <int kGatherWaves>
__global__ void __launch_bounds__(64 * kGatherWaves * kGatherBricks)
    k_surf_gather(SurfGrid G, int reach, const float4* __restrict__ gs, const float4* __restrict__ vs,
                  const uint32_t* __restrict__ start, int nb1, int nb2, int64_t n_bricks, float* __restrict__ density,
                  float* __restrict__ momentum) {
  __shared__ float4 s_part[kGatherWaves > 1 ? kGatherWaves - 1 : 1][kGatherBricks][64];
  const int lane = threadIdx.x & 63;
  // the wave's index, read from the first lane so that the compiler knows it (and every particle index
  // derived from it) to be uniform: the particles then come through scalar loads
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int local = wave % kGatherBricks, half = wave / kGatherBricks;
  const int64_t b = (int64_t)blockIdx.x * kGatherBricks + local;
  const bool live = b < n_bricks;  // a brick past the end still meets the workgroup's barrier
  const int bk = (int)(b % nb2) * 4, bj = (int)((b / nb2) % nb1) * 4, bi = (int)(b / ((int64_t)nb2 * nb1)) * 4;
  const int i = bi + (lane >> 4), j = bj + ((lane >> 2) & 3), k = bk + (lane & 3);
  const float fi = (float)i, fj = (float)j, fk = (float)k;
  const float q_per_cell = G.cell / G.h;
  const int i0 = max(bi - reach, 0), i1 = live ? min(bi + 3 + reach, G.dims[0] - 1) : -1;
  const int j0 = max(bj - reach, 0), j1 = min(bj + 3 + reach, G.dims[1] - 1);
  const float r = G.h * G.inv_cell, r2 = r * r;
  float sd = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f;
  for (int ci = i0 + half; ci <= i1; ci += kGatherWaves)
    for (int cj = j0; cj <= j1; ++cj) {
      // a particle of cell (ci, cj, ck) lies at least di, dj (and dk) cells from the brick's nodes along each
      // axis: rows out of reach are skipped, and each row keeps the cells within the radius left for dk
      const int di = max(0, max(bi - ci - 1, ci - bi - 3)), dj = max(0, max(bj - cj - 1, cj - bj - 3));
      const float lat2 = (float)(di * di + dj * dj);
      if (lat2 >= r2) continue;
      const int kc = (int)ceilf(sqrtf(r2 - lat2));
      const int k0 = max(bk - kc, 0), k1 = min(bk + 3 + kc, G.dims[2] - 1);
      const uint32_t first = start[surf_node(G, ci, cj, k0)], last = start[surf_node(G, ci, cj, k1) + 1];
      if (first >= last) continue;
      // wave-uniform indices: scalar loads, the next particle's issued before the current one is used
      float4 pg = gs[first], pv = vs[first];
      for (uint32_t q = first; q < last; ++q) {
        const uint32_t nq = min(q + 1, last - 1);
        const float4 ng = gs[nq], nv = vs[nq];
        const float dx = fi - pg.x, dy = fj - pg.y, dz = fk - pg.z;
        // the hardware square root (1 ulp) and a select form of the cubic: no branches in the loop
        const float w = surf_cubic_select(__builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz) * q_per_cell);
        sd += w;
        sx += w * pv.x;
        sy += w * pv.y;
        sz += w * pv.z;
        pg = ng;
        pv = nv;
      }
    }
  if (kGatherWaves > 1) {
    if (half > 0) s_part[half - 1][local][lane] = make_float4(sd, sx, sy, sz);
    __syncthreads();
  }
  if (live && half == 0 && i < G.dims[0] && j < G.dims[1] && k < G.dims[2]) {
#pragma unroll
    for (int h = 0; h < kGatherWaves - 1; ++h) {
      const float4 o = s_part[h][local][lane];
      sd += o.x;
      sx += o.y;
      sy += o.z;
      sz += o.w;
    }
    const int64_t c = surf_node(G, i, j, k);
    density[c] = sd;
    momentum[3 * c] = sx;
    momentum[3 * c + 1] = sy;
    momentum[3 * c + 2] = sz;
  }
}

__global__ void __launch_bounds__(256) k_surf_normalize(int64_t n_nodes, float* __restrict__ density,
                                                        float* __restrict__ momentum, float inv_rest) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_nodes) return;
  const float d = density[c];
  if (d > 0.0f) {
    momentum[3 * c] = momentum[3 * c] / d;
    momentum[3 * c + 1] = momentum[3 * c + 1] / d;
    momentum[3 * c + 2] = momentum[3 * c + 2] / d;
  }
  density[c] = d * inv_rest;
}

// velocity at a point in grid units, trilinear over the clamped cell (field.py:66-85)
__device__ __forceinline__ void surf_velocity(const SurfGrid& G, const float* __restrict__ vel, const float (&g)[3],
                                              float (&out)[3]) {
  int b[3];
  float f[3];
  for (int a = 0; a < 3; ++a) {
    b[a] = min(max((int)floorf(g[a]), 0), G.dims[a] - 2);
    f[a] = g[a] - (float)b[a];
  }
  out[0] = out[1] = out[2] = 0.0f;
  for (int ia = 0; ia < 2; ++ia) {
    const float wa = ia == 0 ? 1.0f - f[0] : f[0];
    for (int ib = 0; ib < 2; ++ib) {
      const float wb = ib == 0 ? 1.0f - f[1] : f[1];
      for (int ic = 0; ic < 2; ++ic) {
        const float wc = ic == 0 ? 1.0f - f[2] : f[2];
        const int64_t c = surf_node(G, b[0] + ia, b[1] + ib, b[2] + ic);
        const float w = wa * wb * wc;
        out[0] += w * vel[3 * c];
        out[1] += w * vel[3 * c + 1];
        out[2] += w * vel[3 * c + 2];
      }
    }
  }
}

// edge (node c, axis a) joins node c to its +a neighbour; id 3 c + a
__global__ void __launch_bounds__(256) k_surf_edges(SurfGrid G, const float* __restrict__ density,
                                                    const float* __restrict__ vel, float iso, int pass,
                                                    uint32_t* __restrict__ crossed, const uint32_t* __restrict__ vid,
                                                    float* __restrict__ verts, float* __restrict__ vvel) {
  const int64_t n_nodes = (int64_t)G.dims[0] * G.dims[1] * G.dims[2];
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= 3 * n_nodes) return;
  const int64_t c = id / 3;
  const int a = (int)(id % 3);
  const int k = (int)(c % G.dims[2]), j = (int)((c / G.dims[2]) % G.dims[1]), i = (int)(c / ((int64_t)G.dims[2] * G.dims[1]));
  int nb[3] = {i, j, k};
  nb[a] += 1;
  bool cross = false;
  float v0 = 0.0f, v1 = 0.0f;
  if (nb[a] < G.dims[a]) {
    v0 = density[c];
    v1 = density[surf_node(G, nb[0], nb[1], nb[2])];
    cross = (v0 > iso) != (v1 > iso);
  }
  if (pass == 0) {
    crossed[id] = cross ? 1u : 0u;
    return;
  }
  if (!cross) return;
  const float t = (iso - v0) / (v1 - v0);
  float g[3] = {(float)i, (float)j, (float)k};
  g[a] += t;
  const int64_t o = vid[id];
  verts[3 * o] = g[0];
  verts[3 * o + 1] = g[1];
  verts[3 * o + 2] = g[2];
  float u[3];
  surf_velocity(G, vel, g, u);
  vvel[3 * o] = u[0];
  vvel[3 * o + 1] = u[1];
  vvel[3 * o + 2] = u[2];
}

__global__ void __launch_bounds__(256) k_surf_cubes(SurfGrid G, const float* __restrict__ density, float iso,
                                                    const McTables* __restrict__ T, int pass,
                                                    uint32_t* __restrict__ ntri, const uint32_t* __restrict__ tri_off,
                                                    const uint32_t* __restrict__ vid, int32_t* __restrict__ tris) {
  const int64_t cells = (int64_t)(G.dims[0] - 1) * (G.dims[1] - 1) * (G.dims[2] - 1);
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= cells) return;
  const int k = (int)(q % (G.dims[2] - 1)), j = (int)((q / (G.dims[2] - 1)) % (G.dims[1] - 1)),
            i = (int)(q / ((int64_t)(G.dims[2] - 1) * (G.dims[1] - 1)));
  int cfg = 0;
  for (int c = 0; c < 8; ++c)
    if (density[surf_node(G, i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1))] > iso) cfg |= 1 << c;
  if (pass == 0) {
    ntri[q] = (uint32_t)T->count[cfg];
    return;
  }
  const int n = T->count[cfg];
  int32_t* out = tris + 3 * (int64_t)tri_off[q];
  for (int m = 0; m < 3 * n; ++m) {
    const int e = T->edge[cfg][m];
    const int lower = kMcEdgeEnds[e][0];
    const int64_t node = surf_node(G, i + (lower & 1), j + ((lower >> 1) & 1), k + ((lower >> 2) & 1));
    out[m] = (int32_t)vid[3 * node + kMcEdgeAxis[e]];
  }
}

// per block: the minimum and maximum of each coordinate over its particles (grid-stride), and 1 in slot 6
// when any coordinate is NaN; 8 floats per block
__global__ void __launch_bounds__(256) k_surf_bounds(const float* __restrict__ x, int n, float* __restrict__ partial) {
  __shared__ float s_v[7][256];
  float v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0f};
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
    for (int a = 0; a < 3; ++a) {
      const float c = x[3 * p + a];
      v[a] = fminf(v[a], c);
      v[3 + a] = fmaxf(v[3 + a], c);
      if (isnan(c)) v[6] = 1.0f;
    }
  for (int q = 0; q < 7; ++q) s_v[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int half = 128; half > 0; half >>= 1) {
    if ((int)threadIdx.x < half) {
      const int o = threadIdx.x + half;
      for (int a = 0; a < 3; ++a) {
        s_v[a][threadIdx.x] = fminf(s_v[a][threadIdx.x], s_v[a][o]);
        s_v[3 + a][threadIdx.x] = fmaxf(s_v[3 + a][threadIdx.x], s_v[3 + a][o]);
      }
      s_v[6][threadIdx.x] = fmaxf(s_v[6][threadIdx.x], s_v[6][o]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 7) partial[8 * blockIdx.x + threadIdx.x] = s_v[threadIdx.x][0];
}

// adds the total of the preceding chunk of a chunked exclusive scan (its last input and output, which this
// launch does not touch) to every output of the current chunk
__global__ void __launch_bounds__(256) k_scan_carry(uint32_t* __restrict__ out, int64_t n,
                                                    const uint32_t* __restrict__ prev_in,
                                                    const uint32_t* __restrict__ prev_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += *prev_in + *prev_out;
}

}  // namespace mpm
