// Emission on gfx950: the lattice points inside a closed mesh (sample_mesh_volume, reference
// src/mpm_core/geometry.py:89-155, before its jitter). The lattice is (i + 0.5) spacing on every axis; a
// point is inside when the vertical ray through its (x, y) column, nudged by a fixed sub-spacing offset,
// crosses the mesh an odd number of times below it. Emission counts are part of a bake's result, so every
// quantity is computed in FP64 with exactly the operations of the reference (no contraction into fused
// multiply-adds), and the points come out in its order: column i * ny + j ascending, then z ascending.
//
//   k_emit_tri_columns   per triangle: the columns its nudged xy bounding box covers (0 for a triangle of
//                        zero xy area)
//   k_emit_pair_hits     per (triangle, column) pair: edge functions w0, w1, w2 and, where all three are
//                        >= 0, the crossing z; pass 0 counts crossings per column, pass 1 stores them
//   k_emit_column_points per column: its crossings sorted, then the lattice z with an odd count of
//                        crossings strictly below; pass 0 counts them, pass 1 writes the points
#pragma once

#include "kernels.hpp"

namespace mpm {

struct EmitLattice {
  int64_t first[3];   // lattice index of the first point per axis
  int n[3];           // points per axis
  double spacing;
  double nudge[2];    // ray offset in x and y
};

__device__ __forceinline__ double lattice_coord(const EmitLattice& L, int a, int64_t i) {
#pragma clang fp contract(off)
  return ((double)(L.first[a] + i) + 0.5) * L.spacing;
}

// the triangle's corners and its column range [i0, i0 + ni) x [j0, j0 + nj); false when it covers none
__device__ __forceinline__ bool tri_columns(const EmitLattice& L, const double* __restrict__ verts,
                                            const int32_t* __restrict__ tris, int t, double (&a)[3], double (&b)[3],
                                            double (&c)[3], double& area, int& i0, int& j0, int& ni, int& nj) {
#pragma clang fp contract(off)
  for (int k = 0; k < 3; ++k) {
    a[k] = verts[3 * (int64_t)tris[3 * t] + k];
    b[k] = verts[3 * (int64_t)tris[3 * t + 1] + k];
    c[k] = verts[3 * (int64_t)tris[3 * t + 2] + k];
  }
  area = (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]);
  const double lo0 = fmin(fmin(a[0], b[0]), c[0]), hi0 = fmax(fmax(a[0], b[0]), c[0]);
  const double lo1 = fmin(fmin(a[1], b[1]), c[1]), hi1 = fmax(fmax(a[1], b[1]), c[1]);
  const int64_t ci0 = max((int64_t)ceil((lo0 - L.nudge[0]) / L.spacing - 0.5) - L.first[0], (int64_t)0);
  const int64_t ci1 = min((int64_t)floor((hi0 - L.nudge[0]) / L.spacing - 0.5) - L.first[0], (int64_t)L.n[0] - 1);
  const int64_t cj0 = max((int64_t)ceil((lo1 - L.nudge[1]) / L.spacing - 0.5) - L.first[1], (int64_t)0);
  const int64_t cj1 = min((int64_t)floor((hi1 - L.nudge[1]) / L.spacing - 0.5) - L.first[1], (int64_t)L.n[1] - 1);
  if (ci1 < ci0 || cj1 < cj0 || area == 0.0) return false;
  i0 = (int)ci0;
  j0 = (int)cj0;
  ni = (int)(ci1 - ci0 + 1);
  nj = (int)(cj1 - cj0 + 1);
  return true;
}

__global__ void __launch_bounds__(256) k_emit_tri_columns(EmitLattice L, const double* __restrict__ verts,
                                                          const int32_t* __restrict__ tris, int n_tris,
                                                          uint32_t* __restrict__ pairs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tris) return;
  double a[3], b[3], c[3], area;
  int i0, j0, ni, nj;
  pairs[t] = tri_columns(L, verts, tris, t, a, b, c, area, i0, j0, ni, nj) ? (uint32_t)(ni * nj) : 0u;
}

// pass 0: col_count[col] += 1 per crossing; pass 1: z of each crossing into its column's run of hit_z
// (col_off: exclusive scan of the counts; col_fill: per-column cursors starting at zero)
__global__ void __launch_bounds__(256) k_emit_pair_hits(EmitLattice L, const double* __restrict__ verts,
                                                        const int32_t* __restrict__ tris, int n_tris,
                                                        const uint32_t* __restrict__ pair_off, uint32_t n_pairs,
                                                        int pass, uint32_t* __restrict__ col_count,
                                                        const uint32_t* __restrict__ col_off,
                                                        uint32_t* __restrict__ col_fill, double* __restrict__ hit_z) {
#pragma clang fp contract(off)
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pairs) return;
  // the triangle of pair p: the last one whose first pair is <= p (pair_off is the exclusive scan)
  int lo = 0, hi = n_tris - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pair_off[mid] <= p) lo = mid;
    else hi = mid - 1;
  }
  const int t = lo;
  double a[3], b[3], c[3], area;
  int i0, j0, ni, nj;
  if (!tri_columns(L, verts, tris, t, a, b, c, area, i0, j0, ni, nj)) return;
  const uint32_t local = p - pair_off[t];
  const int ii = i0 + (int)(local / (uint32_t)nj), jj = j0 + (int)(local % (uint32_t)nj);
  const double px = lattice_coord(L, 0, ii) + L.nudge[0];
  const double py = lattice_coord(L, 1, jj) + L.nudge[1];
  const double w0 = ((b[0] - px) * (c[1] - py) - (b[1] - py) * (c[0] - px)) / area;
  const double w1 = ((c[0] - px) * (a[1] - py) - (c[1] - py) * (a[0] - px)) / area;
  const double w2 = 1.0 - w0 - w1;
  if (!(w0 >= 0.0 && w1 >= 0.0 && w2 >= 0.0)) return;
  const int64_t col = (int64_t)ii * L.n[1] + jj;
  if (pass == 0) {
    atomicAdd(&col_count[col], 1u);
  } else {
    const double z = w0 * a[2] + w1 * b[2] + w2 * c[2];
    hit_z[col_off[col] + atomicAdd(&col_fill[col], 1u)] = z;
  }
}

// pass 0: pts_count[col] = inside lattice points of the column (its crossings sorted in place first);
// pass 1: their xyz rows from pts_off[col]
__global__ void __launch_bounds__(256) k_emit_column_points(EmitLattice L, int64_t n_cols,
                                                            const uint32_t* __restrict__ col_count,
                                                            const uint32_t* __restrict__ col_off,
                                                            double* __restrict__ hit_z, int pass,
                                                            uint32_t* __restrict__ pts_count,
                                                            const uint32_t* __restrict__ pts_off,
                                                            double* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n_cols) return;
  const uint32_t m = col_count[col];
  if (m == 0) {
    if (pass == 0) pts_count[col] = 0u;
    return;
  }
  double* z = hit_z + col_off[col];
  if (pass == 0) {
    for (uint32_t i = 1; i < m; ++i) {  // insertion sort: a column crosses a mesh a few times
      const double v = z[i];
      uint32_t j = i;
      while (j > 0 && z[j - 1] > v) {
        z[j] = z[j - 1];
        --j;
      }
      z[j] = v;
    }
  }
  const int ii = (int)(col / L.n[1]), jj = (int)(col % L.n[1]);
  const double x = lattice_coord(L, 0, ii), y = lattice_coord(L, 1, jj);
  uint32_t below = 0, inside = 0;
  double* dst = pass == 1 ? out + 3 * (int64_t)pts_off[col] : nullptr;
  for (int k = 0; k < L.n[2]; ++k) {
    const double zk = lattice_coord(L, 2, k);
    while (below < m && z[below] < zk) ++below;  // crossings strictly below this lattice z
    if (below & 1u) {
      if (pass == 1) {
        dst[3 * inside] = x;
        dst[3 * inside + 1] = y;
        dst[3 * inside + 2] = zk;
      }
      ++inside;
    }
  }
  if (pass == 0) pts_count[col] = inside;
}

}  // namespace mpm
