// Binned P2G and brick-sparse grid update on gfx950.
//
// P2G (kernels.py:92-140 semantics) walks the device-side list of occupied bricks (grid-stride, so the
// launch needs no host-side count). Particles were sorted by brick and cell, so each lane of a wave owns
// one of the brick's 64 cells and sums its particles' 27 stencil contributions in registers; the 64 lanes
// then add their sums into the wave's 8x8x8-node LDS window with plain read-add-writes (no two lanes of
// one instruction touch the same node). The window's 6^3 core goes to the brick's staging slot; the outer
// shell, which only particles that drifted since the sort reach, goes to the HBM accumulator with global
// atomics. A drifted particle is summed by the lane of its current cell when drift routing handed it
// there, else scattered with atomics by the lane that holds it.
//
// The grid update (kernels.py:155-180) gives each wave a strided set of bricks and updates the dirty ones,
// one lane per node: the node's P2G result is the sum of the staged cores that cover it plus its atomic
// part; momentum -> velocity, gravity, walls, then contact with the mesh colliders. It clears the
// accumulator and the dirty flag behind itself and records which bricks hold a fresh velocity.
#pragma once

#include "colliders.hpp"
#include "kernels.hpp"

namespace mpm {

constexpr int kWin = 8;
constexpr int kWinNodes = kWin * kWin * kWin;
// The staged part of a window: its core, window indices 1..6 per axis (nodes 4B .. 4B + 5 of brick B).
// A particle binned in brick B has its stencil base in the brick's cells 4B .. 4B + 3, so its 3 nodes
// per axis lie in the core; only drifted particles scattered by their holder reach the outer shell
// (indices 0 and 7), whose nonzero nodes P2G adds to the HBM accumulator at the end of the brick.
constexpr int kCore = 6;
constexpr int kStageNodes = kCore * kCore * kCore;

// Optional fixed-point P2G accumulator: 4 int64 per node (momentum xyz, mass) scaled by `scale`. Every
// contribution reaching it is rounded to an integer first and integer addition is associative, so the
// grid is bit-identical from run to run whatever order the atomics land in (the per-brick FP32 partial
// sums before it are formed in a fixed order). fx == nullptr selects the FP32 accumulator.
struct FixedAccum {
  unsigned long long* fx;
  double scale, inv_scale;
};

__device__ __forceinline__ void fixed_add(const FixedAccum& fa, int node, int c, float v) {
  atomicAdd(&fa.fx[4 * (int64_t)node + c], (unsigned long long)__double2ll_rn((double)v * fa.scale));
}

__device__ __forceinline__ void mark_brick(const GridGeom& g, uint32_t* dirty, int bx, int by, int bz) {
  if (bx < 0 || by < 0 || bz < 0 || bx >= g.bdims[0] || by >= g.bdims[1] || bz >= g.bdims[2]) return;
  dirty[(bx * g.bdims[1] + by) * g.bdims[2] + bz] = 1u;
}

// The affine momentum of one particle (kernels.py:126-131): A = -dt p_vol 4 / dx^2 tau + m C, with the
// stencil base, the per-axis weights and node offsets of its quadratic stencil.
struct P2GParticle {
  float A[3][3];
  float mv[3];
  float pm;
  int base[3];
  float wx[3][3], dp[3][3];
};

// the per-particle inputs of P2G as stored: x, v from their planes, C, stress and mass from the record
struct P2GRaw {
  float x[3], v[3];
  float4 r[4];
};

__device__ __forceinline__ void load_p2g_raw(const Particles& P, int64_t p, P2GRaw& r) {
  const int64_t cap = P.cap;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    r.x[a] = P.x[a * cap + p];
    r.v[a] = P.v[a * cap + p];
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) r.r[e] = P.rec[4 * p + e];
}

// the parts of a particle's P2G inputs that do not depend on where it lies in its stencil: the affine
// momentum matrix A, m v and m
__device__ __forceinline__ void p2g_affine(const P2GRaw& r, float k, float (&A)[3][3], float (&mv)[3], float& pm) {
  const float C[9] = {r.r[0].x, r.r[0].y, r.r[0].z, r.r[0].w, r.r[1].x, r.r[1].y, r.r[1].z, r.r[1].w, r.r[2].x};
  const float t00 = r.r[2].y, t11 = r.r[2].z, t22 = r.r[2].w, t01 = r.r[3].x, t02 = r.r[3].y, t12 = r.r[3].z;
  pm = r.r[3].w;
  A[0][0] = k * t00 + pm * C[0];
  A[0][1] = k * t01 + pm * C[1];
  A[0][2] = k * t02 + pm * C[2];
  A[1][0] = k * t01 + pm * C[3];
  A[1][1] = k * t11 + pm * C[4];
  A[1][2] = k * t12 + pm * C[5];
  A[2][0] = k * t02 + pm * C[6];
  A[2][1] = k * t12 + pm * C[7];
  A[2][2] = k * t22 + pm * C[8];
  mv[0] = pm * r.v[0];
  mv[1] = pm * r.v[1];
  mv[2] = pm * r.v[2];
}

// per-axis stencil weights and node offsets (x_node - x_p) of a particle at fractional offset fx
__device__ __forceinline__ void p2g_weights(const GridGeom& g, const float (&fx)[3], float (&wx)[3][3],
                                            float (&dp)[3][3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      wx[a][i] = bspline_w(fx[a], i);
      dp[a][i] = ((float)i - fx[a]) * g.dx;
    }
}

__device__ __forceinline__ void make_p2g_particle(const P2GRaw& r, const GridGeom& g, float k, P2GParticle& q) {
  p2g_affine(r, k, q.A, q.mv, q.pm);
  float fx[3];
  stencil_base(g, r.x, q.base, fx);
  p2g_weights(g, fx, q.wx, q.dp);
}

// a particle's 27 stencil contributions (momentum with its affine part, and mass) added to a lane's sums
__device__ __forceinline__ void p2g_accumulate(float (&acc)[3][3][3][4], const float (&A)[3][3], const float (&mv)[3],
                                               float pm, const float (&wx)[3][3], const float (&dp)[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float wij = wx[0][i] * wx[1][j];
      const float ax = mv[0] + A[0][0] * dp[0][i] + A[0][1] * dp[1][j];
      const float ay = mv[1] + A[1][0] * dp[0][i] + A[1][1] * dp[1][j];
      const float az = mv[2] + A[2][0] * dp[0][i] + A[2][1] * dp[1][j];
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) {
        const float wgt = wij * wx[2][kk];
        acc[i][j][kk][0] += wgt * (ax + A[0][2] * dp[2][kk]);
        acc[i][j][kk][1] += wgt * (ay + A[1][2] * dp[2][kk]);
        acc[i][j][kk][2] += wgt * (az + A[2][2] * dp[2][kk]);
        acc[i][j][kk][3] += wgt * pm;
      }
    }
  }
}

// Scatter of a particle whose stencil is not its sorted cell's (it drifted since the last sort): through
// atomics into the wave's window when the stencil fits the window, else straight to HBM. Returns whether
// it wrote to the window's outer shell.
__device__ __forceinline__ bool p2g_scatter_atomic(const P2GParticle& q, const GridGeom& g, const int (&wo)[3],
                                                float (*win)[kWinNodes], float4* __restrict__ accum,
                                                uint32_t* __restrict__ dirty, const FixedAccum& fa) {
  const int l0 = q.base[0] - wo[0], l1 = q.base[1] - wo[1], l2 = q.base[2] - wo[2];
  // fixed point: straight to the integer grid (the window's float atomics would not add in a fixed order)
  const bool in_window = fa.fx == nullptr && l0 >= 0 && l1 >= 0 && l2 >= 0 && l0 <= kWin - 3 &&
                         l1 <= kWin - 3 && l2 <= kWin - 3;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const float wij = q.wx[0][i] * q.wx[1][j];
      for (int kk = 0; kk < 3; ++kk) {
        const float w = wij * q.wx[2][kk];
        const float d0 = q.dp[0][i], d1 = q.dp[1][j], d2 = q.dp[2][kk];
        const float vx = w * (q.mv[0] + (q.A[0][0] * d0 + q.A[0][1] * d1 + q.A[0][2] * d2));
        const float vy = w * (q.mv[1] + (q.A[1][0] * d0 + q.A[1][1] * d1 + q.A[1][2] * d2));
        const float vz = w * (q.mv[2] + (q.A[2][0] * d0 + q.A[2][1] * d1 + q.A[2][2] * d2));
        const float vm = w * q.pm;
        if (in_window) {
          const int li = ((l0 + i) * kWin + (l1 + j)) * kWin + (l2 + kk);
          atomicAdd(&win[0][li], vx);
          atomicAdd(&win[1][li], vy);
          atomicAdd(&win[2][li], vz);
          atomicAdd(&win[3][li], vm);
        } else if (fa.fx) {
          const int node = node_addr(g, q.base[0] + i, q.base[1] + j, q.base[2] + kk);
          fixed_add(fa, node, 0, vx);
          fixed_add(fa, node, 1, vy);
          fixed_add(fa, node, 2, vz);
          fixed_add(fa, node, 3, vm);
        } else {
          float4* node = accum + node_addr(g, q.base[0] + i, q.base[1] + j, q.base[2] + kk);
          atomicAdd(&node->x, vx);
          atomicAdd(&node->y, vy);
          atomicAdd(&node->z, vz);
          atomicAdd(&node->w, vm);
        }
      }
    }
  if (!in_window) {
    // the stencil spans at most two bricks per axis
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        for (int kk = 0; kk < 2; ++kk)
          mark_brick(g, dirty, (q.base[0] + 2 * i) >> 2, (q.base[1] + 2 * j) >> 2, (q.base[2] + 2 * kk) >> 2);
    return false;
  }
  return l0 == 0 || l1 == 0 || l2 == 0 || l0 == kWin - 3 || l1 == kWin - 3 || l2 == kWin - 3;
}

// ---- drift routing --------------------------------------------------------------------------------
// Between two sorts a particle stays in the slot of the cell it was sorted into, but its stencil follows
// its position. Before P2G, k_drift hands every particle whose cell changed ("drifted") to the incoming
// list of its current cell (kInPerCell entries per cell of an occupied brick, indexed by the brick's
// slot), so the P2G lane that owns that cell sums it in registers like its own particles; `fwd` marks the
// forwarded ones for the lane that holds them in sorted order, which then skips them. A drifted particle
// whose cell is in an unoccupied brick, or whose cell's list is full, is left to the holder's atomic
// scatter. P2G clears the counters it reads, so they start every substep at zero.
constexpr int kInPerCell = 4;

struct DriftLists {
  uint32_t* cnt;  // per slot cell: particles routed to it (may exceed kInPerCell); null: no routing
  uint32_t* lst;  // per slot cell: kInPerCell particle indices
  const uint8_t* fwd;  // per particle: routed to its current cell's list
};

// The routing of particles p0 and p1 (p1 < 0: p0 alone) against the cell keys of the last sort (`keys`, in
// slot order). Each particle's accesses form a chain (position and old key, brick slot, list counter, list
// entry); the two chains are issued side by side, so that they cost one chain's latency. The slot is read
// whether or not the cell changed: every cell key lies inside the grid.
__device__ __forceinline__ void drift_route_pair(const Particles& P, const GridGeom& g, int p0, int p1,
                                                 const uint32_t* __restrict__ keys,
                                                 const int32_t* __restrict__ slot, uint32_t* __restrict__ in_cnt,
                                                 uint32_t* __restrict__ in_lst, uint8_t* __restrict__ fwd) {
  const int pp[2] = {p0, p1 >= 0 ? p1 : p0};
  float x[2][3];
  uint32_t old[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    x[j][0] = P.x[pp[j]];
    x[j][1] = P.x[P.cap + pp[j]];
    x[j][2] = P.x[2 * P.cap + pp[j]];
    old[j] = keys[pp[j]];
  }
  uint32_t key[2];
  int32_t s[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    int base[3];
    float fx[3];
    stencil_base(g, x[j], base, fx);
    key[j] = cell_key(g, base);
    s[j] = slot[key[j] >> 6];
  }
  // the slot cell each particle drifted to (~0u: none), both settled before either counter update is
  // issued: memory counters complete in issue order, so a wait for the second slot after the first update
  // would wait for that update's return too
  uint32_t cell[2];
#pragma unroll
  for (int j = 0; j < 2; ++j)
    cell[j] = key[j] != old[j] && s[j] >= 0 && (j == 0 || p1 >= 0) ? (uint32_t)s[j] * 64u + (key[j] & 63u) : ~0u;
  asm volatile("" ::"v"(cell[0]), "v"(cell[1]));
  uint32_t pos[2] = {~0u, ~0u};
#pragma unroll
  for (int j = 0; j < 2; ++j)
    if (cell[j] != ~0u) pos[j] = atomicAdd(&in_cnt[cell[j]], 1u);
#pragma unroll
  for (int j = 0; j < 2; ++j)
    if (pos[j] < (uint32_t)kInPerCell) in_lst[cell[j] * kInPerCell + pos[j]] = (uint32_t)pp[j];
  fwd[p0] = pos[0] < (uint32_t)kInPerCell ? 1 : 0;
  if (p1 >= 0) fwd[p1] = pos[1] < (uint32_t)kInPerCell ? 1 : 0;
}

// two particles per thread, a launch's thread count apart (launched over (n + 1) / 2 threads)
__global__ void __launch_bounds__(256) k_drift(Particles P, GridGeom g, int n, const uint32_t* __restrict__ keys,
                                               const int32_t* __restrict__ slot, uint32_t* __restrict__ in_cnt,
                                               uint32_t* __restrict__ in_lst, uint8_t* __restrict__ fwd) {
  const int stride = (int)(gridDim.x * blockDim.x);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += 2 * stride)
    drift_route_pair(P, g, p, p + stride < n ? p + stride : -1, keys, slot, in_cnt, in_lst, fwd);
}

// A value of another lane of the same 16-lane row (DPP row shift; kRowShr4: lane l - 4, kRowShl8: l + 8, ...);
// zero when that lane lies outside the row. The whole wave must be active.
constexpr int kRowShl4 = 0x104, kRowShl8 = 0x108, kRowShl12 = 0x10c, kRowShr4 = 0x114, kRowShr8 = 0x118;
template <int kCtrl>
__device__ __forceinline__ float row_shift(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xf, 0xf, true));
}

// The drift routing of the next substep done by the kernel after G2P (k_fp64_fixup): keys == null for none.
struct DriftAhead {
  GridGeom g;
  int n;
  const uint32_t* keys;
  const int32_t* slot;
  uint32_t *in_cnt, *in_lst;
  uint8_t* fwd;
};

// orders this wave's LDS accesses across lanes (a wave's LDS operations complete in issue order)
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// P2G of occupied brick b by one wave, one lane per cell of the brick. Particles are sorted by cell, so
// lane l walks the particles of cell l and sums their 27 stencil contributions in registers. The 64 lanes
// of a wave own 64 different cells, so when they then add their sums into the wave's private LDS window
// every store of one instruction hits a different node: plain read-add-write, no LDS atomics. The window
// is flushed to the bricked HBM grid with global atomics (neighbouring windows overlap).
// kSplit waves per brick (1, 2 or 4; above 1 the caller is a workgroup of kSplit waves on one brick, and
// `half` is the wave): wave h takes the particles h, h + kSplit, ... of every cell's list into its own
// window wins[h], and the workgroup stages the sum of the windows, added in wave order.
template <int kSplit>
__device__ __forceinline__ void p2g_brick(const Particles& P, const GridGeom& g, float k,
                                          const uint32_t* __restrict__ cell_off, uint32_t b,
                                          float (*wins)[4][kWinNodes], float4* __restrict__ accum,
                                          uint32_t* __restrict__ dirty, int lane, const FixedAccum& fa,
                                          float4* __restrict__ stage, uint32_t slot, const DriftLists& dl,
                                          int half = 0) {
  static_assert(kSplit == 1 || kSplit == 2 || kSplit == 4, "1, 2 or 4 waves per brick");
  float (*const win)[kWinNodes] = wins[half];
  const int cx = lane >> 4, cy = (lane >> 2) & 3, cz = lane & 3;
  const int bz = (int)(b % (uint32_t)g.bdims[2]);
  const int by = (int)((b / (uint32_t)g.bdims[2]) % (uint32_t)g.bdims[1]);
  const int bx = (int)(b / ((uint32_t)g.bdims[2] * (uint32_t)g.bdims[1]));
  const int wo[3] = {4 * bx - 1, 4 * by - 1, 4 * bz - 1};
  for (int i = lane; i < kWinNodes; i += 64) {
    win[0][i] = 0.0f;
    win[1][i] = 0.0f;
    win[2][i] = 0.0f;
    win[3][i] = 0.0f;
  }
  wave_lds_fence();
  const int cb[3] = {4 * bx + cx, 4 * by + cy, 4 * bz + cz};
  const uint32_t p0 = cell_off[64 * b + lane], p1 = cell_off[64 * b + lane + 1];
  // the lane's particles: its cell's sorted ones, then those routed to its cell since the last sort
  const uint32_t own_n = p1 - p0;
  uint32_t n_in = 0;
  uint32_t* in_list = nullptr;
  if (dl.cnt) {
    const uint32_t c = slot * 64u + (uint32_t)lane;
    const uint32_t m = dl.cnt[c];
    n_in = m < (uint32_t)kInPerCell ? m : (uint32_t)kInPerCell;
    in_list = dl.lst + (size_t)c * kInPerCell;
    if (kSplit == 1) dl.cnt[c] = 0u;  // two waves: cleared once both have read it (below)
    if (n_in > 1 && half == 0) {
      // in particle order, so the lane's sums do not depend on the order of the routing atomics
      static_assert(kInPerCell == 4, "sorting network for 4 entries");
      uint32_t e[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) e[j] = (uint32_t)j < n_in ? in_list[j] : 0xffffffffu;
      const auto order = [](uint32_t& a, uint32_t& b) {
        const uint32_t lo = a < b ? a : b;
        b = a < b ? b : a;
        a = lo;
      };
      order(e[0], e[1]);
      order(e[2], e[3]);
      order(e[0], e[2]);
      order(e[1], e[3]);
      order(e[1], e[2]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if ((uint32_t)j < n_in) in_list[j] = e[j];
    }
  }
  if (kSplit > 1) {
    __threadfence_block();
    __syncthreads();  // wave 0's sorted lists are visible to the others; all have read the counts
    if (dl.cnt && half == 0) dl.cnt[slot * 64u + (uint32_t)lane] = 0u;
  }
  const uint32_t total = own_n + n_in;
  // The routed particles' indices are held in registers, so no particle's loads wait on an index load
  // inside the loop below (a wait for one load is a wait for every load issued before it, the prefetch
  // of the next particle included).
  uint32_t inl[kInPerCell];
#pragma unroll
  for (int j = 0; j < kInPerCell; ++j) inl[j] = (uint32_t)j < n_in ? in_list[j] : 0u;
  const auto index_of = [&](uint32_t t) -> uint32_t {
    const uint32_t j = t - own_n;
    return t < own_n ? p0 + t : (j == 0 ? inl[0] : j == 1 ? inl[1] : j == 2 ? inl[2] : inl[3]);
  };
  float acc[3][3][3][4];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int kk = 0; kk < 3; ++kk)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[i][j][kk][c] = 0.0f;
  bool shell = false;  // this lane scattered into the window's outer shell
  // The next particle's inputs load while the current one is scattered. Its drift flag is loaded as a raw
  // byte and tested only once it is the current particle: testing it at once would wait for the whole
  // prefetch.
  P2GRaw next;
  uint32_t next_fwd = 0u;
  const uint32_t t0 = (uint32_t)half;
  if (t0 < total) {
    const uint32_t pn = index_of(t0);
    load_p2g_raw(P, (int64_t)pn, next);
    if (dl.cnt) next_fwd = dl.fwd[pn];
  }
  for (uint32_t t = t0; t < total; t += kSplit) {
    const P2GRaw cur = next;
    const bool cur_skip = t < own_n && next_fwd != 0u;  // routed to its current cell's lane
    const uint32_t tn = t + kSplit;
    if (tn < total) {
      const uint32_t pn = index_of(tn);
      load_p2g_raw(P, (int64_t)pn, next);
      if (dl.cnt) next_fwd = dl.fwd[pn];
    }
    P2GParticle q;
    make_p2g_particle(cur, g, k, q);
    if (q.base[0] == cb[0] && q.base[1] == cb[1] && q.base[2] == cb[2]) {
      p2g_accumulate(acc, q.A, q.mv, q.pm, q.wx, q.dp);
    } else if (!cur_skip) {
      shell |= p2g_scatter_atomic(q, g, wo, win, accum, dirty, fa);
    }
  }
  wave_lds_fence();
  // The lane sums reach the window in two steps. Along y they are added across lanes first: lane
  // 16 cx + 4 cy + cz, so a 16-lane row is one cx and a row shift by 4 lanes moves one cell along y (a
  // lane whose source lies outside its row reads zero). For each (i, k), window node y index cy + 1
  // collects acc[i][0][k] of cell cy, acc[i][1][k] of cell cy - 1 and acc[i][2][k] of cell cy - 2; node
  // cy + 5 (cy < 2) collects acc[i][1][k] of cell cy + 3 and acc[i][2][k] of cell cy + 2. Those nine (i, k)
  // sums are then added into the window with read-add-writes: within one (i, k) no two lanes touch the same
  // node; across them they may, so the fence after each keeps the nine in program order (a wave's LDS
  // operations complete in issue order). Every lane takes part (an empty cell's lane still carries its
  // neighbours' sums).
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
      float lo[4], hi[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float a0 = acc[i][0][kk][c], a1 = acc[i][1][kk][c], a2 = acc[i][2][kk][c];
        lo[c] = (a0 + row_shift<kRowShr4>(a1)) + row_shift<kRowShr8>(a2);
        hi[c] = row_shift<kRowShl12>(a1) + row_shift<kRowShl8>(a2);
      }
      const int li = ((cx + 1 + i) * kWin + (cy + 1)) * kWin + (cz + 1 + kk);
#pragma unroll
      for (int c = 0; c < 4; ++c) win[c][li] += lo[c];
      if (cy < 2) {
#pragma unroll
        for (int c = 0; c < 4; ++c) win[c][li + 4 * kWin] += hi[c];
      }
      wave_lds_fence();
    }
  // the brick's window: this wave's, or the sum of the workgroup's (each thread then takes every
  // (64 kSplit)-th node)
  // Whether any lane of the brick's workgroup wrote the outer shell (uniform over the workgroup). With
  // several waves the vote is also the barrier after which every wave's window is complete (each wave's
  // LDS writes are made visible to the workgroup first): one barrier instead of two (1 us per launch).
  if (kSplit > 1) __threadfence_block();
  const bool any_shell = kSplit > 1 ? __syncthreads_or(shell) != 0 : __ballot(shell) != 0;
  const int tid = half * 64 + lane;
  const auto node_value = [&](int i, int c) {
    float v = wins[0][c][i];
#pragma unroll
    for (int h = 1; h < kSplit; ++h) v += wins[h][c][i];
    return v;
  };
  if (stage && !fa.fx) {
    // the window's core goes to the brick's staging slot with plain stores; the grid update sums the (1 to
    // 8) cores that cover each node, in a fixed order
    float4* out = stage + (int64_t)slot * kStageNodes;
    for (int c = tid; c < kStageNodes; c += 64 * kSplit) {
      const int x = c / (kCore * kCore), y = (c / kCore) % kCore, z = c % kCore;
      const int i = ((x + 1) * kWin + (y + 1)) * kWin + (z + 1);
      out[c] = make_float4(node_value(i, 0), node_value(i, 1), node_value(i, 2), node_value(i, 3));
    }
    // the outer shell holds only what drifted particles scattered there: its nonzero nodes go to the HBM
    // accumulator (four atomics per node, with the node's brick marked) instead of being staged
    for (int i = tid; any_shell && i < kWinNodes; i += 64 * kSplit) {
      const int x = i >> 6, y = (i >> 3) & 7, z = i & 7;
      if (x != 0 && x != kWin - 1 && y != 0 && y != kWin - 1 && z != 0 && z != kWin - 1) continue;
      const float vx = node_value(i, 0), vy = node_value(i, 1), vz = node_value(i, 2), vm = node_value(i, 3);
      if (vx == 0.0f && vy == 0.0f && vz == 0.0f && vm == 0.0f) continue;
      const int gi = wo[0] + x, gj = wo[1] + y, gk = wo[2] + z;
      if (gi < 0 || gj < 0 || gk < 0 || gi >= g.dims[0] || gj >= g.dims[1] || gk >= g.dims[2]) continue;
      float4* node = accum + node_addr(g, gi, gj, gk);
      if (vx != 0.0f) atomicAdd(&node->x, vx);
      if (vy != 0.0f) atomicAdd(&node->y, vy);
      if (vz != 0.0f) atomicAdd(&node->z, vz);
      if (vm != 0.0f) atomicAdd(&node->w, vm);
      mark_brick(g, dirty, gi >> 2, gj >> 2, gk >> 2);
    }
  } else for (int i = tid; i < kWinNodes; i += 64 * kSplit) {
    const int x = i >> 6, y = (i >> 3) & 7, z = i & 7;
    const int gi = wo[0] + x, gj = wo[1] + y, gk = wo[2] + z;
    if (gi < 0 || gj < 0 || gk < 0 || gi >= g.dims[0] || gj >= g.dims[1] || gk >= g.dims[2]) continue;
    const float vx = node_value(i, 0), vy = node_value(i, 1), vz = node_value(i, 2), vm = node_value(i, 3);
    if (vx == 0.0f && vy == 0.0f && vz == 0.0f && vm == 0.0f) continue;
    // a shell node may lie in the brick below (the core's bricks are marked below for every node)
    if (x == 0 || x == kWin - 1 || y == 0 || y == kWin - 1 || z == 0 || z == kWin - 1)
      mark_brick(g, dirty, gi >> 2, gj >> 2, gk >> 2);
    if (fa.fx) {
      const int nd = node_addr(g, gi, gj, gk);
      if (vx != 0.0f) fixed_add(fa, nd, 0, vx);
      if (vy != 0.0f) fixed_add(fa, nd, 1, vy);
      if (vz != 0.0f) fixed_add(fa, nd, 2, vz);
      if (vm != 0.0f) fixed_add(fa, nd, 3, vm);
      continue;
    }
    float4* node = accum + node_addr(g, gi, gj, gk);
    if (vx != 0.0f) atomicAdd(&node->x, vx);
    if (vy != 0.0f) atomicAdd(&node->y, vy);
    if (vz != 0.0f) atomicAdd(&node->z, vz);
    if (vm != 0.0f) atomicAdd(&node->w, vm);
  }
  // the core spans this brick and the next one along each axis
  if (half == 0 && lane < 8) mark_brick(g, dirty, bx + (lane >> 2), by + ((lane >> 1) & 1), bz + (lane & 1));
  if (kSplit > 1) __syncthreads();  // every window read before the next brick clears them
  else wave_lds_fence();
}

// The CFL step of an advance (solver.py:176-185) from the reduced speed, in FP64 with the host's operations
// in the host's order (no contraction), so it equals the dt the host derives from the same speed: the
// advance's first P2G derives it itself and can start before the host has read the speed.
struct CflStep {
  const float* vmax;  // the reduced speed (device copy); null: use the dt argument
  double wave, dx, frame_dt;
};

__device__ __forceinline__ float cfl_dt(const CflStep& c) {
#pragma clang fp contract(off)
  double vmax = (double)*c.vmax;
  if (!isfinite(vmax)) vmax = 0.0;
  const double a = 0.4 * c.dx / (c.wave + vmax);
  const double cfl = c.frame_dt < a ? c.frame_dt : a;
  const double steps = ceil(c.frame_dt / cfl - 1e-9);
  return (float)(c.frame_dt / steps);
}

// P2G over the device-side list of occupied bricks: kWaves waves (bricks at a time) per workgroup,
// grid-stride, so the launch needs no host-side count
template <int kWaves>
__global__ void __launch_bounds__(64 * kWaves) k_p2g_cells(Particles P, GridGeom g, float p_vol, float dt,
                                                   const uint32_t* __restrict__ cell_off,
                                                   const uint32_t* __restrict__ active,
                                                   const uint32_t* __restrict__ n_active_ptr,
                                                   float4* __restrict__ accum, uint32_t* __restrict__ dirty,
                                                   FixedAccum fa, float4* __restrict__ stage, DriftLists dl,
                                                   CflStep cfl) {
  __shared__ float s_win[kWaves][4][kWinNodes];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t n_active = *n_active_ptr;
  if (cfl.vmax) dt = cfl_dt(cfl);
  const float k = -dt * p_vol * 4.0f * g.inv_dx * g.inv_dx;
  for (uint32_t w = blockIdx.x * kWaves + wave; w < n_active; w += gridDim.x * kWaves)
    p2g_brick<1>(P, g, k, cell_off, active[w], &s_win[wave], accum, dirty, lane, fa, stage, w, dl);
}

// kSplit waves per brick (one workgroup of 64 kSplit per brick at a time, grid-stride over the brick
// list): each brick's chain of particle visits is 1/kSplit as long and kSplit times as many waves are in
// flight, for scenes whose occupied bricks are too few to cover the chains' latency
template <int kSplit>
__global__ void __launch_bounds__(64 * kSplit) k_p2g_split(Particles P, GridGeom g, float p_vol, float dt,
                                                           const uint32_t* __restrict__ cell_off,
                                                           const uint32_t* __restrict__ active,
                                                           const uint32_t* __restrict__ n_active_ptr,
                                                           float4* __restrict__ accum, uint32_t* __restrict__ dirty,
                                                           FixedAccum fa, float4* __restrict__ stage, DriftLists dl,
                                                           CflStep cfl) {
  __shared__ float s_win[kSplit][4][kWinNodes];
  const int half = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t n_active = *n_active_ptr;
  if (cfl.vmax) dt = cfl_dt(cfl);  // the advance's first P2G: its CFL step derived here
  const float k = -dt * p_vol * 4.0f * g.inv_dx * g.inv_dx;
  for (uint32_t w = blockIdx.x; w < n_active; w += gridDim.x)
    p2g_brick<kSplit>(P, g, k, cell_off, active[w], s_win, accum, dirty, lane, fa, stage, w, dl, half);
}

// grid update of one dirty brick by one wave (lane = node): momentum -> velocity, gravity, walls,
// collider contact; clears the accumulator behind itself
// P2G result of node (i, j, k): what reached the HBM accumulator (drifted particles) plus, when windows
// are staged, the staged cores of the occupied bricks that cover the node. Core B' spans nodes 4B' .. 4B' + 5
// per axis, so node 4B + r is covered by bricks B - 1 and B (r <= 1) or by B alone (r >= 2); the sum runs
// over bricks in increasing order.
__device__ __forceinline__ float4 staged_sum(const GridGeom& g, int i, int j, int k, float4 a,
                                             const float4* __restrict__ stage, const int32_t* __restrict__ slot) {
  const int n[3] = {i, j, k};
  int b0[3], nb[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const bool two = (n[ax] & 3) <= 1;
    b0[ax] = (n[ax] >> 2) - (two ? 1 : 0);
    nb[ax] = two ? 2 : 1;
  }
  for (int dx = 0; dx < nb[0]; ++dx)
    for (int dy = 0; dy < nb[1]; ++dy)
      for (int dz = 0; dz < nb[2]; ++dz) {
        const int bx = b0[0] + dx, by = b0[1] + dy, bz = b0[2] + dz;
        if (bx < 0 || by < 0 || bz < 0 || bx >= g.bdims[0] || by >= g.bdims[1] || bz >= g.bdims[2]) continue;
        const int32_t s = slot[(bx * g.bdims[1] + by) * g.bdims[2] + bz];
        if (s < 0) continue;
        const int li = ((i - 4 * bx) * kCore + (j - 4 * by)) * kCore + (k - 4 * bz);
        const float4 w = stage[(int64_t)s * kStageNodes + li];
        a.x += w.x;
        a.y += w.y;
        a.z += w.z;
        a.w += w.w;
      }
  return a;
}

// staged_sum for every node of one brick (bx, by, bz), as the grid update needs it: the cores that can cover
// a node of the brick are those of the 8 bricks {bx - 1, bx} x {by - 1, by} x {bz - 1, bz}, the same for every
// node, so their slots are loaded once per wave (wave-uniform) and a node's (up to 8) stage loads are issued
// together instead of one slot-then-stage pair after another. Node 4B + r needs brick B - 1 along an axis
// only when r <= 1. The cores are added in the order staged_sum adds them (bricks in increasing order, x
// outermost), so the sums are the same.
__device__ __forceinline__ float4 staged_sum_brick(const GridGeom& g, int bx, int by, int bz, int i, int j, int k,
                                                   float4 a, const float4* __restrict__ stage,
                                                   const int32_t* __restrict__ slot) {
  int32_t s[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int cx = bx - 1 + (q >> 2), cy = by - 1 + ((q >> 1) & 1), cz = bz - 1 + (q & 1);
    const bool in = cx >= 0 && cy >= 0 && cz >= 0 && cx < g.bdims[0] && cy < g.bdims[1] && cz < g.bdims[2];
    s[q] = in ? slot[(cx * g.bdims[1] + cy) * g.bdims[2] + cz] : -1;
  }
  const bool lo[3] = {i - 4 * bx <= 1, j - 4 * by <= 1, k - 4 * bz <= 1};
  float4 w[8];
  bool use[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int dx = q >> 2, dy = (q >> 1) & 1, dz = q & 1;
    use[q] = s[q] >= 0 && (dx == 1 || lo[0]) && (dy == 1 || lo[1]) && (dz == 1 || lo[2]);
    const int cx = bx - 1 + dx, cy = by - 1 + dy, cz = bz - 1 + dz;
    const int li = ((i - 4 * cx) * kCore + (j - 4 * cy)) * kCore + (k - 4 * cz);
    w[q] = use[q] ? stage[(int64_t)s[q] * kStageNodes + li] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (use[q]) {
      a.x += w[q].x;
      a.y += w[q].y;
      a.z += w[q].z;
      a.w += w[q].w;
    }
  return a;
}

// the P2G result of every node (accumulator plus staged windows) as float4, for the grid download
__global__ void __launch_bounds__(256) k_p2g_resolve(GridGeom g, const float4* __restrict__ accum, int n_nodes,
                                                     const float4* __restrict__ stage,
                                                     const int32_t* __restrict__ slot, float4* __restrict__ out) {
  const int addr = blockIdx.x * blockDim.x + threadIdx.x;
  if (addr >= n_nodes) return;
  if (!stage) {  // windows flushed with atomics: the accumulator holds everything
    out[addr] = accum[addr];
    return;
  }
  int i, j, k;
  node_ijk(g, addr, i, j, k);
  out[addr] = staged_sum(g, i, j, k, accum[addr], stage, slot);
}

__device__ __forceinline__ void grid_brick(const GridGeom& g, const StepConsts& sc, float dt, int brick,
                                           float4* __restrict__ accum, float4* __restrict__ vel,
                                           const ColliderSet& cs, const ColliderFrame* frames, int lane,
                                           const FixedAccum& fa, const float4* __restrict__ stage,
                                           const int32_t* __restrict__ slot) {
  brick = __builtin_amdgcn_readfirstlane(brick);  // one brick per wave
  const int addr = brick * 64 + lane;
  int i, j, k;
  node_ijk(g, addr, i, j, k);
  float4 a;
  if (fa.fx) {
    unsigned long long* f = fa.fx + 4 * (int64_t)addr;
    a = make_float4((float)((double)(long long)f[0] * fa.inv_scale), (float)((double)(long long)f[1] * fa.inv_scale),
                    (float)((double)(long long)f[2] * fa.inv_scale), (float)((double)(long long)f[3] * fa.inv_scale));
    f[0] = f[1] = f[2] = f[3] = 0ull;
  } else {
    a = accum[addr];
  }
  const float4 drift = a;
  if (stage && !fa.fx) {
    const int bz = brick % g.bdims[2], by = (brick / g.bdims[2]) % g.bdims[1], bx = brick / (g.bdims[2] * g.bdims[1]);
    a = staged_sum_brick(g, bx, by, bz, i, j, k, a, stage, slot);
  }
  float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (i < g.dims[0] && j < g.dims[1] && k < g.dims[2]) {
    out = update_node(g, sc, dt, i, j, k, a);
    if (cs.count > 0 && out.w > 0.0f) collide_node_frames(cs, frames, g, i, j, k, out);
  }
  vel[addr] = out;
  if (drift.x != 0.0f || drift.y != 0.0f || drift.z != 0.0f || drift.w != 0.0f)
    accum[addr] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// one wave per brick; non-dirty bricks are skipped (and flagged stale for the grid download)
constexpr int kGridBricksPerWave = 8;

__global__ void __launch_bounds__(256) k_grid_bricks(GridGeom g, StepConsts sc, float dt, int n_bricks,
                                                     uint32_t* __restrict__ dirty, uint8_t* __restrict__ live,
                                                     float4* __restrict__ accum, float4* __restrict__ vel,
                                                     ColliderSet cs, ColliderFrames fr, FixedAccum fa,
                                                     const float4* __restrict__ stage,
                                                     const int32_t* __restrict__ slot,
                                                     uint32_t* __restrict__ n_fix) {
  const ColliderFrame* frames = fr.f;
  cs = with_records(cs, fr);
  // the FP64 work list of the G2P that follows starts empty
  if (n_fix && blockIdx.x == 0 && threadIdx.x == 0) *n_fix = 0u;
  // wave w examines the kGridBricksPerWave bricks w, w + W, w + 2W, ... (W waves in the launch): a few
  // thousand waves instead of one per brick of the grid, most of which would find their brick clean, and
  // the dirty bricks (contiguous bodies in brick order) still spread over the waves
  const int waves = (int)(gridDim.x * (blockDim.x >> 6));
  const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
  const int lane = threadIdx.x & 63;
  const int mine = w + lane * waves;
  bool d = false;
  if (lane < kGridBricksPerWave && mine < n_bricks) {
    d = dirty[mine] != 0u;
    live[mine] = d ? 1 : 0;
  }
  uint64_t todo = __ballot(d);
  while (todo != 0ull) {
    const int l = __builtin_ctzll(todo);
    todo &= todo - 1ull;
    const int brick = w + l * waves;
    grid_brick(g, sc, dt, brick, accum, vel, cs, frames, lane, fa, stage, slot);
    if (lane == 0) dirty[brick] = 0u;
  }
}

// dense variant for the reference-order P2G (which flags no bricks): every node of the grid
__global__ void __launch_bounds__(256) k_grid_dense(GridGeom g, StepConsts sc, float dt, int n_nodes,
                                                    float4* __restrict__ accum, float4* __restrict__ vel,
                                                    uint8_t* __restrict__ live, ColliderSet cs, ColliderFrames fr,
                                                    uint32_t* __restrict__ n_fix) {
  const ColliderFrame* frames = fr.f;
  cs = with_records(cs, fr);
  if (n_fix && blockIdx.x == 0 && threadIdx.x == 0) *n_fix = 0u;
  const int addr = blockIdx.x * blockDim.x + threadIdx.x;
  if (addr >= n_nodes) return;
  if ((addr & 63) == 0) live[addr >> 6] = 1;
  const float4 a = accum[addr];
  int i, j, k;
  node_ijk(g, addr, i, j, k);
  float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (i < g.dims[0] && j < g.dims[1] && k < g.dims[2]) {
    out = update_node(g, sc, dt, i, j, k, a);
    if (cs.count > 0 && out.w > 0.0f) collide_node_frames(cs, frames, g, i, j, k, out);
  }
  vel[addr] = out;
  if (a.x != 0.0f || a.y != 0.0f || a.z != 0.0f || a.w != 0.0f) accum[addr] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ---------------------------------------------------------------------------------------------
// G2P (kernels.py:183-263): gather v and C, advect with wall projection and finite-only clamp,
// F <- (I + dt C) F, per-model return map. One SVD of the new F per particle also gives the Kirchhoff
// stress that the next P2G scatters (stored as 6 symmetric entries), so P2G needs no SVD. The
// largest |v| of each workgroup goes to block_max for the next CFL step.
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Per-model return map of the new deformation gradient Fn = U diag(sig) V^T (kernels.py:242-263, 63-89)
// and the stress of the result, in precision T; F, Jp and tau are stored in FP32.
template <class T>
__device__ __forceinline__ void material_update(const MatDev& m, const M3<T>& Fn, const M3<T>& U, const T (&sig)[3],
                                                const M3<T>& V, Particles& P, float (&tau_out)[6], int64_t p) {
  const int64_t cap = P.cap;
  T t[6];
  if (m.model == kElastic) {
#pragma unroll
    for (int e = 0; e < 9; ++e) P.F[e * cap + p] = (float)Fn.a[e / 3][e % 3];
    stress_from_svd(m, U, sig, T(1), t);
  } else {
    T ns[3];
    T jp = T(1);
    if (m.model == kVonMises) {
      project_von_mises(sig, T(m.mu), T(m.yield_stress), ns);
    } else if (m.model == kDruckerPrager) {
      project_drucker_prager(sig, T(m.mu), T(m.lam), T(m.alpha), ns);
    } else {  // snow (Stomakhin et al. 2013)
      const T lo = T(1) - T(m.crit_c), hi = T(1) + T(m.crit_s);
#pragma unroll
      for (int a = 0; a < 3; ++a) ns[a] = min_t(max_t(sig[a], lo), hi);
      jp = quot(T(P.Jp[p]) * (sig[0] * sig[1] * sig[2]), ns[0] * ns[1] * ns[2]);
      P.Jp[p] = (float)jp;
    }
    const M3<T> Fe = usv(U, ns, V);
#pragma unroll
    for (int e = 0; e < 9; ++e) P.F[e * cap + p] = (float)Fe.a[e / 3][e % 3];
    stress_from_svd(m, U, ns, jp, t);
  }
#pragma unroll
  for (int e = 0; e < 6; ++e) tau_out[e] = (float)t[e];
}

// The FP64 pass of material_update for one particle whose FP32 pass was deferred: F holds the updated
// (not yet projected) deformation gradient; SVD, return map and stress in double precision.
__device__ __forceinline__ void material_update_fp64(const MatDev& m, Particles& P, int64_t p) {
  M3<double> Fd, Ud, Vd;
#pragma unroll
  for (int e = 0; e < 9; ++e) Fd.a[e / 3][e % 3] = (double)P.F[e * P.cap + p];
  double sd[3];
  svd3(Fd, Ud, sd, Vd);
  float t[6];
  material_update(m, Fd, Ud, sd, Vd, P, t, p);
  store_tau_pm(P, p, t, m.p_mass);
}

// one workgroup of 256: the largest of count per-workgroup maxima into *out (NaN counts as +inf). Each
// thread keeps 8 loads in flight per round (independent partial maxima): one at a time, the reduction of
// a 1M-particle G2P's 3,990 maxima paid the L2 latency 16 times in a row on the substep's critical path.
// With seq_out, the maximum is published to the host: after *out, a system-scope fence, then *seq_out = seq,
// so a host that sees seq also sees the maximum (and needs no event behind the kernel).
// dev_out (optional): a device copy of the maximum, for a kernel that derives the next CFL step from it.
__device__ void reduce_block_maxima(const float* __restrict__ block_max, int count, float* __restrict__ out,
                                    uint32_t* seq_out = nullptr, uint32_t seq = 0u, float* dev_out = nullptr) {
  constexpr int kInFlight = 8;
  float part[kInFlight];
#pragma unroll
  for (int u = 0; u < kInFlight; ++u) part[u] = 0.0f;
  for (int i0 = threadIdx.x; i0 < count; i0 += kInFlight * blockDim.x) {
#pragma unroll
    for (int u = 0; u < kInFlight; ++u) {
      // past the end: the last entry again (a maximum does not change), so the loads need no branch
      const float v = block_max[min(i0 + u * (int)blockDim.x, count - 1)];
      part[u] = (v == v) ? fmaxf(part[u], v) : __int_as_float(0x7f800000);
    }
  }
  float m = part[0];
#pragma unroll
  for (int u = 1; u < kInFlight; ++u) m = fmaxf(m, part[u]);
  __shared__ float s_max[4];
  const float wm = wave_max(m);
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = wm;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float m = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
    *out = m;
    if (dev_out) *dev_out = m;
    if (seq_out) {
      __threadfence_system();
      *static_cast<volatile uint32_t*>(seq_out) = seq;
    }
  }
  __syncthreads();
}

// FP64 pass over the deferred particles (grid-stride over the work list). With vmax_out, workgroup 0
// first reduces G2P's per-workgroup speed maxima into it (the pass changes no velocity), which saves the
// reduction its own launch.
__global__ void __launch_bounds__(256) k_fp64_fixup(Particles P, const MatDev* __restrict__ mats,
                                                    const uint32_t* __restrict__ list,
                                                    const uint32_t* __restrict__ count,
                                                    const float* __restrict__ block_max, int block_count,
                                                    float* __restrict__ vmax_out, uint32_t* seq_out,
                                                    uint32_t seq, float* vmax_dev, DriftAhead da) {
  if (vmax_out && blockIdx.x == 0) reduce_block_maxima(block_max, block_count, vmax_out, seq_out, seq, vmax_dev);
  const uint32_t n = count ? *count : 0u;
  // the next substep's drift routing (it reads only the positions G2P just wrote; the FP64 pass below
  // changes no position), two particles a stride apart per thread
  if (da.keys) {
    const int stride = (int)(gridDim.x * blockDim.x);
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < da.n; p += 2 * stride)
      drift_route_pair(P, da.g, p, p + stride < da.n ? p + stride : -1, da.keys, da.slot, da.in_cnt, da.in_lst,
                       da.fwd);
  }
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int64_t p = list[i];
    material_update_fp64(mats[P.mat[p]], P, p);
  }
}

// c_stash: this thread's LDS words, `stride` apart (one column per thread of the workgroup: no bank
// conflicts). The new C waits there across the SVD instead of in registers, so the 64-byte record is written
// whole at the end: written a float4 at a time at two points of the kernel, its lines left the L2 partly
// written and cost about a quarter more HBM write traffic. kStageRecord: the whole record (C, stress,
// mass: 16 words) is left in the stash for the caller to store, instead of stored here.
template <bool kStageRecord = false>
__device__ __forceinline__ bool g2p_particle(Particles& P, const MatDev* __restrict__ mats, const GridGeom& g,
                                             const StepConsts& sc, float dt, const float4* __restrict__ vel,
                                             int64_t p, const ColliderSet& cs,
                                             const ColliderFrame* frames, float& speed, float* c_stash,
                                             int stride) {
  const int64_t cap = P.cap;
  const float xo[3] = {P.x[p], P.x[cap + p], P.x[2 * cap + p]};
  float fx[3];
  int base[3];
  stencil_base(g, xo, base, fx);
  float wx[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) wx[a][i] = bspline_w(fx[a], i);
  // v = sum w v_n and C = 4/dx^2 sum w v_n (x_n - x)^T (kernels.py:208-224). With o_b in {-1, 0, 1} the
  // stencil offset of the node along axis b less one, (x_n - x)_b = (o_b - (fx_b - 1)) dx, so
  // C_ab = 4/dx (M_ab - (fx_b - 1) v_a) with M_ab = sum w v_n,a o_b. The weight is a product over the axes,
  // so the sums are taken one axis at a time: along k, s = sum_k wz_k v and t = sum_k wz_k o_k v for each
  // (i, j); along j, S = sum_j wy_j s, T = sum_j wy_j t and U = sum_j wy_j o_j s for each i; along i, v,
  // M_.2 = sum_i wx_i T, M_.1 = sum_i wx_i U and M_.0 = sum_i wx_i o_i S (267 flops, not 27 x 13).
  // Node addresses split per axis in the brick-major layout: addr = ax[i] + ay[j] + az[k].
  int ad[3][3];
  {
    const int sx = g.bdims[1] * g.bdims[2] * 64, sy = g.bdims[2] * 64;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int X = base[0] + i, Y = base[1] + i, Z = base[2] + i;
      ad[0][i] = (X >> 2) * sx + ((X & 3) << 4);
      ad[1][i] = (Y >> 2) * sy + ((Y & 3) << 2);
      ad[2][i] = (Z >> 2) * 64 + (Z & 3);
    }
  }
  float Si[3][3], Ti[3][3], Ui[3][3];  // [i][component]
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float sj[3][3], tj[3][3];  // [j][component]
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float4* row = vel + (ad[0][i] + ad[1][j]);
      const float4 g0 = row[ad[2][0]], g1 = row[ad[2][1]], g2 = row[ad[2][2]];
      const float v0[3] = {g0.x, g0.y, g0.z}, v1[3] = {g1.x, g1.y, g1.z}, v2[3] = {g2.x, g2.y, g2.z};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float p0 = wx[2][0] * v0[a], p2 = wx[2][2] * v2[a];
        sj[j][a] = (p0 + wx[2][1] * v1[a]) + p2;
        tj[j][a] = p2 - p0;
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float s0 = wx[1][0] * sj[0][a], s2 = wx[1][2] * sj[2][a];
      Si[i][a] = (s0 + wx[1][1] * sj[1][a]) + s2;
      Ti[i][a] = wx[1][0] * tj[0][a] + wx[1][1] * tj[1][a] + wx[1][2] * tj[2][a];
      Ui[i][a] = s2 - s0;
    }
  }
  float nv[3], M[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float S0 = wx[0][0] * Si[0][a], S2 = wx[0][2] * Si[2][a];
    nv[a] = (S0 + wx[0][1] * Si[1][a]) + S2;
    M[a][0] = S2 - S0;
    M[a][1] = wx[0][0] * Ui[0][a] + wx[0][1] * Ui[1][a] + wx[0][2] * Ui[2][a];
    M[a][2] = wx[0][0] * Ti[0][a] + wx[0][1] * Ti[1][a] + wx[0][2] * Ti[2][a];
  }
  M3<float> nC;
  const float c4dx = 4.0f * g.inv_dx;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) nC.a[a][b] = c4dx * (M[a][b] - (fx[b] - 1.0f) * nv[a]);
#pragma unroll
  for (int e = 0; e < 9; ++e) c_stash[e * stride] = nC.a[e / 3][e % 3];
  float nx[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    nx[a] = xo[a] + dt * nv[a];
    if (sc.lo_on[a] != 0 && nx[a] < sc.wall_lo[a]) nx[a] = sc.wall_lo[a];
    if (sc.hi_on[a] != 0 && nx[a] > sc.wall_hi[a]) nx[a] = sc.wall_hi[a];
  }
  // a non-finite position is stored as it is so the NaN check of the runner sees it (kernels.py:236-241)
  if (isfinite(nx[0]) && isfinite(nx[1]) && isfinite(nx[2])) {
#pragma unroll
    for (int a = 0; a < 3; ++a) nx[a] = fminf(fmaxf(nx[a], sc.clamp_lo[a]), sc.clamp_hi[a]);
  }
  // project_particles (colliders.py:120-149) on the advected position and the new velocity
  if (cs.count > 0) project_particle(cs, frames, nx, nv);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    P.x[a * cap + p] = nx[a];
    P.v[a * cap + p] = nv[a];
  }
  speed = sqrtf(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
  M3<float> Fo;
  load_m3(P.F, cap, p, Fo);
  M3<float> G;
#pragma unroll
  for (int e = 0; e < 9; ++e) G.a[e / 3][e % 3] = (e % 4 == 0 ? 1.0f : 0.0f) + dt * nC.a[e / 3][e % 3];
  const M3<float> Fn = mul(G, Fo);
  const MatDev m = mats[P.mat[p]];
  M3<float> U, V;
  float sig[3];
  svd3(Fn, U, sig, V);
  float t[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  const bool defer = needs_fp64(sc.precision, sig, m.model == kSnow ? P.Jp[p] : 1.0f);
  if (defer) {
    // deferred to the FP64 pass (k_fp64_fixup), which reads Fn back from F and writes the stress
    store_m3(P.F, cap, p, Fn);
  } else {
    material_update(m, Fn, U, sig, V, P, t, p);
  }
  if constexpr (kStageRecord) {
#pragma unroll
    for (int e = 0; e < 6; ++e) c_stash[(9 + e) * stride] = t[e];
    c_stash[15 * stride] = m.p_mass;
    return defer;
  }
  float c[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) c[e] = c_stash[e * stride];
  P.rec[4 * p] = make_float4(c[0], c[1], c[2], c[3]);
  P.rec[4 * p + 1] = make_float4(c[4], c[5], c[6], c[7]);
  P.rec[4 * p + 2] = make_float4(c[8], t[0], t[1], t[2]);
  P.rec[4 * p + 3] = make_float4(t[3], t[4], t[5], m.p_mass);
  return defer;
}

// The records of a wave's 64 consecutive particles, slots p0 + 0..63 (those below n), from a stash of 16 rows
// `stride` words apart whose column c0 + l holds particle p0 + l's record (g2p_particle<true>):
// four lanes per record, one float4 each, so that each store instruction writes 16 whole records (one
// float4 of 64 records per instruction: G2P 45.2 instead of 41.9 us on the 1M mix). A stride of 257 puts
// the four words a lane reads in four banks. The whole wave calls it.
__device__ __forceinline__ void store_staged_records(const Particles& P, const float* stash, int stride, int c0,
                                                     int64_t p0, int64_t n) {
  wave_lds_fence();
  const int lane = threadIdx.x & 63, e = lane & 3;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int l = 16 * q + (lane >> 2);
    if (p0 + l < n) {
      const float* col = stash + c0 + l;
      P.rec[4 * (p0 + l) + e] = make_float4(col[4 * e * stride], col[(4 * e + 1) * stride],
                                            col[(4 * e + 2) * stride], col[(4 * e + 3) * stride]);
    }
  }
}

// largest speed of the workgroup into block_max[blockIdx.x]; NaN (which fmaxf would drop) counts as +inf,
// which the host also reads as "not finite"
__device__ __forceinline__ void block_max_speed(float speed, float* __restrict__ block_max) {
  __shared__ float s_max[4];
  if (!(speed == speed)) speed = __int_as_float(0x7f800000);
  const float wm = wave_max(speed);
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = wm;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = s_max[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = fmaxf(b, s_max[w]);
    block_max[blockIdx.x] = b;
  }
}

__global__ void __launch_bounds__(256) k_g2p(Particles P, const MatDev* __restrict__ mats, GridGeom g,
                                             StepConsts sc, int n, float dt, const float4* __restrict__ vel,
                                             float* __restrict__ block_max,
                                             ColliderSet cs, ColliderFrames fr, FixList fix) {
  const ColliderFrame* frames = fr.f;
  cs = with_records(cs, fr);
  __shared__ float s_c[16][257];  // each thread's record (g2p_particle stages it)
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  float speed = 0.0f;
  bool defer = false;
  if (p < n) defer = g2p_particle<true>(P, mats, g, sc, dt, vel, p, cs, frames, speed, &s_c[0][threadIdx.x], 257);
  const int w0 = threadIdx.x & ~63;
  store_staged_records(P, &s_c[0][0], 257, w0, (int64_t)blockIdx.x * blockDim.x + w0, n);
  if (sc.precision != kPrecisionFp32) wave_append(fix, defer, (uint32_t)p);
  block_max_speed(speed, block_max);
}

// |v| of every particle reduced per workgroup (kernels.py:266-269), when no G2P has produced it
__global__ void __launch_bounds__(256) k_speed_partials(Particles P, int n, float* __restrict__ block_max) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  float speed = 0.0f;
  if (p < n) {
    const float a = P.v[p], b = P.v[P.cap + p], c = P.v[2 * P.cap + p];
    speed = sqrtf(a * a + b * b + c * c);
  }
  block_max_speed(speed, block_max);
}

// one workgroup: the maximum of the per-workgroup maxima
__global__ void __launch_bounds__(256) k_reduce_max(const float* __restrict__ block_max, int count,
                                                    float* __restrict__ out, uint32_t* seq_out, uint32_t seq,
                                                    float* vmax_dev) {
  reduce_block_maxima(block_max, count, out, seq_out, seq, vmax_dev);
}

}  // namespace mpm
