/* mpm_hip.h: C-ABI of libmpm_hip.so, the MI355X (gfx950) MLS-MPM substep library.
 *
 * The reference's boundary for this path is the Python ABC mpm_core.solver.Solver
 * (src/mpm_core/solver.py:37-82) implemented by WarpSolver (src/mpm_core/backends/warp/solver.py:26-255).
 * Its FFI is Python: mpm_core/backends/hip/solver.py binds every entry point below with ctypes. Each
 * declaration names the reference method or kernel it replaces.
 *
 * Conventions
 *   - Plain pointers and sizes only; host arrays use the reference's numpy layouts: positions (N,3)
 *     float32, C and F (N,3,3) float32 row major, Jp (N,) float32, material/emitter ids (N,) int32,
 *     grids (D0,D1,D2,3) / (D0,D1,D2) float32 in C order with D = grid_shape + 6 (PAD = 3, origin =
 *     domain min - 3 dx).
 *   - Every call is synchronous with respect to host-visible outputs and returns MPM_OK or a negative
 *     code; mpm_last_error() gives the message of the last failure on the calling thread.
 *   - One HIP stream per context. Contexts share no state, so several may live on different threads.
 */
#ifndef MPM_HIP_H
#define MPM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_ABI_VERSION 1

#define MPM_OK 0
#define MPM_ERR_VALUE (-1)           /* bad argument: ValueError in the reference */
#define MPM_ERR_NOT_IMPLEMENTED (-2) /* unsupported material model: NotImplementedError */
#define MPM_ERR_MEMORY (-3)          /* device allocation failed: MemoryError */
#define MPM_ERR_RUNTIME (-4)         /* HIP runtime failure: RuntimeError */

#define MPM_MODEL_ELASTIC 0
#define MPM_MODEL_VON_MISES 1
#define MPM_MODEL_DRUCKER_PRAGER 2
#define MPM_MODEL_SNOW 3

typedef struct mpm_ctx mpm_ctx;
typedef struct mpm_batch mpm_batch;

/* Domain constants, derived on the host in FP64 and rounded to FP32 as in
 * backends/warp/solver.py:37-54. */
typedef struct mpm_domain_desc {
  int32_t dims[3];         /* grid_shape + 2 * PAD */
  float origin[3];         /* domain.min - PAD * dx */
  float dx, inv_dx, p_vol;
  float gravity[3];
  int32_t wall_lo_node[3]; /* floor((min - origin) / dx + 1e-3) */
  int32_t wall_hi_node[3]; /* ceil((max - origin) / dx - 1e-3) */
  int32_t lo_on[3], hi_on[3];
  float wall_lo[3], wall_hi[3];
  float clamp_lo[3], clamp_hi[3];
  float wall_friction;
  float cell_vol;          /* dx^3, for is_occupied */
  float remove_lo[3];      /* float32(min - dx) per open lo side, -inf when walled (solver.py:222-240) */
  float remove_hi[3];      /* float32(max + dx) per open hi side, +inf when walled */
  double dx_f64;           /* dx in double precision for the CFL step count (solver.py:176-185) */
} mpm_domain_desc;

/* One material (solver.py:56-71; materials.py:48-51; _dp_alpha solver.py:263-265). */
typedef struct mpm_material_desc {
  int32_t model;
  float mu, lam, yield_stress, alpha;
  float hardening, max_hardening, crit_c, crit_s;
  float p_mass;            /* float32(density) * float32(p_vol) (solver.py:105) */
  double wave_speed;       /* sqrt((lam + 2 mu) * h / rho), h = max_hardening for snow (solver.py:60-62) */
} mpm_material_desc;

/* Start/end pose of one collider over the next advance (colliders.py:20-28), decomposed on the host
 * (solver.py:160-172): M = [R(q) S | t], q as (x, y, z, w), q1 on q0's hemisphere. */
typedef struct mpm_collider_pose {
  float t0[3], t1[3];
  float q0[4], q1[4];
  float S0[9], S1[9];
  float margin;            /* AABB early-out margin in local units: 3 dx / smallest singular value */
} mpm_collider_pose;

typedef struct mpm_advance_stats {
  int64_t substeps;
  double dt;
  int64_t removed;
} mpm_advance_stats;

/* ---- lifecycle ---------------------------------------------------------------------------- */
int32_t mpm_abi_version(void);
const char* mpm_last_error(void);
/* replaces WarpSolver.__init__ (solver.py:27-78) */
int32_t mpm_create(int32_t device, const mpm_domain_desc* domain, const mpm_material_desc* materials,
                   int32_t n_materials, mpm_ctx** out);
int32_t mpm_destroy(mpm_ctx* ctx);

/* ---- particles ---------------------------------------------------------------------------- */
/* replaces WarpSolver.add_particles (solver.py:100-125): C = 0, F = I, Jp = 1 */
int32_t mpm_add_particles(mpm_ctx* ctx, int64_t n, const float* x, const float* v, const int32_t* mat,
                          const int32_t* emitter);
/* replaces WarpSolver.load_state (solver.py:137-138) */
int32_t mpm_load_state(mpm_ctx* ctx, int64_t n, const float* x, const float* v, const float* C, const float* F,
                       const float* Jp, const int32_t* mat, const int32_t* emitter);
/* replaces WarpSolver.particle_count (solver.py:127-129) */
int64_t mpm_particle_count(const mpm_ctx* ctx);
/* replaces get_particles / save_state (solver.py:131-135, 96-98); NULL skips a field.
 * Particles come back in insertion order whatever order the device keeps them in. */
int32_t mpm_get_state(mpm_ctx* ctx, float* x, float* v, float* C, float* F, float* Jp, int32_t* mat,
                      int32_t* emitter);

/* ---- colliders ---------------------------------------------------------------------------- */
/* replaces WarpSolver.set_colliders (solver.py:142-152): meshes concatenated, local space */
int32_t mpm_set_colliders(mpm_ctx* ctx, int32_t n, const int32_t* n_verts, const int32_t* n_tris,
                          const float* verts, const int32_t* tris, const int32_t* closed, const float* friction);
/* replaces WarpSolver.set_collider_transforms (solver.py:154-172) with poses decomposed by the caller */
int32_t mpm_set_collider_poses(mpm_ctx* ctx, int32_t n, const mpm_collider_pose* poses);
/* replaces WarpSolver.set_collider_transforms (solver.py:154-172) with _decompose (:258-260,
 * geometry.py:30-49) done here in FP64: start/end are n row-major 4x4 float32 world matrices. A count other
 * than the colliders' is a ValueError (solver.py:157-158); so is a non-finite entry. */
int32_t mpm_set_collider_transforms(mpm_ctx* ctx, int32_t n, const float* start, const float* end);
/* the decomposition mpm_set_collider_transforms applies (geometry.py:30-49 decompose_transform): of one
 * row-major 4x4 float32 matrix, translation t[3], rotation quaternion q[4] (x, y, z, w), symmetric stretch
 * S[9] (row-major) and the smallest singular value of the linear part; no device work */
int32_t mpm_decompose_transform(const float* m, double* t, double* q, double* S, double* smin);

/* ---- stepping ----------------------------------------------------------------------------- */
/* replaces WarpSolver.advance (solver.py:213-220) with _choose_dt (:176-185) and _remove_escaped (:222-240) */
int32_t mpm_advance(mpm_ctx* ctx, double frame_dt, mpm_advance_stats* out);
/* one fixed-dt substep (solver.py:187-211) at frame fraction s of a frame lasting frame_dt */
int32_t mpm_substep(mpm_ctx* ctx, double dt, double s, double frame_dt);

/* ---- queries ------------------------------------------------------------------------------ */
/* replaces WarpSolver.is_occupied (solver.py:244-255; kernels.py:272-298) */
int32_t mpm_is_occupied(mpm_ctx* ctx, int64_t m, const float* points, uint8_t* out);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): mpm_is_occupied with the query points
 * (m,3) float32 and the result (m,) uint8 in DEVICE memory the caller owns, and the volume grid in a fixed
 * array of MPM_NOALLOC_MAX_NODES floats (32 MiB) that each GPU holds for the life of the process. The call
 * allocates nothing; it enqueues the work on the context's stream and returns without waiting, so
 * mpm_synchronize(ctx) comes before `out` is read. Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (contexts on other streams, other host threads) race. A grid of more nodes (bricks of 4^3
 * nodes over grid_shape + 6) than the array holds is refused with MPM_ERR_VALUE: use mpm_is_occupied. */
#define MPM_NOALLOC_MAX_NODES 8388608
int32_t mpm_is_occupied_noalloc(mpm_ctx* ctx, int64_t m, const float* points, uint8_t* out);
/* the lattice part of sample_mesh_volume (src/mpm_core/geometry.py:89-155; the caller adds the jitter):
   points (i + 0.5) spacing inside the closed mesh (verts: n_verts x 3 float64, tris: n_tris x 3 indices)
   by z-ray crossing parity, as n x 3 float64 rows in the reference's order (column i * ny + j, then z).
   *count is set to n; the rows are written only when out is not null and capacity >= n (call with
   capacity 0 to size the buffer). Same FP64 arithmetic as the reference, so the points are identical. */
int32_t mpm_lattice_inside(mpm_ctx* ctx, const double* verts, int64_t n_verts, const int32_t* tris, int64_t n_tris,
                           double spacing, int64_t capacity, double* out, int64_t* count);

/* ---- phase hooks for parity tests (one reference kernel each) ------------------------------ */
/* grid zero + p2g (kernels.py:92-140) */
int32_t mpm_debug_p2g(mpm_ctx* ctx, double dt);
/* grid_update (kernels.py:155-180) then grid_colliders (colliders.py:77-117) */
int32_t mpm_debug_grid_update(mpm_ctx* ctx, double dt, double s, double frame_dt);
/* g2p (kernels.py:183-263) then project_particles (colliders.py:120-149) */
int32_t mpm_debug_g2p(mpm_ctx* ctx, double dt, double s, double frame_dt);
/* which = 0: P2G output (momentum, mass); which = 1: grid velocity after the update (velocity, mass) */
int32_t mpm_get_grid(mpm_ctx* ctx, int32_t which, float* grid_v, float* grid_m);
/* overwrite the grid velocity and mass (which = 1 layout) before mpm_debug_g2p */
int32_t mpm_set_grid(mpm_ctx* ctx, const float* grid_v, const float* grid_m);
/* particle binning of the last sort: per-particle cell key (brick * 64 + cell within the 4x4x4-node
 * brick, in slot order before the sort), the stable permutation, and per-cell offsets (n_bins + 1);
 * call with null arrays to get n_bins */
int32_t mpm_get_binning(mpm_ctx* ctx, int32_t* keys, int32_t* perm, int32_t* offsets, int64_t* n_bins);

/* run the brick binning (stable sort + reorder) now, as the next substep would (parity tests) */
int32_t mpm_debug_sort(mpm_ctx* ctx);

/* ---- options ------------------------------------------------------------------------------ */
#define MPM_OPT_P2G_MODE 0       /* MPM_P2G_BINNED (default), MPM_P2G_ATOMIC or MPM_P2G_FIXED */
#define MPM_OPT_SORT_INTERVAL 1  /* substeps between brick re-sorts (default 24; >= 1) */
#define MPM_OPT_P2G_BLOCKS 2     /* workgroups of the binned P2G launch (0: automatic) */
#define MPM_OPT_PRECISION 3      /* SVD, return maps and stress: MPM_PRECISION_FP32, _FALLBACK (default:
                                    FP64 only where FP32 cannot resolve F: sigma_min/sigma_max < 1e-2 or
                                    J = Jp det F < 0.25), _FP64 */
#define MPM_OPT_P2G_WAVES 4      /* waves per occupied brick in the binned P2G: 1, 2 (default) or 4; wave h of
                                    n takes the particles h, h + n, ... of each cell (more waves in flight
                                    for scenes with few bricks) */
#define MPM_OPT_P2G_FLUSH 5      /* binned P2G: how a brick's window reaches the grid update. MPM_P2G_FLUSH_STAGED
                                    (default): its 6^3-node core stored into the brick's staging slot, and the grid
                                    update sums the (1 to 8) cores covering each node in a fixed order;
                                    MPM_P2G_FLUSH_ATOMIC: its non-zero nodes added into the HBM accumulator with
                                    global float atomics */
#define MPM_P2G_FLUSH_STAGED 0
#define MPM_P2G_FLUSH_ATOMIC 1
#define MPM_P2G_BINNED 0         /* per-brick LDS window, global FP32 atomics for the flush */
#define MPM_P2G_ATOMIC 1         /* one thread per particle, 27 x 4 FP32 global atomics (reference flow) */
#define MPM_P2G_FIXED 2          /* as BINNED, accumulated in 64-bit fixed point: run-to-run bit-exact grids */
#define MPM_PRECISION_FP32 0
#define MPM_PRECISION_FALLBACK 1
#define MPM_PRECISION_FP64 2
int32_t mpm_set_option(mpm_ctx* ctx, int32_t option, int64_t value);
int64_t mpm_get_option(const mpm_ctx* ctx, int32_t option);

/* ---- measurement -------------------------------------------------------------------------- */
/* kernel ids for mpm_get_kernel_time */
#define MPM_K_P2G 0
#define MPM_K_GRID 1
#define MPM_K_G2P 2
#define MPM_K_BIN 3
#define MPM_K_COLLIDERS 4
#define MPM_K_OTHER 5
#define MPM_K_COUNT 6
#define MPM_TIMING_ALL ((1 << MPM_K_COUNT) - 1)
/* record HIP events around the launches of the kernel classes in `mask` (bit 1 << MPM_K_*, MPM_TIMING_ALL
   for all, 0 off; off by default) on the context's stream. Each class brackets its own kernel only: P2G is
   the P2G kernel (drift routing and the FP64 pass count as OTHER), G2P the G2P kernel. */
int32_t mpm_set_timing(mpm_ctx* ctx, int32_t mask);
/* accumulated device time and launch count of one kernel class since the last reset */
int32_t mpm_get_kernel_time(mpm_ctx* ctx, int32_t kernel, double* total_ms, int64_t* launches);
int32_t mpm_reset_timing(mpm_ctx* ctx);
/* wait for all work on the context's stream */
int32_t mpm_synchronize(mpm_ctx* ctx);

/* ---- batches of scenes ---------------------------------------------------------------------
 * No counterpart in the reference, whose runner bakes one scene per process (bake.py:116-208): a batch
 * advances many contexts (scenes) by one frame each through one launch sequence per substep, captured
 * as HIP graphs. Each scene keeps its own state and still answers every mpm_* call between frames;
 * while it belongs to a batch its operations run on the batch's stream. A context belongs to at most
 * one batch; mpm_destroy drops it from its batch. */
int32_t mpm_batch_create(int32_t device, mpm_batch** out);
int32_t mpm_batch_destroy(mpm_batch* batch);
/* replace the batch's scenes (all on the batch's device; order fixes the index of frame_dt/out below) */
int32_t mpm_batch_set_scenes(mpm_batch* batch, mpm_ctx* const* scenes, int32_t n);
int32_t mpm_batch_scene_count(const mpm_batch* batch);
/* one frame of every scene i of length frame_dt[i] (<= 0: scene i does not advance), each with its own
 * CFL substep count and escaped-particle removal, exactly as mpm_advance; out[i] as mpm_advance */
int32_t mpm_batch_advance(mpm_batch* batch, const double* frame_dt, mpm_advance_stats* out);
/* Frames that start and end independently (continuous batching): mpm_batch_begin starts a frame of
 * length frame_dt[i] for every scene i with frame_dt[i] > 0 (the scene must not be in a frame); then each
 * mpm_batch_run replays substeps of every scene in a frame: max_launches > 0 replays that many, 0 until
 * every frame in flight has ended, < 0 until the first one has. Scenes whose frame ended in the call get
 * finished[i] = 1 and out[i]. While a scene is in a frame, the mpm_* calls on it fail with MPM_ERR_VALUE. */
int32_t mpm_batch_begin(mpm_batch* batch, const double* frame_dt);
int32_t mpm_batch_run(mpm_batch* batch, int64_t max_launches, int64_t* launched, int32_t* finished,
                      mpm_advance_stats* out);
/* substeps done / in the frame in flight, per scene (equal when the scene is not in a frame) */
int32_t mpm_batch_progress(const mpm_batch* batch, int32_t* substeps_done, int32_t* substeps_total);
#define MPM_BATCH_OPT_P2G_BLOCKS 0        /* workgroups of the persistent P2G launch (0: automatic) */
#define MPM_BATCH_OPT_GRAPH 1             /* 1: replay captured graphs (default); 0: launch directly */
#define MPM_BATCH_OPT_SORT_INTERVAL 2     /* substeps between re-sorts for every scene (default 16; 0: each scene's own) */
#define MPM_BATCH_OPT_SUBSTEPS_LAUNCHED 3 /* read only: substep launches since creation */
#define MPM_BATCH_OPT_G2P_BLOCKS 4        /* workgroups of the persistent G2P launch (0: automatic) */
#define MPM_BATCH_OPT_GRID_BLOCKS 5       /* workgroups of the persistent grid-update launch (0: automatic) */
#define MPM_BATCH_OPT_G2P_TILE 6          /* particles per G2P workgroup: 64, 128 or 256 (default) */
#define MPM_BATCH_OPT_P2G_WAVES 7         /* waves per occupied brick in P2G: 1 (default), 2 or 4 */
int32_t mpm_batch_set_option(mpm_batch* batch, int32_t option, int64_t value);
int64_t mpm_batch_get_option(const mpm_batch* batch, int32_t option);
/* kernel classes of a batched substep, for mpm_batch_get_kernel_time */
#define MPM_BK_PLAN 0       /* k_batch_plan: which scenes run and sort, work prefixes, collider frames */
#define MPM_BK_SORT 1       /* binning of the scenes that sort: cell keys, radix passes, cell bounds, reorder */
#define MPM_BK_DRIFT 2      /* drift routing (with the occupied-brick prefix) */
#define MPM_BK_P2G 3
#define MPM_BK_GRID_SCAN 4  /* dirty bricks to work lists (with their prefix) */
#define MPM_BK_GRID 5
#define MPM_BK_G2P 6
#define MPM_BK_FIXUP 7      /* the FP64 pass */
#define MPM_BK_ADVANCE 8    /* end-of-substep bookkeeping */
#define MPM_BK_COUNT 9
/* HIP events around every kernel class of the batch's substeps (1: on, 0: off, the default). No counterpart
   in the reference. While on, substeps are launched directly instead of replayed from the captured graphs
   (a graph's kernels cannot be bracketed by events): the device time of each class is the same, the host's
   launch overhead is not. */
int32_t mpm_batch_set_timing(mpm_batch* batch, int32_t on);
/* accumulated device time and launch count of one class since the last reset */
int32_t mpm_batch_get_kernel_time(mpm_batch* batch, int32_t kernel, double* total_ms, int64_t* launches);
int32_t mpm_batch_reset_timing(mpm_batch* batch);

/* ---- surfacing -----------------------------------------------------------------------------
 * Replaces the device work of surface_particles (src/mpm_surface/field.py:100-142): the density splat
 * (field.py:35-55), its normalisation (:58-63), marching cubes (Warp's MarchingCubes, :130-131) and the
 * vertex velocity sample (:66-85). The host part (grid placement, rest density, memory budget, world
 * coordinates) stays in mpm_surface.field. A surfacer keeps its device buffers between calls, so meshing
 * frame after frame reallocates only when a frame needs more. */
typedef struct mpm_surface mpm_surface;
typedef struct {
  int32_t dims[3];  /* field nodes per axis (>= 2 each), node (i, j, k) at origin + (i, j, k) cell */
  float origin[3];
  float cell;       /* node spacing */
  float h;          /* kernel support radius */
  float inv_rest;   /* 1 / rest density: the normalised density of a uniform particle lattice reads 1 */
  float iso;        /* the surface's level of the normalised density */
} mpm_surface_desc;
int32_t mpm_surface_create(int32_t device, mpm_surface** out);
int32_t mpm_surface_destroy(mpm_surface* surf);
/* splat n particles (positions x, velocities v: (n,3) float32 host rows), normalise, and extract the
 * desc->iso surface of the density; the mesh stays on the device until the next build. n_verts and n_tris
 * receive its sizes (0 and 0 when nothing crosses iso). */
int32_t mpm_surface_build(mpm_surface* surf, const mpm_surface_desc* desc, int64_t n, const float* x,
                          const float* v, int64_t* n_verts, int64_t* n_tris);
/* the particles' upload on its own, for a caller that places the field from their bounds (as
 * surface_particles does, field.py:112-118): copies x and v (n > 0) to the device and returns the minimum
 * and maximum of each coordinate, computed there (lo[3], hi[3]; all NaN when any coordinate is NaN, as
 * numpy's min and max give for the column). mpm_surface_build_uploaded then meshes them. */
int32_t mpm_surface_upload(mpm_surface* surf, int64_t n, const float* x, const float* v, float* lo, float* hi);
int32_t mpm_surface_build_uploaded(mpm_surface* surf, const mpm_surface_desc* desc, int64_t* n_verts,
                                   int64_t* n_tris);
/* the last build's mesh: vertices (n_verts,3) float32 in field-grid units (world = origin + cell * g, as
 * Warp's marching cubes returns them), triangles (n_tris,3) int32 wound with normals pointing out of the
 * material (the reference flips Warp's inward winding, field.py:138), vertex velocities (n_verts,3)
 * float32. A null output is skipped. */
int32_t mpm_surface_fetch(mpm_surface* surf, float* verts, int32_t* tris, float* vel);
/* the last build's field (no counterpart: for tests): normalised density (D0,D1,D2) float32 and velocity
 * (D0,D1,D2,3) float32 (momentum / density where density > 0, else 0), C order; a null output is skipped */
int32_t mpm_surface_get_field(mpm_surface* surf, float* density, float* velocity);
/* device time of the last build in ms: [0] upload of the particles (with the bounds, when they came by
 * mpm_surface_upload), [1] splat (cell sort and gather), [2] normalisation, [3] marching cubes with the
 * vertex velocities (both passes, their scans and the size read-back) */
int32_t mpm_surface_times(const mpm_surface* surf, float* ms4);
/* the marching cubes tables (host only, no device call): count[256] triangles per corner configuration
 * (bit c set: corner c = (c & 1, c >> 1 & 1, c >> 2 & 1) above iso) and edges[256 * 36], three cell edges
 * per triangle; edge e joins corners MC_EDGE_ENDS[e] (x edges 0-3 from corners 0, 2, 4, 6; y edges 4-7
 * from corners 0, 1, 4, 5; z edges 8-11 from corners 0, 1, 2, 3) */
int32_t mpm_surface_tables(int32_t* count, int8_t* edges);
/* free and total device memory in bytes (the reference's default surfacing budget is 70% of free) */
int32_t mpm_device_memory(int32_t device, int64_t* free_bytes, int64_t* total_bytes);

#ifdef __cplusplus
}
#endif

#endif /* MPM_HIP_H */
