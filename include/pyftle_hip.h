/*
 * pyftle_hip.h -- C ABI of the MI355X FTLE engine (libpyftle_hip.so).
 *
 * Every entry point takes plain device pointers and sizes, is enqueued on the
 * HIP stream passed as `stream` (NULL = the null stream), and returns
 * PYFTLE_OK (0) or a negative error code; pyftle_last_error() then holds the
 * message (thread-local).  Argument errors are detected on the host before any
 * launch.  Nothing is allocated on the device by these calls unless noted.
 *
 * Reference interfaces replaced (pyFTLE 1.0.3, paths under src/pyftle/):
 *   grid_interp/GridInterp.cpp:93-142  interp2d_vec_inplace  -> pyftle_interp2d_vec
 *   grid_interp/GridInterp.cpp:15-91   interp3d_vec_inplace  -> pyftle_interp3d_vec
 *   interpolate.py:342-472  HighPerfInterpolator layout/update -> pyftle_pack_snapshot
 *   interpolate.py:475-543  HighPerfInterpolator.interpolate   -> pyftle_field_sample
 *   integrate.py:161-342    Euler/AB2/RK4 .integrate(h, p)     -> pyftle_advect
 *   ftle_solver.py:68-161   FTLESolver.run window + _compute_ftle -> pyftle_advect_ftle
 *   cauchy_green.py:15-210  compute_flow_map_jacobian_2x2/3x3  -> pyftle_flow_map_jacobian
 *   ftle.py:9-96            compute_cauchy_green_2x2/3x3       -> pyftle_cauchy_green
 *   ftle.py:99-163          max_eigenvalue_2x2/3x3             -> pyftle_max_eigenvalue
 *   ftle.py:166-233         compute_ftle_2x2/3x3               -> pyftle_ftle_from_jacobian
 *   cauchy_green.py + ftle.py (fused)                          -> pyftle_ftle_from_groups
 *   interpolate.py:546-762  GridInterpolator (SciPy RegularGridInterpolator, methods
 *                           linear / nearest / cubic) -> pyftle_field.interp modes,
 *                           pyftle_pack_spline (cubic coefficients)
 *   interpolate.py:114-294  CubicInterpolator / LinearInterpolator /
 *                           NearestNeighborInterpolator (SciPy CloughTocher2DInterpolator,
 *                           LinearNDInterpolator, NearestNDInterpolator on scattered
 *                           points) -> PYFTLE_INTERP_SCATTER_* with pyftle_scatter
 */
#ifndef PYFTLE_HIP_H
#define PYFTLE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define PYFTLE_OK 0
#define PYFTLE_EINVAL (-1) /* bad argument (shape, dtype, pointer); maps to ValueError/RuntimeError */
#define PYFTLE_EHIP (-2)   /* HIP runtime error at launch */

/* ---- element types ---- */
#define PYFTLE_F32 0
#define PYFTLE_F64 1

/* ---- snapshot record layouts in the device ring ---- */
#define PYFTLE_LAYOUT_SINGLE 0 /* record = one time level: 2D (u,v); 3D (u,v,w,0) */
#define PYFTLE_LAYOUT_PAIR 1   /* record = levels k,k+1: 2D (u0,v0,u1,v1); 3D (u0,v0,w0,0,u1,v1,w1,0) */
/* Per-cell records (FP32 GridInterp-mode fields, read by the FP32 advection and batch
 * kernels and by pyftle_field_sample): the record of cell (i, j[, k]) holds the 2^dim corner nodes of the cell,
 * corner q = (i + (q&1), j + (q>>1 & 1)[, k + (q>>2 & 1)]), dim floats each, corner after
 * corner; one level (CELL) or levels k, k+1 one after the other (CELL_PAIR).  n cells per
 * axis (the last one holds valid corners on periodic axes only).  One interpolation reads
 * one contiguous record (96 bytes in 3D) instead of 2^dim node records. */
#define PYFTLE_LAYOUT_CELL 2
#define PYFTLE_LAYOUT_CELL_PAIR 3

/* ---- integrators (integrate.py:350-387 names) ---- */
#define PYFTLE_EULER 0
#define PYFTLE_AB2 1
#define PYFTLE_RK4 2

/* ---- time treatment of the velocity field ---- */
#define PYFTLE_TIME_STEADY 0 /* every stage reads snapshot `first_snapshot` */
#define PYFTLE_TIME_FROZEN 1 /* step i reads snapshot first_snapshot+i for all stages (ftle_solver.py:99-102) */
#define PYFTLE_TIME_LINEAR 2 /* stage time t blends snapshots k,k+1 linearly (needs PYFTLE_LAYOUT_PAIR) */

/* ---- compute precision of the advection (FP32 fields only; FP64 fields are FP64 throughout) ---- */
#define PYFTLE_COMPUTE_NATIVE 0 /* FP32: compensated FP32 positions, FP32 index and blend */
#define PYFTLE_COMPUTE_MIXED 1  /* FP64 positions, cell index and weights; FP32 corner blend */
#define PYFTLE_COMPUTE_F64 2    /* FP64 positions and FP64 blend of the FP32 records */

/* ---- spatial interpolation of a field (pyftle_field.interp) ---- */
#define PYFTLE_INTERP_GINTERP 0     /* GridInterp.cpp: multilinear, zero outside, upper edge excluded */
#define PYFTLE_INTERP_RGI_LINEAR 1  /* RegularGridInterpolator(method="linear", bounds_error=False,
                                       fill_value=0): upper edge included, NaN -> NaN */
#define PYFTLE_INTERP_RGI_NEAREST 2 /* method="nearest": nearest node, ties to the lower node */
#define PYFTLE_INTERP_RGI_CUBIC 3   /* method="cubic": tensor-product not-a-knot cubic spline; the
                                       slot holds B-spline coefficients (pyftle_pack_spline) */
#define PYFTLE_INTERP_SCATTER_LINEAR 4  /* LinearNDInterpolator (interpolate.py:188-246): barycentric
                                           blend in the Delaunay simplex holding the point, NaN outside
                                           the convex hull; the slot holds one record per data point */
#define PYFTLE_INTERP_SCATTER_NEAREST 5 /* NearestNDInterpolator (interpolate.py:249-294): the record
                                           of the nearest data point (ties: lower index) */
#define PYFTLE_INTERP_SCATTER_CUBIC 6   /* CloughTocher2DInterpolator (interpolate.py:114-185), 2D: the
                                           Clough-Tocher cubic patch of the containing triangle, NaN
                                           outside the hull; one record per data point of 8 components
                                           (u, v, du/dx, du/dy, dv/dx, dv/dy, 0, 0) */

/* Scattered data (the two SCATTER modes): device arrays, described on the host.
 * The arrays must be consistent (ids below their counts, offsets of length
 * buckets + 1); the library reads them as given.
 * Simplices and points are indexed through uniform bucket grids (CSR lists):
 * bucket (bx,by,bz) covers [origin + b*size, origin + (b+1)*size) per axis and is
 * entry bx + dims[0]*(by + dims[1]*bz). */
typedef struct pyftle_scatter {
  int64_t n_points;            /* data points = records per slot */
  const double *points;        /* (n_points, dim) FP64 (NEAREST) */
  int64_t n_simplices;
  const int32_t *simplices;    /* (n_simplices, dim+1) point indices (LINEAR) */
  const double *transform;     /* (n_simplices, dim+1, dim): scipy Delaunay.transform (LINEAR) */
  const int32_t *hull;         /* (n_simplices,) 1 if a facet lies on the convex hull (LINEAR) */
  int32_t simplex_dims[3];     /* simplex bucket grid */
  double simplex_origin[3];
  double simplex_size[3];
  const int32_t *simplex_offsets; /* (buckets + 1) */
  const int32_t *simplex_ids;
  int32_t point_dims[3];       /* point bucket grid (NEAREST) */
  double point_origin[3];
  double point_size[3];
  const int32_t *point_offsets;   /* (buckets + 1) */
  const int32_t *point_ids;
  const double *ct_g;          /* (n_simplices, 3) cross-edge direction factors (CUBIC) */
} pyftle_scatter;

/* ---- work-to-lane order for grouped kernels ---- */
#define PYFTLE_ORDER_SEED_MAJOR 0     /* the nb particles of a seed sit in adjacent lanes */
#define PYFTLE_ORDER_NEIGHBOR_MAJOR 1 /* a wave holds one neighbour direction of consecutive seeds */

/*
 * A gridded velocity field resident in HBM as a ring of snapshot slots.
 * Node (ix,iy[,iz]) of slot s is record  ix + px*(iy + py*iz)  at
 * data + s*slot_stride, with px = n[0] + periodic[0], py = n[1] + periodic[1]:
 * a periodic axis stores one ghost node (a copy of node 0) at index n, written
 * by pyftle_pack_snapshot.  Snapshot (or pair) k lives in slot k % ring_slots.
 * FP32 slots are limited to 2 GiB; rows and xy-slices to < 2^24 records.
 * Physical -> index mapping is (p - origin)/spacing (grid_interp.py:78-79).
 * The RGI interp modes follow interpolate.py:546-762 instead: axes are
 * numpy.linspace(origin, upper, n) (node k = k*spacing + origin, the last node
 * is `upper`), no periodic axes, SINGLE layout, STEADY/FROZEN time; the CUBIC
 * mode stores coefficient c[ix][iy][iz] (ix in -1..n) at record
 * (ix+1) + (n[0]+2)*((iy+1) + (n[1]+2)*(iz+1)).
 */
typedef struct pyftle_field {
  int32_t dim;          /* 2 or 3 */
  int32_t n[3];         /* grid points along x, y, z (n[2] = 1 in 2D) */
  double origin[3];     /* grid minimum per axis */
  double spacing[3];    /* (max - min)/(n - 1) per axis (grid_interp.py:34-35) */
  int32_t periodic[3];  /* 1: index wraps modulo n (extension); 0: zero outside (GridInterp.cpp:58,126) */
  int32_t dtype;        /* PYFTLE_F32 or PYFTLE_F64 record components; also the compute precision */
  int32_t layout;       /* PYFTLE_LAYOUT_* */
  int32_t ring_slots;   /* number of resident slots (>= 1) */
  int32_t n_snapshots;  /* snapshots in the time series (time clamping in LINEAR mode) */
  int64_t slot_stride;  /* bytes between slots (>= record bytes * nodes) */
  double t0;            /* time of snapshot 0 (LINEAR mode) */
  double dt_snap;       /* time between snapshots (LINEAR mode, non-zero) */
  void *data;           /* device pointer to slot 0 */
  int32_t interp;       /* PYFTLE_INTERP_* (0 = GridInterp.cpp semantics) */
  int32_t reserved;     /* 0 */
  double upper[3];      /* grid maximum per axis (last linspace node), RGI modes */
  const pyftle_scatter *scatter; /* SCATTER modes (n[0] = n_points, n[1] = n[2] = 1); else NULL */
} pyftle_field;

typedef struct pyftle_advect_opts {
  int32_t method;         /* PYFTLE_EULER / PYFTLE_AB2 / PYFTLE_RK4 */
  int32_t time_mode;      /* PYFTLE_TIME_* */
  int32_t n_steps;        /* steps taken in this launch (>= 0) */
  int32_t first_snapshot; /* STEADY: snapshot read; FROZEN: snapshot of step 0 */
  double t_start;         /* LINEAR: time of step 0 */
  double h;               /* signed step size */
  int32_t ab2_have_prev;  /* AB2: 1 if `vprev` already holds the previous velocity */
  int32_t lane_order;     /* PYFTLE_ORDER_* (grouped kernels) */
  int32_t groups_per_block; /* seeds per work-group (grouped kernels), 0 = default */
  int32_t xcd_remap;      /* 1: map consecutive work to one XCD's L2 (8 XCDs) */
  int32_t compute;        /* PYFTLE_COMPUTE_* */
  int32_t reserved;       /* 0 */
  const int32_t *seed_order; /* device (n_groups,) permutation or NULL: the k-th seed group
                                processed is group seed_order[k] (positions read and results
                                written at its own index, so results do not depend on it; an
                                entry outside [0, n_groups) processes group k instead);
                                see pyftle_seed_morton_codes */
} pyftle_advect_opts;

/* ---- library ---- */
int pyftle_version(void);             /* 10000*major + 100*minor + patch */
const char *pyftle_last_error(void);  /* message of the last failing call on this thread */
int pyftle_device_sync(void);         /* hipDeviceSynchronize, for tests */

/* ---- gather-only interpolation: GridInterp.cpp:93-142 / 15-91 ----
 * v (nx,ny) or (X,Y,Z) row-major FP64, points (n,dim) in index space, out (n,).
 * Same arithmetic and evaluation order as the reference (bit-exact in FP64);
 * zero outside [0,n-2] cells, including a point exactly on the upper edge. */
int pyftle_interp2d_vec(const double *v, int32_t nx, int32_t ny, const double *points,
                        int64_t n, double *out, void *stream);
int pyftle_interp3d_vec(const double *v, int32_t X, int32_t Y, int32_t Z, const double *points,
                        int64_t n, double *out, void *stream);

/* ---- snapshot ingest: interpolate.py:342-472 (layout of one snapshot) ----
 * Copies snapshot `snapshot` from a component-planar source into the ring.
 * Source element (c, ix, iy, iz) is at src[c*s[0] + ix*s[1] + iy*s[2] + iz*s[3]]
 * (element strides), so the reference's 2D F-order and 3D axis-swapped layouts
 * are both expressed by strides.  PAIR layout writes level 0 of pair slot k and
 * level 1 of pair slot k-1 (and level 1 of slot k when k is the last snapshot). */
int pyftle_pack_snapshot(const pyftle_field *field, int32_t snapshot, const void *src,
                         int32_t src_dtype, const int64_t src_strides[4], void *stream);

/* ---- cubic spline ingest: GridInterpolator(method="cubic").update (interpolate.py:546-762) ----
 * Fits the tensor-product not-a-knot cubic spline of one snapshot (what SciPy's
 * RegularGridInterpolator method="cubic" builds, solved exactly instead of by its
 * default iterative solver) and stores its uniform B-spline coefficients in slot
 * `snapshot` of a PYFTLE_INTERP_RGI_CUBIC field.  Source strides as for
 * pyftle_pack_snapshot.  `work` is device scratch of at least
 * pyftle_spline_work_bytes(field) bytes (FP64 node copy and two sweep buffers). */
int64_t pyftle_spline_work_bytes(const pyftle_field *field);
int pyftle_pack_spline(const pyftle_field *field, int32_t snapshot, const void *src,
                       int32_t src_dtype, const int64_t src_strides[4], void *work, void *stream);

/* ---- physical-space sampling: HighPerfInterpolator.interpolate (interpolate.py:475-543)
 * points (n,dim) of `pts_dtype`, out (n,dim) of the same dtype; velocity at time
 * t (LINEAR) or of snapshot `snapshot` (STEADY/FROZEN). */
int pyftle_field_sample(const pyftle_field *field, int32_t time_mode, int32_t snapshot, double t,
                        const void *points, int32_t pts_dtype, int64_t n, void *out, void *stream);

/* ---- advection: Integrator.integrate (integrate.py:161-342), n_steps at once ----
 * positions (n_particles, dim) of pos_dtype, updated in place.  vprev (same
 * shape/dtype) is the AB2 state (may be NULL for Euler/RK4).  Compute runs in
 * the field's dtype; FP32 keeps a compensated (hi, lo) position. */
int pyftle_advect(const pyftle_field *field, const pyftle_advect_opts *opts, void *positions,
                  int32_t pos_dtype, int64_t n_particles, void *vprev, void *stream);

/* ---- integrator stage arithmetic for caller-evaluated velocities: integrate.py:16-109 ----
 * The analytical strategy evaluates the user's velocity_fn(t, points) on device
 * arrays at every stage (interpolate.py:765-812); these element-wise updates,
 * in the reference's operation order, replace euler_step_inplace,
 * adams_bashforth_2_step_inplace and runge_kutta_4_step_inplace plus the stage
 * positions of RungeKutta4Integrator.integrate (integrate.py:291-342).
 * x, k1..k4, out: n contiguous elements of `dtype` (n = particles * dim).
 *   STAGE: out = x + c*k1          c = 0.5*h or h  (tmp of integrate.py:327-338)
 *   RK4:   x = x + c*(k1 + 2 k2 + 2 k3 + k4)       c = h/6 (integrate.py:109)
 *   EULER: x = x + c*k1; out (may be NULL) = k1    c = h   (integrate.py:44)
 *   AB2:   x = x + c*(1.5 k1 - 0.5 k2); out (may be NULL, may equal k2) = k1 */
#define PYFTLE_COMBINE_STAGE 0
#define PYFTLE_COMBINE_RK4 1
#define PYFTLE_COMBINE_EULER 2
#define PYFTLE_COMBINE_AB2 3
int pyftle_stage_combine(int32_t op, int32_t dtype, int64_t n, double c, void *x, const void *k1,
                         const void *k2, const void *k3, const void *k4, void *out, void *stream);

/* ---- seed order (no reference counterpart): Morton (Z-order) code of the cell
 * floor((centroid - lo) * inv_cell) (clamped to [0, 2^20) per axis) of each seed group of a
 * block-ordered (nb*n_groups, dim) position array; codes (n_groups,) in int64 storage.
 * Sorting the groups by code (stable) and passing the permutation as opts->seed_order
 * puts spatially close seeds in the same waves, so a wave's gathers touch fewer lines. */
int pyftle_seed_morton_codes(const void *positions, int32_t pos_dtype, int64_t n_groups, int32_t nb,
                             int32_t dim, const double lo[3], const double inv_cell[3], int64_t *codes,
                             void *stream);

/* ---- fused window: FTLESolver.run + _compute_ftle (ftle_solver.py:68-161) ----
 * positions_in: block-ordered (nb*n_groups, dim) [L,R,T,B(,F,Bk)] of pos_dtype
 * (particles.py:170,191,212); nb = 2*dim.  Advects every particle for
 * opts->n_steps, then per seed forms J, C and lambda_max in FP64 and writes
 * ftle_out[s] = (1/map_period) * log(sqrt(lambda_max)).  positions_out (may be
 * NULL, may equal positions_in) receives the final positions. */
int pyftle_advect_ftle(const pyftle_field *field, const pyftle_advect_opts *opts,
                       const void *positions_in, int32_t pos_dtype, int64_t n_groups,
                       void *positions_out, double map_period, double *ftle_out, void *stream);

/* ---- instrumentation (no reference counterpart): measuring the gather-bound kernel ----
 * pyftle_advect_ftle_count_loads: the same window as pyftle_advect_ftle (identical
 * results) for RK4 on FP32 CELL / CELL_PAIR records, launched with a counting copy of
 * the kernel that adds the wave-level 16-byte record loads it issues to *loads (device
 * uint64, accumulated; a wave issues a record's loads when any of its lanes needs the
 * record).  Run outside timed regions: the count is a property of the inputs and the
 * launch shape, the timed launches use the plain kernel.
 * pyftle_probe_load_rate: `blocks` work-groups of 256 lanes each issue `iters` (a
 * multiple of 4) 16-byte-per-lane wave loads, coalesced and L1-resident, from the first
 * 16 KiB of `buf`; one float per lane goes to `out` (blocks * 256).  Its rate is the peak
 * of the vector-memory path for such loads.
 * pyftle_probe_issue_rates: the issue cost of the fused MLP kernels' instruction mix.
 * `blocks` work-groups, each reserving 96 KB of LDS (one per CU), run `iters` iterations of
 * 8 slots per wave: mode 0 one MFMA 16x16x32 bf16 per slot, 1 two FP32 multiplies, 2 two
 * exp2, 3 two reciprocals, 4 one MFMA and two multiplies in the same wave (one wave per
 * SIMD); modes 5, 6, 7 pair an MFMA wave (waves 0-3) with a multiply, exp2 or reciprocal
 * wave (waves 4-7) on every SIMD.  cycles[2 b + k] (device uint64, 2 * blocks) receives
 * the s_memtime cycles of wave 0 (k = 0) and wave 4 (k = 1, pairs only) of work-group b;
 * out (blocks * 512 floats) keeps the results live. */
int pyftle_advect_ftle_count_loads(const pyftle_field *field, const pyftle_advect_opts *opts,
                                   const void *positions_in, int32_t pos_dtype, int64_t n_groups,
                                   void *positions_out, double map_period, double *ftle_out,
                                   unsigned long long *loads, void *stream);
int pyftle_probe_load_rate(const void *buf, int32_t blocks, int32_t iters, float *out, void *stream);
int pyftle_probe_issue_rates(int32_t mode, int32_t blocks, int32_t iters, unsigned long long *cycles, float *out,
                             void *stream);

/* ---- FTLE tensors (FP64 in/out) ----
 * pyftle_flow_map_jacobian: cauchy_green.py:15-153.  positions / initial
 *   positions block-ordered (nb*n, dim) of pos_dtype -> J (n,dim,dim).
 * pyftle_cauchy_green: ftle.py:9-96 (2D F^T F, 3D F F^T) -> (n,dim,dim).
 * pyftle_max_eigenvalue: ftle.py:99-163; 2D closed form for a general 2x2
 *   (NaN for complex pairs, as the reference), 3D symmetric trigonometric form.
 * pyftle_ftle_from_jacobian: ftle.py:166-233.
 * pyftle_ftle_from_groups: Jacobian + FTLE fused. */
int pyftle_flow_map_jacobian(const void *positions, const void *initial_positions,
                             int32_t pos_dtype, int32_t dim, int64_t n, double *J_out,
                             void *stream);
int pyftle_cauchy_green(const double *J, int32_t dim, int64_t n, double *C_out, void *stream);
int pyftle_max_eigenvalue(const double *C, int32_t dim, int64_t n, double *lam_out,
                          void *stream);
int pyftle_ftle_from_jacobian(const double *J, int32_t dim, int64_t n, double map_period,
                              double *ftle_out, void *stream);
int pyftle_ftle_from_groups(const void *positions, const void *initial_positions,
                            int32_t pos_dtype, int32_t dim, int64_t n, double map_period,
                            double *ftle_out, void *stream);

/* ---- batched windows: many FTLE fields in one persistent launch ----
 * No reference counterpart: pyFTLE runs one window per worker process
 * (parallel.py:224-228, __main__.py:111-166); here every window of a batch is
 * split into work items (field, block of seeds), ordered longest-first, and
 * one persistent kernel steals items from a device counter.  All fields of a
 * batch share dim, integrator, time mode, record layout and dtype; grids,
 * snapshot rings, seed counts, windows and step sizes may all differ.
 * pyftle_batch_create allocates the device work list and, with use_graph,
 * captures counter reset + persistent kernel into a hipGraph; launches then
 * replay it.  The fields' buffers must outlive the batch.  Launches of one
 * batch share its work counter: they must not overlap on different streams.
 * pyftle_batch_destroy waits for the device (a replay may still run) and
 * frees the batch. */
typedef struct pyftle_batch_field {
  pyftle_field field;
  const void *positions_in; /* block-ordered (2*dim*n_groups, dim) of pyftle_batch_opts.pos_dtype */
  void *positions_out;      /* final positions (may be NULL) */
  double *ftle_out;         /* (n_groups,) FP64 */
  int64_t n_groups;
  int32_t n_steps;
  int32_t first_snapshot;
  double t_start;
  double h;
  double map_period;
  const int32_t *seed_order; /* processing order of the seed groups or NULL (as in pyftle_advect_opts) */
} pyftle_batch_field;

typedef struct pyftle_batch_opts {
  int32_t method;           /* PYFTLE_EULER / PYFTLE_AB2 / PYFTLE_RK4 */
  int32_t time_mode;        /* PYFTLE_TIME_* */
  int32_t pos_dtype;        /* PYFTLE_F32 / PYFTLE_F64 */
  int32_t groups_per_block; /* seeds per work item, 0 = 64 */
  int32_t lane_order;       /* PYFTLE_ORDER_* */
  int32_t blocks_per_cu;    /* persistent work-groups per CU, 0 = occupancy limit */
  int32_t longest_first;    /* 1: items sorted by steps, longest first */
  int32_t use_graph;        /* 1: replay a captured hipGraph */
} pyftle_batch_opts;

typedef struct pyftle_batch pyftle_batch; /* opaque */

int pyftle_batch_create(const pyftle_batch_field *fields, int32_t n_fields,
                        const pyftle_batch_opts *opts, pyftle_batch **out);
int pyftle_batch_launch(pyftle_batch *batch, void *stream);
int pyftle_batch_info(const pyftle_batch *batch, int64_t *n_items, int32_t *grid_blocks,
                      int32_t *block_threads);
int pyftle_batch_destroy(pyftle_batch *batch);
/* Instrumentation (no reference counterpart): pyftle_batch_count_loads runs the batch once
 * (not through its graph) with a counting copy of the persistent kernel, for RK4 batches on
 * CELL / CELL_PAIR records: the same results, and the wave-level 16-byte record loads the
 * launch issues are added to *loads (device uint64, accumulated), as
 * pyftle_advect_ftle_count_loads does for one window.  Run outside timed regions. */
int pyftle_batch_count_loads(pyftle_batch *batch, unsigned long long *loads, void *stream);

/* ---- batched windows without device allocation ----
 * No reference counterpart (pyFTLE allocates nothing on a GPU).  For callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc -- this runs one batch like
 * pyftle_batch_create + pyftle_batch_launch (same validation, same per-item
 * arithmetic, bit-identical FTLE and positions) without allocating: the
 * descriptor table, the prefix sums of the fields' block counts and the work
 * counter live in one static device array of PYFTLE_BATCH_NOALLOC_SCRATCH_BYTES
 * per GPU, which the library keeps until the process ends.  At most
 * PYFTLE_BATCH_NOALLOC_MAX_FIELDS fields per call (more -> PYFTLE_EINVAL).
 * opts->use_graph is ignored (nothing is captured).  Enqueued on `stream`; the
 * descriptors are copied from host memory before the call returns, so `fields`
 * may be reused at once.  Every call on one GPU shares that GPU's array (the
 * counter is cleared by a kernel at each call): calls that overlap on one GPU
 * (other host threads, or other streams whose work has not finished) race. */
#define PYFTLE_BATCH_NOALLOC_MAX_FIELDS 256
#define PYFTLE_BATCH_NOALLOC_SCRATCH_BYTES (192 * 1024)
int pyftle_batch_run_noalloc(const pyftle_batch_field *fields, int32_t n_fields,
                             const pyftle_batch_opts *opts, void *stream);

/* ---- flow-map surrogate: Fourier-feature MLP kernels ----
 * No reference counterpart: a learned approximation of the flow map that
 * pyFTLE's RK4 path (integrate.py:265-342) computes; RK4 stays the fallback
 * and the accuracy reference.  BF16 operands, FP32 accumulation (MFMA 32x32x16). */
#define PYFTLE_EPI_F32 0       /* C (FP32, split-K partials at C + split*M*ldc) = A B */
#define PYFTLE_EPI_BIAS_SILU 1 /* Z = A B + bias (bf16, optional); C = SiLU(Z) (bf16) */
#define PYFTLE_EPI_DSILU 2     /* C = (A B) * SiLU'(Z) (bf16), Z read at the same row modulo z_rows */
#define PYFTLE_EPI_BIAS 3      /* C = A B (+ bias) (bf16) */

/* C[M,N] = A[M,K] B[K,N].  a_k_contig: A stored row-major MxK (else A^T row-major);
 * b_k_contig: B^T stored row-major NxK, e.g. a weight [out][in] (else B row-major KxN).
 * M, N multiples of 128; K a multiple of 64*splits; leading dimensions multiples of 8.
 * z_rows: row period of Z for EPI_DSILU (M for a plain backward pass; the primal
 * row count when tangent blocks are stacked below the primal rows). */
int pyftle_mlp_gemm(const void *A, int64_t lda, int32_t a_k_contig, const void *B, int64_t ldb,
                    int32_t b_k_contig, int32_t M, int32_t N, int32_t K, int32_t splits,
                    int32_t epilogue, const float *bias, void *Z, int32_t z_rows, void *C,
                    int64_t ldc, void *stream);
/* Fourier features [sin(2 pi B v), cos(2 pi B v)] of v = in*scale (B: F x d, FP32) as bf16
 * rows of ld_out; with n_tangents, d/d in_a of the features for a < n_tangents in row
 * blocks (1+a)*M .. of `out`. */
int pyftle_mlp_fourier(const float *in, int32_t d, const float *scale, const float *Bm, int32_t F,
                       int32_t M, int32_t ld_out, int32_t n_tangents, void *out, void *stream);
/* Narrow output layer Y[M,O] = H W^T (+ b on the first rows_with_bias rows), FP32. */
int pyftle_mlp_linear_out(const void *H, int64_t ldh, int32_t K, const float *W, const float *b,
                          int32_t O, int32_t M, int32_t rows_with_bias, float *Y, void *stream);
/* Its backward: dZ = (dY W) * SiLU'(Z) (bf16), and per-chunk partial dW (chunks x O x K)
 * and db (chunks x O) over chunk_rows rows each. */
int pyftle_mlp_linear_out_backward(const float *dY, const float *W, int32_t O, int32_t K, int32_t M,
                                   const void *H, const void *Z, int64_t ld, int32_t chunk_rows,
                                   float *dW_part, float *db_part, void *dZ, void *stream);
int pyftle_mlp_reduce_splits(const float *part, int32_t splits, int64_t n, float *out, void *stream);
/* Per-chunk partial column sums of a bf16 matrix (bias gradients): part (chunks x N). */
int pyftle_mlp_colsum(const void *X, int64_t ld, int32_t M, int32_t N, int32_t chunk_rows, float *part,
                      void *stream);
/* dY = 2 (Y - T) * scale; loss += sum (Y - T)^2 * scale (FP32 atomics). */
int pyftle_mlp_mse(const float *Y, const float *T, int64_t n, float scale, float *dY, float *loss,
                   void *stream);
/* Adam on FP32 master weights (bias-corrected, step >= 1), optional bf16 copy. */
int pyftle_mlp_adam(float *w, float *m, float *v, const float *g, int64_t n, float lr, float beta1,
                    float beta2, float eps, int32_t step, float weight_decay, void *w_bf16,
                    void *stream);
/* Adam with a cosine learning-rate schedule evaluated on the device from *step (the
 * number k of completed schedule steps; incremented by the call on the stream; bias
 * corrections for Adam step adam_step0 + k + 1), so a training step holds no host-side
 * scalars and can be captured in a graph. */
int pyftle_mlp_adam_scheduled(float *w, float *m, float *v, const float *g, int64_t n, float beta1,
                              float beta2, float eps, float weight_decay, float lr_max, float lr_min,
                              int32_t total_steps, int32_t adam_step0, int32_t *step, void *w_bf16,
                              void *stream);
int pyftle_mlp_cast_bf16(const float *x, int64_t n, void *y, void *stream);

/* Fused inference (Fourier features, every hidden layer and the output layer in one
 * kernel; each work-group keeps its rows' activations in LDS and streams the weights from
 * L2 into registers): 64 frequencies, hidden width 256 or 512, depth 2..8, (width,
 * outputs, tangents) in {(256, 2, 0), (256, 2, 2), (256, 3, 0), (512, 2, 0), (512, 3, 0),
 * (512, 3, 3)}.  W_packed: every layer (width x K, layer 0 with K = 128 first) as written
 * by pyftle_mlp_pack_fused (rows a multiple of 16, K of 32); W_packed, bias (depth x width)
 * and Wout (O x width) 16-byte aligned; Y (1 + n_tangents, M, O) with the output scale
 * applied. */
int pyftle_mlp_pack_fused(const float *W, int32_t rows, int32_t K, void *out, void *stream);
int pyftle_mlp_fused_forward(const float *x, int32_t d_in, int32_t M, const float *scale,
                             const float *Bf, int32_t n_freq, const void *W_packed,
                             const float *bias, int32_t depth, int32_t width, const float *Wout,
                             const float *bout, int32_t O, float out_scale, int32_t n_tangents,
                             float *Y, void *stream);

/* Fused training step (one minibatch; BF16 MFMA, FP32 accumulation).  Rows
 * *base_row .. *base_row + M - 1 of x (rows, d_in) and targets (rows, d_out; the targets
 * of the network output, i.e. already divided by the output scale).  params is the flat
 * FP32 layout of FlowMapMLP (per hidden layer W_l (width x fan_l) then b_l, then Wout
 * (d_out x width), bout); W_packed holds the same weights in bf16 as written by
 * pyftle_mlp_pack_train: the hidden layers in pyftle_mlp_pack_fused order followed by the
 * transposed layers 1..depth-1.  Writes every parameter's MSE gradient (mean over the
 * M x d_out outputs) to grads, adds the loss to *loss and advances *base_row by M, all on
 * the stream, so a step has no host-side arguments that change.  M a multiple of 64 and
 * of 128 * splits (split-K of the weight-gradient GEMMs); 64 Fourier frequencies, width
 * 256 or 512, depth 2..8, d_in <= 4, d_out <= 3; work of pyftle_mlp_train_work_bytes,
 * 256-byte aligned.  Gradients are summed in a fixed order (deterministic). */
typedef struct pyftle_mlp_train_args {
  const float *x;
  const float *targets;
  int64_t *base_row;      /* device */
  int32_t M;
  int32_t d_in, d_out, n_freq, width, depth;
  int32_t splits;
  const float *scale;     /* (d_in) */
  const float *Bf;        /* (n_freq, d_in) */
  const void *W_packed;
  const float *params;
  float *grads;
  float *loss;            /* device FP32 */
  void *work;
  int32_t *step;          /* device step index, += 1 by the call (may be NULL) */
} pyftle_mlp_train_args;
int64_t pyftle_mlp_train_work_bytes(int32_t M, int32_t width, int32_t depth, int32_t d_out, int32_t splits);
int pyftle_mlp_train_grads(const pyftle_mlp_train_args *args, void *stream);
/* W_packed for pyftle_mlp_train_grads from the flat FP32 parameters. */
int pyftle_mlp_pack_train(const float *params, int32_t width, int32_t depth, void *packed, void *stream);
/* Adam as pyftle_mlp_adam_scheduled, with the learning rate of step *step (the step's
 * index, which pyftle_mlp_train_grads advanced: start it at the first index minus one),
 * writing W_packed for the next step instead of a flat bf16 copy. */
int pyftle_mlp_train_adam(float *params, float *m, float *v, const float *grads, int32_t width, int32_t depth,
                          int32_t d_out, float beta1, float beta2, float eps, float weight_decay, float lr_max,
                          float lr_min, int32_t total_steps, int32_t adam_step0, const int32_t *step,
                          void *packed, void *stream);
/* A whole step: pyftle_mlp_train_grads followed by pyftle_mlp_train_adam (with
 * packed = args->W_packed, params = args->params, step = args->step, which must not be NULL),
 * with the Adam update applied by the gradient reduction itself: three launches instead of
 * four, the same parameters, moments, packed weights, gradients, loss, batch offset and step
 * index bit for bit.  args->params and args->W_packed are updated in place. */
int pyftle_mlp_train_step(const pyftle_mlp_train_args *args, float *m, float *v, float beta1, float beta2,
                          float eps, float weight_decay, float lr_max, float lr_min, int32_t total_steps,
                          int32_t adam_step0, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* PYFTLE_HIP_H */
