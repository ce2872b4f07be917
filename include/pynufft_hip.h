/*
 * pynufft_hip.h -- C ABI of the MI355X min-max NUFFT library (libpynufft_hip.so).
 *
 * The reference (pynufft) is pure Python and has no FFI of its own; each
 * entry point below replaces one run-time stage of the reference's NUFFT
 * class and is what a ctypes binding of that stage calls.  Citations are to
 * the reference checkout (pynufft/pynufft).
 *
 * Conventions
 *  - Plain pointers and sizes only.  Every buffer is device memory owned by
 *    the caller (PyTorch tensors in pynufft_amd); the library allocates no
 *    device memory.  Scratch space is passed in as `ws` and sized by
 *    pnh_workspace_bytes().
 *  - Complex data are interleaved (re, im): complex64 = 2 x float,
 *    complex128 = 2 x double.  Arrays are C order with the coil (batch)
 *    index fastest: image Nd x nb, grid Kd x nb, samples M x nb.
 *  - Every call is asynchronous on `stream` (a hipStream_t; NULL = default).
 *  - Return value: 0 on success, < 0 on error (-1 bad argument, -2
 *    unsupported configuration, -3 HIP error); pnh_last_error() gives the
 *    message of the calling thread's last failure.  Nothing throws across
 *    the ABI.
 *  - Distinct plans may be used from different threads concurrently.  Calls
 *    on one plan that share a workspace must be stream-ordered.
 */
#ifndef PYNUFFT_HIP_H
#define PYNUFFT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PNH_MAXD 3

enum pnh_dtype { PNH_C64 = 0, PNH_C128 = 1 };

/* Operators whose scratch needs are reported by pnh_workspace_bytes(). */
enum pnh_op {
  PNH_OP_FORWARD = 0,
  PNH_OP_ADJOINT = 1,
  PNH_OP_SELFADJOINT = 2,
  PNH_OP_Y2K = 3,
  PNH_OP_K2Y2K = 4,
  PNH_OP_SOLVE_CG = 5,
  PNH_OP_FORWARD_ONE2MANY = 6,
  PNH_OP_ADJOINT_MANY2ONE = 7,
  PNH_OP_PRECOMPUTE_SP = 8,
  PNH_OP_SELFADJOINT2 = 9,
  PNH_OP_SOLVE_DC = 10,
  PNH_OP_SOLVE_L1TVOLS = 11,
  PNH_OP_SOLVE_LSQR = 12,
  PNH_OP_SOLVE_LSMR = 13,
  PNH_OP_SOLVE_BICG = 14
};

/* Krylov solvers of pnh_solve_bicg. */
enum pnh_bicg_kind { PNH_BICGSTAB = 0, PNH_BICG = 1 };

/*
 * Plan descriptor: geometry plus device pointers to the interpolator built on
 * the host in FP64 (reference src/_helper/helper.py:620-802, plan()).  Samples
 * are held in a trajectory-sorted row order; row r is sample perm[r].
 */
typedef struct pnh_plan_desc {
  int32_t ndim;              /* 1..3 */
  int32_t batch;             /* coils sharing the trajectory (>= 1) */
  int64_t M;                 /* number of samples */
  int64_t Nd[PNH_MAXD];      /* image shape */
  int64_t Kd[PNH_MAXD];      /* oversampled grid shape */
  int32_t Jd[PNH_MAXD];      /* interpolator size per dim */

  const int32_t* perm;       /* [M] sorted row -> sample index */
  const int32_t* base;       /* [ndim][M] first support cell per dim (row order) */
  const void* u_c64;         /* complex64 per-dim factors u_d[row][j], concatenated */
  const void* u_c128;        /* complex128 factors (FP64 operators), may be NULL */
  const void* du_c64;        /* complex64 d u_d / d om_d, same layout as u_c64;
                                needed only by pnh_forward_grad_om, may be NULL */
  int64_t u_off[PNH_MAXD];   /* element offset of dim d's block */
  const void* csr_c64;       /* complex64 full coefficients (may be NULL), in
                                slices of S rows: [ceil(M/S)][G][S][J_last],
                                G = prod(Jd[:-1]), S = 64 / L where
                                L = J_last (<= 4 or 6), 8 (<= 8) or 16 */
  const int32_t* csr_col;    /* int32 grid cell of each entry, same layout */

  const float* sn_f32;       /* [prod Nd] apodization (reference st['sn']) */
  const double* sn_f64;      /* [prod Nd] FP64 apodization, may be NULL */
  const void* tw_c64[PNH_MAXD];   /* exp(-2 pi i t / Kd[d]), t < Kd[d] */
  const void* tw_c128[PNH_MAXD];  /* same in complex128, may be NULL */

  /* Gridding (y2k) work list.  Rows are sorted by the tile that holds their
   * first support cell (C order over the tile grid), so a tile owns one
   * contiguous run of rows, cut into items of at most 4096 rows.  Per axis
   * either tile >= Kd (one tile spans the axis) or tile >= Jd - 1 and the
   * last tile's extent Kd - (ceil(Kd / tile) - 1) * tile >= Jd - 1. */
  int32_t tile[PNH_MAXD];    /* gridding tile extent per dim */
  double u_bound;            /* >= |coefficient| for every row and support offset */
  int64_t n_items;           /* gridding work items */
  const int32_t* items;      /* [n_items][4]: tile id, first row, rows, partial slot */
  int64_t n_tiles;           /* prod over dims of ceil(Kd / tile) */
  const int32_t* tile_slots; /* [n_tiles][2]: first partial slot, slot count; the
                                slots of one tile are consecutive, in row order */
  int64_t n_slots;           /* partial slots (one per item) */
  int32_t max_tile_slots;    /* largest slot count of any tile (sizes the gather) */
  int32_t threads;           /* gridding workgroup size: multiple of 64 in
                                64..1024, or 0 to let the library choose */
  /* 1: pnh_forward, pnh_adjoint and pnh_selfadjoint (and the operators
   * built on them) keep complex64 inputs and outputs but run the
   * apodization, FFTs and interpolation in FP64 (needs u_c128, tw_c128 and
   * sn_f64).  For plans whose apodization spans a large range
   * (max|sn| / min|sn|, e.g. Kd == Nd), where the interpolation cancels
   * large terms and FP32 arithmetic loses digits that the reference's
   * complex128 CSR products (nufft/_nufft_class_methods_cpu.py:269-314) and
   * complex128 ifftn (:316-329) keep.  0: FP32 arithmetic. */
  int32_t fp64_ops;
  /* [n_items] per-item bounds written by pnh_item_bounds for this work list
   * (may be NULL).  With them the complex64 gridding keeps one 64-bit
   * fixed-point word per cell (real and imaginary parts packed) instead of
   * two: half the LDS atomics and accumulator.  No counterpart in the
   * reference (an implementation choice of the gridding). */
  const double* item_bound;
  /* >= max over rows of sum_j |coefficient(row, j)| over the row's support
   * (for per-axis factors prod_d sum_j |u_d[row][j]|), or 0.  With it the
   * FP64 P^H P (pnh_k2y2k in complex128, and the solvers) grids each work
   * item in one pass over its rows, reading each row's factors once; the
   * fixed-point exponent then comes from this bound and the largest input
   * value of the item's box instead of the largest sample.  No counterpart
   * in the reference (an implementation choice of the gridding). */
  double row_bound;
} pnh_plan_desc;

typedef struct pnh_plan pnh_plan;

/* Library version string. */
const char* pnh_version(void);
/* Message of the calling thread's last error ("" if none). */
const char* pnh_last_error(void);
/* Number of visible HIP devices, or < 0 on error.
 * Replaces reference src/_helper/helper.py:1172-1224 (device_list discovery). */
int pnh_device_count(void);

/* Validate a descriptor and build the per-axis FFT passes.  The descriptor's
 * pointers must stay valid for the plan's lifetime.
 * Replaces the offload step of reference nufft/_nufft_class_methods_device.py:183-254. */
int pnh_plan_create(const pnh_plan_desc* desc, pnh_plan** out);
int pnh_plan_destroy(pnh_plan* plan);
/* Bytes of `ws` needed by operator `op` in precision `dtype`. */
int pnh_workspace_bytes(const pnh_plan* plan, int op, int dtype, size_t* bytes);
/* bound[i] (device double[n_items]) = for work item i, the largest over the
 * cells of its padded tile of the sum over its rows of the product of the
 * per-axis complex64 factor magnitudes, rounded up.  A constant of the
 * trajectory and work list: computed once after planning and passed back in
 * pnh_plan_desc.item_bound of a new plan.  No counterpart in the reference
 * (it sizes the fixed-point scale of the packed gridding accumulators). */
int pnh_item_bounds(const pnh_plan* plan, double* bound, void* stream);

/* xx = x * sn.  Reference nufft/_nufft_class_methods_cpu.py:224-230 (_x2xx_cpu). */
int pnh_x2xx(const pnh_plan* plan, const void* x, void* xx, void* stream);
/* x = xx * sn.  Reference nufft/_nufft_class_methods_cpu.py:346-351 (_xx2x_cpu). */
int pnh_xx2x(const pnh_plan* plan, const void* xx, void* x, void* stream);
/* k = fftn(zero_pad(xx)) over all axes (front-corner pad), complex64.
 * Reference nufft/_nufft_class_methods_cpu.py:232-249 (_xx2k_cpu). */
int pnh_xx2k(const pnh_plan* plan, const void* xx, void* k, void* stream);
/* xx = crop(ifftn(k)); k is overwritten (transformed in place, as
 * reference nufft/_nufft_class_methods_device.py:432 does).
 * Reference nufft/_nufft_class_methods_cpu.py:316-329 (_k2xx_cpu). */
int pnh_k2xx(const pnh_plan* plan, void* k, void* xx, void* stream);
/* y = P k.  dtype selects complex64 (fp32) or complex128 (FP64) operands.
 * Reference nufft/_nufft_class_methods_cpu.py:269-288 (_k2y_cpu). */
int pnh_k2y(const pnh_plan* plan, int dtype, const void* k, void* y, void* stream);
/* k = P^H y (deterministic binned gridding).
 * Reference nufft/_nufft_class_methods_cpu.py:290-314 (_y2k_cpu). */
int pnh_y2k(const pnh_plan* plan, int dtype, const void* y, void* k, void* ws,
            void* stream);
/* k_out = P^H P k.  Reference nufft/_nufft_class_methods_cpu.py:353-362 (_k2y2k_cpu). */
int pnh_k2y2k(const pnh_plan* plan, int dtype, const void* k, void* k_out,
              void* ws, void* stream);
/* y = k2y(xx2k(x2xx(x))), complex64.
 * Reference nufft/_nufft_class_methods_cpu.py:168-179 (_forward_cpu). */
int pnh_forward(const pnh_plan* plan, const void* x, void* y, void* ws, void* stream);
/* x = xx2x(k2xx(y2k(y))), complex64.
 * Reference nufft/_nufft_class_methods_cpu.py:181-193 (_adjoint_cpu). */
int pnh_adjoint(const pnh_plan* plan, const void* y, void* x, void* ws, void* stream);
/* x_out = xx2x(k2xx(k2y2k(xx2k(x2xx(x))))), complex64.
 * Reference nufft/_nufft_class_methods_cpu.py:201-214 (_selfadjoint_cpu). */
int pnh_selfadjoint(const pnh_plan* plan, const void* x, void* x_out, void* ws,
                    void* stream);
/* Gradient of a real loss with respect to the sample coordinates, for
 * y = k2y(k): om_grad[m][d] = sum_c Re(conj(g[m][c]) * dy[m][c] / d om[m][d]),
 * where g = dL/d conj(y) (torch's convention), k is the oversampled spectrum
 * xx2k(x2xx(x)) of the forward pass and om_grad is float32 [M][ndim] in the
 * original sample order.  The derivative goes through the per-axis min-max
 * factors (product rule across axes); the support cells are piecewise
 * constant in om.  No counterpart in the reference (new); it backs the
 * trajectory gradient of pynufft_amd.autograd. */
int pnh_forward_grad_om(const pnh_plan* plan, const void* k, const void* g,
                        float* om_grad, void* stream);
/* Multi-coil operators of a plan with batch = ncoil; sense is the coil
 * sensitivity array, Nd x ncoil complex64 with the coil index fastest.
 * forward_one2many: y[:, c] = forward(x * sense[..., c]) for one Nd image x
 * (reference linalg/nufft_cpu.py:177-189; device kernel cDistribute,
 * src/re_subroutine.py:355-374).
 * adjoint_many2one: x = mean over c of adjoint(y)[..., c] * conj(sense[..., c])
 * (reference linalg/nufft_cpu.py:191-203, numpy.mean over the coil axis;
 * device kernel cMerge, src/re_subroutine.py:378-439). */
int pnh_forward_one2many(const pnh_plan* plan, const void* x, const void* sense, void* y,
                         void* ws, void* stream);
int pnh_adjoint_many2one(const pnh_plan* plan, const void* y, const void* sense, void* x,
                         void* ws, void* stream);
/* Toeplitz-style self-adjoint (complex64).  precompute_sp writes
 * W = |xx2k(adjoint(1))| (first coil) as float32 [prod(Kd)]; selfadjoint2
 * returns k2xx(W * xx2k(x)) with no apodization, each coil weighted by W.
 * Reference nufft/_nufft_class_methods_cpu.py:126-145 (_precompute_sp_cpu)
 * and :216-222 (_selfadjoint2_cpu). */
int pnh_precompute_sp(const pnh_plan* plan, float* W, void* ws, void* stream);
int pnh_selfadjoint2(const pnh_plan* plan, const void* x, const float* W, void* x_out,
                     void* ws, void* stream);
/* solve(y, 'dc', maxiter): pipe density W <- W / forward(adjoint(W)) from
 * W = 1, maxiter times, then x = adjoint(W * y) (per coil).  As in the
 * reference, W is complex128 and the interpolation runs in FP64 (the plan
 * needs u_c128); the FFTs are complex64 forward, complex128 inverse.
 * Reference linalg/solve_cpu.py:165-183 (_pipe_density) and :214-235. */
int pnh_solve_dc(const pnh_plan* plan, const void* y, void* x, int maxiter, void* ws,
                 void* stream);
/* solve(y, 'L1TVOLS', maxiter, rho): L1 total-variation regularised least
 * squares by split Bregman with mu = 1, lambda = rho (complex64, batch 1).
 * lap: device complex64 [prod Kd], the Laplacian spectrum of the reference's
 * create_laplacian_kernel (the complex64 fftn of the periodic stencil; the
 * solve is sensitive to its rounding), or NULL for the closed form
 * sum_d 2 cos(2 pi k_d / K_d) - 2.
 * Reference linalg/solve_cpu.py:42-163 (L1TVOLS), src/_helper/helper.py:11-70
 * (Laplacian kernel, periodic difference maps). */
int pnh_solve_l1tvols(const pnh_plan* plan, const void* y, void* x, int maxiter, double rho,
                      const void* lap, void* ws, void* stream);
/* solve(y, 'lsqr', ...): scipy.sparse.linalg.lsqr on A = P (k2y, adjoint
 * y2k) with k-space unknowns, x0 = 0, in FP64 on the device (batch 1), then
 * x = crop(ifftn(k)) / sn.  iter_lim <= 0 means scipy's default 2 prod(Kd).
 * info_out (device int32[2], may be NULL) receives istop and the iteration
 * count.  Reference linalg/solve_cpu.py:236-256 and scipy's lsqr. */
int pnh_solve_lsqr(const pnh_plan* plan, const void* y, void* x, int iter_lim, double damp,
                   double atol, double btol, double conlim, void* ws, int32_t* info_out,
                   void* stream);
/* solve(y, 'lsmr', ...): scipy.sparse.linalg.lsmr on A = P, as
 * pnh_solve_lsqr (maxiter <= 0 means scipy's default min(M, prod(Kd))).
 * Reference linalg/solve_cpu.py:236-256 and scipy's lsmr. */
int pnh_solve_lsmr(const pnh_plan* plan, const void* y, void* x, int maxiter, double damp,
                   double atol, double btol, double conlim, void* ws, int32_t* info_out,
                   void* stream);
/* solve(y, 'bicgstab' | 'bicg', ...): scipy.sparse.linalg.bicgstab / bicg
 * (no preconditioner, x0 = 0) on the k-space normal equations
 * P^H P k = P^H y in FP64 on the device (batch 1); stops when
 * ||r|| < max(atol, rtol ||b||) at the start of an iteration (bicgstab also
 * after its half step), on scipy's breakdown tests, or after maxiter
 * (<= 0: scipy's default 10 prod(Kd)); then x = crop(ifftn(k)) / sn.
 * info_out (device int32[2], may be NULL) receives scipy's info and the
 * iterations run.  Reference linalg/solve_cpu.py:256-288 with scipy's
 * bicgstab / bicg. */
int pnh_solve_bicg(const pnh_plan* plan, int kind, const void* y, void* x, int maxiter,
                   double rtol, double atol, void* ws, int32_t* info_out, void* stream);
/* solve(y, 'gmres', restart=, maxiter=, rtol=, atol=): scipy's restarted
 * GMRES (modified Gram-Schmidt, Givens rotations, no preconditioner) on
 * P^H P in FP64 on the device (batch 1); restart 1..64 (<= 0: scipy's 20,
 * capped at prod(Kd)), maxiter = restart cycles (<= 0: 10 prod(Kd)).  Its
 * scratch depends on restart: pnh_gmres_workspace_bytes.  info_out (device
 * int32[2], may be NULL) receives scipy's info and the inner iterations.
 * Reference linalg/solve_cpu.py:256-288 with scipy's gmres. */
int pnh_gmres_workspace_bytes(const pnh_plan* plan, int restart, size_t* bytes);
int pnh_solve_gmres(const pnh_plan* plan, const void* y, void* x, int restart, int maxiter,
                    double rtol, double atol, void* ws, int32_t* info_out, void* stream);
/* solve(y, 'lgmres', maxiter=, inner_m=, outer_k=, rtol=, atol=): scipy's
 * LGMRES (GMRES cycles of inner_m Krylov columns augmented with the last
 * outer_k corrections, stored with their products A z; no preconditioner)
 * on P^H P in FP64 on the device (batch 1); inner_m + outer_k <= 64,
 * outer_k <= 8, maxiter outer iterations.  info_out (device int32[2], may be
 * NULL) receives scipy's info and the inner iterations.
 * Reference linalg/solve_cpu.py:256-288 with scipy's lgmres. */
int pnh_lgmres_workspace_bytes(const pnh_plan* plan, int inner_m, int outer_k, size_t* bytes);
int pnh_solve_lgmres(const pnh_plan* plan, const void* y, void* x, int maxiter, int inner_m,
                     int outer_k, double rtol, double atol, void* ws, int32_t* info_out,
                     void* stream);
/* Conjugate-gradient solve on the k-space normal equations P^H P k = P^H y in
 * FP64 (complex128), x0 = 0, stop when ||r|| < max(rtol ||b||, atol) at the
 * start of an iteration or after maxiter iterations; result
 * x = crop(ifftn(k)) / sn cast to complex64.  The loop runs on the device with
 * no host synchronisation; iters_out (device int32[batch], may be NULL)
 * receives the iteration count per coil and resid_out (device double[batch],
 * may be NULL) the final ||r|| per coil: the value of the stopping test
 * when it held, else the norm after the last iteration (||b|| for
 * maxiter = 0).
 * Reference linalg/solve_cpu.py:256-288 with scipy.sparse.linalg.cg. */
int pnh_solve_cg(const pnh_plan* plan, const void* y, void* x, int maxiter,
                 double rtol, double atol, void* ws, int32_t* iters_out,
                 double* resid_out, void* stream);
/* pnh_solve_cg without a workspace argument and without any allocation: the
 * scratch is a static device array of PNH_NOALLOC_SCRATCH_BYTES in the
 * library, one per device, held until the process ends.  A solve whose
 * pnh_workspace_bytes(PNH_OP_SOLVE_CG) exceeds it fails with -1 (the radial
 * 2-D 256^2 / 512^2 solve, M 205 824, needs 27 MiB).  Every call on a device
 * shares that device's array, so calls that overlap on one device (other
 * streams, other threads) race; calls ordered on one stream do not.  The
 * plan's complex128 factors must exist before the call.  No counterpart in
 * the reference (new): for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). */
#define PNH_NOALLOC_SCRATCH_BYTES (32u << 20)
int pnh_solve_cg_noalloc(const pnh_plan* plan, const void* y, void* x, int maxiter,
                         double rtol, double atol, int32_t* iters_out,
                         double* resid_out, void* stream);

/* ---- many independent problems (uneven Nd, Kd, M; same ndim and Jd) ----
 * The reference handles one problem per NUFFT object; a batch here runs every
 * stage of all problems in one launch.  Vectors are packed: problem p's image,
 * samples and grid follow problem p-1's (complex64 for forward/adjoint and the
 * solve's y/x).  Each problem is described by the same pnh_plan_desc as a
 * single plan (batch == 1, factor format, complex64 and complex128 factors);
 * the per-problem and per-stage work tables live in `table`, a device buffer
 * of pnh_batch_table_bytes() bytes written once by pnh_batch_create (which
 * synchronizes `stream`).  No counterpart in the reference (new). */
typedef struct pnh_batch pnh_batch;
typedef struct pnh_graph pnh_graph;
int pnh_batch_table_bytes(const pnh_plan_desc* descs, int n, size_t* bytes);
int pnh_batch_create(const pnh_plan_desc* descs, int n, void* table, void* stream,
                     pnh_batch** out);
int pnh_batch_destroy(pnh_batch* batch);
/* Packed totals: sum of prod(Nd), of M and of prod(Kd). */
int pnh_batch_sizes(const pnh_batch* batch, int64_t* n_image, int64_t* n_samples,
                    int64_t* n_grid);
/* Scratch for PNH_OP_FORWARD, PNH_OP_ADJOINT or PNH_OP_SOLVE_CG. */
int pnh_batch_workspace_bytes(const pnh_batch* batch, int op, int dtype, size_t* bytes);
/* Batched forward / adjoint (complex64), one launch per stage for all problems. */
int pnh_batch_forward(const pnh_batch* batch, const void* x, void* y, void* ws, void* stream);
int pnh_batch_adjoint(const pnh_batch* batch, const void* y, void* x, void* ws, void* stream);
/* Batched FP64 CG, semantics of pnh_solve_cg per problem (iters_out: int32[n],
 * resid_out: double[n], each may be NULL). */
int pnh_batch_solve_cg(const pnh_batch* batch, const void* y, void* x, int maxiter,
                       double rtol, double atol, void* ws, int32_t* iters_out,
                       double* resid_out, void* stream);
/* Capture pnh_batch_solve_cg with these buffers into a HIP graph, and replay
 * it on any stream; the buffers must outlive the graph. */
int pnh_batch_cg_graph_create(const pnh_batch* batch, const void* y, void* x, int maxiter,
                              double rtol, double atol, void* ws, int32_t* iters_out,
                              double* resid_out, pnh_graph** out);
/* Capture pnh_solve_cg of one plan with these buffers into a HIP graph
 * (every iteration launched; converged coils skip their work on the device);
 * replay with pnh_graph_launch.  No counterpart in the reference (new): for
 * repeated solves on one trajectory without per-kernel launch cost. */
int pnh_cg_graph_create(const pnh_plan* plan, const void* y, void* x, int maxiter,
                        double rtol, double atol, void* ws, int32_t* iters_out,
                        double* resid_out, pnh_graph** out);
int pnh_graph_launch(pnh_graph* graph, void* stream);
int pnh_graph_destroy(pnh_graph* graph);

#ifdef __cplusplus
}
#endif
#endif /* PYNUFFT_HIP_H */
