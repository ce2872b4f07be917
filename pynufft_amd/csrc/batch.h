// Batched independent problems: per-problem descriptors and work tables.
#pragma once
#include "common.h"
#include "fft.h"
#include "interp.h"

// One problem of a batch, resident in device memory.  Vector offsets are in
// complex elements of the packed batch vectors (problems back to back).
struct BatchProblem {
  InterpArgs ia;             // complex64 factors
  InterpArgs ia64;           // complex128 factors
  BinArgs bins;              // this problem's gridding work list
  int64_t x_off, y_off, k_off;
  int64_t part_off;          // first element of this problem's partial slots
  int64_t n_img, n_grid;     // prod(Nd), prod(Kd)
  const float* sn32;
  FftDesc fwd[PNH_MAXD];     // forward passes, fp32, x2xx fused in the first
  FftDesc inv[PNH_MAXD];     // inverse passes, fp32, xx2x fused in the last
  FftDesc inv64[PNH_MAXD];   // inverse passes, FP64, no apodization
  const void* tw32[PNH_MAXD];
  const void* tw64[PNH_MAXD];
};

// Per-launch table: block b of the launch works on problem tab[b] as its
// (b - first[tab[b]])-th block.
struct BatchTable {
  const int* tab;            // [n_blocks]
  const int* first;          // [n_problems]
  int64_t n_blocks;
};

// Device-resident CG state of a batch (arrays of n_problems entries).  rr
// and done have two parity buffers ([2][n]): iteration `it` reads buffer
// it&1, which its p update writes from buffer (it+1)&1, so no workgroup
// reads what another writes in the same launch.  stopped / res / iters are
// written once per problem, when its stopping test first holds.  The block
// partials are per block of the grid table; every workgroup of a problem
// sums its problem's partials itself, in one fixed order.
struct BatchCg {
  double* tol;
  double* rr;                // [2][n]
  int* done;                 // [2][n]
  int* iters;
  int* stopped;
  double* res;
  double2* part_rr;          // |r|^2 (|b|^2 before the first iteration)
  double2* part_pq;          // p . q
  int n;
  int maxiter;
  double rtol, atol;
};

// Interpolates every problem's items (tile ownership, padded boxes in LDS).
int pnh_batch_k2y_launch(const BatchProblem* probs, const int2* items, int64_t n_items, int D,
                         const int* Jd, int fp64, int64_t pad_cells, int threads, const void* k,
                         void* y, const int* done, const int* all_done, hipStream_t s);
// Grids every problem's items into partial slots, then gathers the grids
// (one block per tile, by table `gather`).
// k_out = P^H P k for every problem (fused k2y + y2k per work item, then the
// gather); y receives P k (in row order within each problem).
// fixed_pad: every problem's work list passes pnh_fixed_pad (boxes of the
// compile-time extent PNH_FIXED_PAD) and has a row bound.
int pnh_batch_k2y2k_launch(const BatchProblem* probs, const int2* items, int64_t n_items,
                           const BatchTable& gather, int D, const int* Jd, int fp64,
                           int64_t pad_cells, int threads, int fixed_pad, const void* k, void* y,
                           void* k_out, void* part, const int* done, const int* all_done,
                           hipStream_t s);
// packed: every problem has per-item bounds (complex64 packed accumulators).
int pnh_batch_y2k_launch(const BatchProblem* probs, const int2* items, int64_t n_items,
                         const BatchTable& gather, int D, const int* Jd, int fp64, int packed,
                         int64_t pad_cells, int threads, const void* y, void* k, void* part,
                         const int* done, const int* all_done, hipStream_t s);
int pnh_batch_fft_launch(const BatchProblem* probs, const BatchTable& t, int axis,
                         int kind, int fp64, int pow2, size_t lds, int threads,
                         const void* in, int in_image, void* out, int out_image,
                         const int* done, const int* all_done, hipStream_t s);
// One element-wise / CG step over every problem's packed vector (op
// BVEC_*); `it` is the CG iteration of the CG steps.
int pnh_batch_vec_launch(int op, const BatchProblem* probs, const BatchTable& t,
                         int64_t chunk, void* a, void* b, void* c, void* d,
                         const BatchCg& cg, int it, hipStream_t s);
// Iterations run and final ||r|| per problem (either output may be null);
// t is the grid table whose block partials the last CG step wrote.
int pnh_batch_cg_out_launch(const BatchTable& t, const BatchCg& cg, int32_t* iters_out,
                            double* resid_out, hipStream_t s);

enum { BFFT_FWD = 0, BFFT_INV = 1, BFFT_INV64 = 2 };
enum { BVEC_CAST_Y = 0, BVEC_ZERO_X = 1, BVEC_DOT_RR = 2, BVEC_P_UPDATE = 3,
       BVEC_DOT_PQ = 4, BVEC_XR_UPDATE = 5, BVEC_CAST_DIV = 6 };
