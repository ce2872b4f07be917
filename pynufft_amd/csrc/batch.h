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

// Device-resident CG state of a batch (arrays of n_problems entries).
struct BatchCg {
  double* rr;
  double* rr_prev;
  double* tol;
  double2* pq;
  int* done;
  int* iters;
  int* all_done;
  double2* part;             // per-block partial sums
};

// Interpolates every problem's items (tile ownership, padded boxes in LDS).
int pnh_batch_k2y_launch(const BatchProblem* probs, const int2* items, int64_t n_items, int D,
                         const int* Jd, int fp64, int64_t pad_cells, int threads, const void* k,
                         void* y, const int* done, const int* all_done, hipStream_t s);
// Grids every problem's items into partial slots, then gathers the grids
// (one block per tile, by table `gather`).
// k_out = P^H P k for every problem (fused k2y + y2k per work item, then the
// gather); y receives P k.
int pnh_batch_k2y2k_launch(const BatchProblem* probs, const int2* items, int64_t n_items,
                           const BatchTable& gather, int D, const int* Jd, int fp64,
                           int64_t pad_cells, int threads, const void* k, void* y, void* k_out,
                           void* part, const int* done, const int* all_done, hipStream_t s);
int pnh_batch_y2k_launch(const BatchProblem* probs, const int2* items, int64_t n_items,
                         const BatchTable& gather, int D, const int* Jd, int fp64,
                         int64_t pad_cells, int threads, const void* y, void* k, void* part,
                         const int* done, const int* all_done, hipStream_t s);
int pnh_batch_fft_launch(const BatchProblem* probs, const BatchTable& t, int axis,
                         int kind, int fp64, int pow2, size_t lds, int threads,
                         const void* in, int in_image, void* out, int out_image,
                         const int* done, const int* all_done, hipStream_t s);
int pnh_batch_vec_launch(int op, const BatchProblem* probs, const BatchTable& t,
                         int64_t chunk, void* a, void* b, void* c, void* d,
                         const BatchCg& cg, int first_iter, hipStream_t s);
int pnh_batch_finalize_launch(const BatchTable& t, int n_problems, const BatchCg& cg,
                              int phase, int iter, int maxiter, double rtol, double atol,
                              hipStream_t s);

enum { BFFT_FWD = 0, BFFT_INV = 1, BFFT_INV64 = 2 };
enum { BVEC_CAST_Y = 0, BVEC_ZERO_X = 1, BVEC_DOT_RR = 2, BVEC_P_UPDATE = 3,
       BVEC_DOT_PQ = 4, BVEC_XR_UPDATE = 5, BVEC_CAST_DIV = 6 };
