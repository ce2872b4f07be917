// Element-wise and CG vector kernels (host launch wrappers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PNH_RED_BLOCKS 512
#define PNH_CG_MAXNB 64

enum { CG_PHASE_INIT = 0, CG_PHASE_PQ = 1, CG_PHASE_RR = 2 };

// Device-resident CG state, one entry per coil (independent systems).  rr,
// done and all_done have two parity buffers: iteration `it` reads buffer
// it&1, which its p update writes from buffer (it+1)&1 (the init phase fills
// buffer 1).
// res and stopped are written once per coil, when its stopping test first
// holds: ||r|| then, and 1.
struct CgScalars {
  double tol[PNH_CG_MAXNB];
  double rr[2][PNH_CG_MAXNB];
  double res[PNH_CG_MAXNB];
  int done[2][PNH_CG_MAXNB];
  int iters[PNH_CG_MAXNB];
  int stopped[PNH_CG_MAXNB];
  int all_done[2];
};

int pnh_scale_sn(int fp64, const void* in, void* out, const void* sn,
                 int64_t n_img, int nb, hipStream_t s);
int pnh_c64_to_c128(const void* in, void* out, int64_t n, hipStream_t s);
// complex128 -> complex64 (round to nearest per part)
int pnh_c128_to_c64(const void* in, void* out, int64_t n, hipStream_t s);
int pnh_cast_div_sn(const void* in, void* out, const float* sn, int64_t n_img,
                    int nb, hipStream_t s);
// out[p*nb + c] = x[p] * sense[p*nb + c] (complex64)
int pnh_coil_expand(const void* x, const void* sense, void* out, int64_t n_img, int nb,
                    hipStream_t s);
// out[p] = mean_c xc[p*nb + c] * conj(sense[p*nb + c]) (complex64)
int pnh_coil_combine(const void* xc, const void* sense, void* out, int64_t n_img, int nb,
                     hipStream_t s);
int pnh_zero(int fp64, void* p, int64_t n, hipStream_t s);
int pnh_cg_dot(int mode, const void* a, const void* b, int64_t n_cells, int nb,
               void* part, const int* skip, hipStream_t s);
// tolerance, first rr and stopping state from the |b|^2 block partials
int pnh_cg_init(const void* part, int nb, CgScalars* sc, int maxiter, double rtol,
                double atol, hipStream_t s);
// iteration it: rr from the previous |r|^2 partials rr_part (it > 0), the
// stopping test, then p = r or p = r + beta p
int pnh_cg_p_update(void* p, const void* r, int64_t n_cells, int nb, CgScalars* sc,
                    const void* rr_part, int it, hipStream_t s);
// alpha = rr / (p.q) with p.q summed from the block partials pq_part (every
// workgroup sums them itself); x += alpha p; r -= alpha q; |r|^2 partials
int pnh_cg_xr_update(void* x, void* r, const void* p, const void* q, int64_t n_cells,
                     int nb, const CgScalars* sc, int it, const void* pq_part, void* part,
                     hipStream_t s);
// out[c] = ||r|| of coil c at the end of the solve: the value of its
// stopping test, or (a coil that ran out of iterations) the norm summed from
// the |r|^2 partials rr_part of the last x/r update (of b when maxiter = 0)
int pnh_cg_resid(const void* rr_part, int nb, const CgScalars* sc, double* out,
                 hipStream_t s);
