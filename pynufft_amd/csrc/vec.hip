// Element-wise apodization, dtype casts and the FP64 conjugate-gradient
// vector kernels.  Reductions are deterministic: a fixed grid of PNH_RED_BLOCKS
// blocks, each summing a fixed strided subset in a fixed order and reducing
// in LDS, then one block summing the partials in block order.
#include "common.h"
#include "vec.h"

namespace {

template <typename T>
__global__ void scale_sn_kernel(const void* __restrict__ in_, void* __restrict__ out_,
                                const T* __restrict__ sn, int64_t n_img, int nb) {
  typedef typename cplx<T>::t C;
  const C* in = (const C*)in_;
  C* out = (C*)out_;
  const int64_t n = n_img * nb;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T s = sn[i / nb];
    C v = in[i];
    v.x *= s; v.y *= s;
    out[i] = v;
  }
}

__global__ void c64_to_c128_kernel(const float2* __restrict__ in, double2* __restrict__ out,
                                   int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 v = in[i];
    out[i] = make_double2((double)v.x, (double)v.y);
  }
}

__global__ void c128_to_c64_kernel(const double2* __restrict__ in, float2* __restrict__ out,
                                   int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = in[i];
    out[i] = make_float2((float)v.x, (float)v.y);
  }
}

// x = complex64(xx) / sn (numpy complex64 division by a real-valued
// complex64: both parts multiplied by 1/sn in single precision).
__global__ void cast_div_sn_kernel(const double2* __restrict__ in, float2* __restrict__ out,
                                   const float* __restrict__ sn, int64_t n_img, int nb) {
  const int64_t n = n_img * nb;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float inv = 1.0f / sn[i / nb];
    const double2 v = in[i];
    out[i] = make_float2((float)v.x * inv, (float)v.y * inv);
  }
}

// out[p*nb + c] = x[p] * sense[p*nb + c]: one image broadcast to the coils
// (complex64 product as numpy forms it: (ac - bd, ad + bc)).
__global__ void coil_expand_kernel(const float2* __restrict__ x, const float2* __restrict__ sense,
                                   float2* __restrict__ out, int64_t n_img, int nb) {
  const int64_t n = n_img * nb;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 v = x[i / nb], s = sense[i];
    out[i] = make_float2(v.x * s.x - v.y * s.y, v.x * s.y + v.y * s.x);
  }
}

// out[p] = (sum_c xc[p*nb + c] * conj(sense[p*nb + c])) / nb: the coil
// images combined with the conjugate sensitivities and averaged over coils
// (complex64, coils summed in order).
__global__ void coil_combine_kernel(const float2* __restrict__ xc,
                                    const float2* __restrict__ sense, float2* __restrict__ out,
                                    int64_t n_img, int nb) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n_img;
       p += (int64_t)gridDim.x * blockDim.x) {
    float2 acc = make_float2(0.f, 0.f);
    for (int c = 0; c < nb; ++c) {
      const float2 v = xc[p * nb + c], s = sense[p * nb + c];
      acc.x += v.x * s.x + v.y * s.y;
      acc.y += v.y * s.x - v.x * s.y;
    }
    out[p] = make_float2(acc.x / (float)nb, acc.y / (float)nb);
  }
}

template <typename C>
__global__ void zero_kernel(C* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    C z; z.x = 0; z.y = 0; p[i] = z;
  }
}

// --- deterministic reductions over K cells x nb coils ---------------------
// mode 0: sum conj(a) b (complex); mode 1: sum |a|^2 (real part only)
template <int MODE>
__global__ void __launch_bounds__(256)
dot_partials_kernel(const double2* __restrict__ a, const double2* __restrict__ b,
                    int64_t n_cells, int nb, double2* __restrict__ part,
                    const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double2 red[256];
  for (int c = 0; c < nb; ++c) {
    double sx = 0.0, sy = 0.0;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n_cells;
         g += (int64_t)gridDim.x * blockDim.x) {
      const double2 u = a[g * nb + c];
      if (MODE == 0) {
        const double2 v = b[g * nb + c];
        sx += u.x * v.x + u.y * v.y;
        sy += u.x * v.y - u.y * v.x;
      } else {
        sx += u.x * u.x + u.y * u.y;
      }
    }
    red[threadIdx.x] = make_double2(sx, sy);
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) {
        red[threadIdx.x].x += red[threadIdx.x + s].x;
        red[threadIdx.x].y += red[threadIdx.x + s].y;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * nb + c] = red[0];
    __syncthreads();
  }
}

// Sum of the block partials of coil `c` (nblk blocks) in block order with a
// fixed tree over 256 threads; every thread of the block gets the total.
// Every caller uses this one routine, so each gets the same bits.
__device__ double2 sum_partials(const double2* __restrict__ part, int nblk, int nb, int c,
                                double2* red) {
  double sx = 0.0, sy = 0.0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
    const double2 v = part[(int64_t)i * nb + c];
    sx += v.x; sy += v.y;
  }
  red[threadIdx.x] = make_double2(sx, sy);
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      red[threadIdx.x].x += red[threadIdx.x + st].x;
      red[threadIdx.x].y += red[threadIdx.x + st].y;
    }
    __syncthreads();
  }
  const double2 tot = red[0];
  __syncthreads();   // red may be reused right away
  return tot;
}

// Init phase of the CG: |b|^2 from the block partials of coil `c` (one block
// per coil) gives the tolerance and the first rr, in parity buffer 1 (the
// one the first p update reads).
__global__ void __launch_bounds__(256)
cg_init_kernel(const double2* __restrict__ part, int nblk, int nb,
               CgScalars* __restrict__ s, int maxiter, double rtol, double atol) {
  __shared__ double2 red[256];
  const int c = blockIdx.x;
  const double2 tot = sum_partials(part, nblk, nb, c, red);
  if (threadIdx.x != 0) return;
  const double bn = sqrt(tot.x);
  const double tol = fmax(atol, rtol * bn);
  s->tol[c] = tol;
  s->rr[1][c] = tot.x;
  const int done = (bn == 0.0) || (bn < tol);
  s->done[1][c] = done;
  s->iters[c] = done ? 0 : maxiter;
  s->stopped[c] = done;
  s->res[c] = bn;
  if (nb == 1) s->all_done[1] = done;   // one coil: no separate AND pass
}

// Final residual norm per coil (one block per coil).
__global__ void __launch_bounds__(256)
cg_resid_kernel(const double2* __restrict__ part, int nblk, int nb,
                const CgScalars* __restrict__ s, double* __restrict__ out) {
  __shared__ double2 red[256];
  const int c = blockIdx.x;
  if (s->stopped[c]) {
    if (threadIdx.x == 0) out[c] = s->res[c];
    return;
  }
  const double2 tot = sum_partials(part, nblk, nb, c, red);
  if (threadIdx.x == 0) out[c] = sqrt(tot.x);
}

// all_done[1] = AND over coils after the init phase (single thread)
__global__ void cg_all_done_kernel(CgScalars* __restrict__ s, int nb) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int all = 1;
  for (int c = 0; c < nb; ++c) all &= s->done[1][c];
  s->all_done[1] = all;
}

// Iteration `it` of the CG, first step.  Reads the state of parity buffer
// (it+1)&1 and writes buffer it&1, which the rest of the iteration reads, so
// no workgroup reads what another writes.  For it > 0 every workgroup sums
// the |r|^2 block partials of the previous x/r update itself (one fixed order
// and tree): rr_new, the stopping test and beta = rr_new / rr_old; workgroup 0
// stores the new state.  Then p = r (it == 0) or p = r + beta p.
__global__ void __launch_bounds__(256)
cg_p_update_kernel(double2* __restrict__ p, const double2* __restrict__ r,
                   int64_t n_cells, int nb, CgScalars* __restrict__ s,
                   const double2* __restrict__ rr_part, int it) {
  __shared__ double2 red[256];
  __shared__ double beta[PNH_CG_MAXNB];
  __shared__ int skip[PNH_CG_MAXNB];
  const int ob = (it + 1) & 1, nbuf = it & 1;
  const bool writer = blockIdx.x == 0 && threadIdx.x == 0;
  if (s->all_done[ob]) {
    if (writer) s->all_done[nbuf] = 1;
    return;
  }
  for (int c = 0; c < nb; ++c) {
    const int dold = s->done[ob][c];
    const double rr_old = s->rr[ob][c];
    double rr_new = rr_old;
    int dnew = dold;
    if (it > 0 && !dold) {
      rr_new = sum_partials(rr_part, PNH_RED_BLOCKS, nb, c, red).x;
      dnew = sqrt(rr_new) < s->tol[c];
    }
    if (threadIdx.x == 0) {
      beta[c] = it > 0 && !dnew ? rr_new / rr_old : 0.0;
      skip[c] = dnew;
    }
    if (writer) {
      s->rr[nbuf][c] = rr_new;
      s->done[nbuf][c] = dnew;
      if (dnew && !dold) {
        s->iters[c] = it;
        s->res[c] = sqrt(rr_new);
        s->stopped[c] = 1;
      }
    }
  }
  __syncthreads();
  if (writer) {
    int all = 1;
    for (int c = 0; c < nb; ++c) all &= skip[c];
    s->all_done[nbuf] = all;
  }
  const int64_t n = n_cells * nb;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % nb);
    if (skip[c]) continue;
    const double2 rv = r[i];
    if (it == 0) { p[i] = rv; continue; }
    const double bt = beta[c];
    double2 pv = p[i];
    pv.x = rv.x + bt * pv.x;
    pv.y = rv.y + bt * pv.y;
    p[i] = pv;
  }
}

// alpha = rr / pq; x += alpha p; r -= alpha q; block partial of |r|^2.
// Every block sums the p.q block partials itself (one fixed order and tree).
// State from parity buffer `buf` (the iteration's).
__global__ void __launch_bounds__(256)
cg_xr_update_kernel(double2* __restrict__ x, double2* __restrict__ r,
                    const double2* __restrict__ p, const double2* __restrict__ q,
                    int64_t n_cells, int nb, const CgScalars* __restrict__ s, int buf,
                    const double2* __restrict__ pq_part, int pq_blocks,
                    double2* __restrict__ part) {
  if (s->all_done[buf]) return;
  __shared__ double2 red[256];
  for (int c = 0; c < nb; ++c) {
    double acc = 0.0;
    if (!s->done[buf][c]) {
      const double2 pq = sum_partials(pq_part, pq_blocks, nb, c, red);
      const double rho = s->rr[buf][c];
      const double den = pq.x * pq.x + pq.y * pq.y;
      const double ax = rho * pq.x / den, ay = -rho * pq.y / den;
      for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n_cells;
           g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = g * nb + c;
        const double2 pv = p[i], qv = q[i];
        double2 xv = x[i], rv = r[i];
        xv.x += ax * pv.x - ay * pv.y;
        xv.y += ax * pv.y + ay * pv.x;
        rv.x -= ax * qv.x - ay * qv.y;
        rv.y -= ax * qv.y + ay * qv.x;
        x[i] = xv;
        r[i] = rv;
        acc += rv.x * rv.x + rv.y * rv.y;
      }
    }
    red[threadIdx.x] = make_double2(acc, 0.0);
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) red[threadIdx.x].x += red[threadIdx.x + st].x;
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * nb + c] = red[0];
    __syncthreads();
  }
}

inline unsigned grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

int pnh_scale_sn(int fp64, const void* in, void* out, const void* sn,
                 int64_t n_img, int nb, hipStream_t s) {
  if (fp64)
    hipLaunchKernelGGL(scale_sn_kernel<double>, dim3(grid_for(n_img * nb)), dim3(256), 0, s,
                       in, out, (const double*)sn, n_img, nb);
  else
    hipLaunchKernelGGL(scale_sn_kernel<float>, dim3(grid_for(n_img * nb)), dim3(256), 0, s,
                       in, out, (const float*)sn, n_img, nb);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_c64_to_c128(const void* in, void* out, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(c64_to_c128_kernel, dim3(grid_for(n)), dim3(256), 0, s,
                     (const float2*)in, (double2*)out, n);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_c128_to_c64(const void* in, void* out, int64_t n, hipStream_t s) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(c128_to_c64_kernel, dim3(grid_for(n)), dim3(256), 0, s,
                     (const double2*)in, (float2*)out, n);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_cast_div_sn(const void* in, void* out, const float* sn, int64_t n_img,
                    int nb, hipStream_t s) {
  hipLaunchKernelGGL(cast_div_sn_kernel, dim3(grid_for(n_img * nb)), dim3(256), 0, s,
                     (const double2*)in, (float2*)out, sn, n_img, nb);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_coil_expand(const void* x, const void* sense, void* out, int64_t n_img, int nb,
                    hipStream_t s) {
  if (n_img * nb == 0) return 0;
  hipLaunchKernelGGL(coil_expand_kernel, dim3(grid_for(n_img * nb)), dim3(256), 0, s,
                     (const float2*)x, (const float2*)sense, (float2*)out, n_img, nb);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_coil_combine(const void* xc, const void* sense, void* out, int64_t n_img, int nb,
                     hipStream_t s) {
  if (n_img == 0) return 0;
  hipLaunchKernelGGL(coil_combine_kernel, dim3(grid_for(n_img)), dim3(256), 0, s,
                     (const float2*)xc, (const float2*)sense, (float2*)out, n_img, nb);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_zero(int fp64, void* p, int64_t n, hipStream_t s) {
  if (fp64)
    hipLaunchKernelGGL(zero_kernel<double2>, dim3(grid_for(n)), dim3(256), 0, s,
                       (double2*)p, n);
  else
    hipLaunchKernelGGL(zero_kernel<float2>, dim3(grid_for(n)), dim3(256), 0, s,
                       (float2*)p, n);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_cg_dot(int mode, const void* a, const void* b, int64_t n_cells, int nb,
               void* part, const int* skip, hipStream_t s) {
  if (mode == 0)
    hipLaunchKernelGGL(dot_partials_kernel<0>, dim3(PNH_RED_BLOCKS), dim3(256), 0, s,
                       (const double2*)a, (const double2*)b, n_cells, nb, (double2*)part, skip);
  else
    hipLaunchKernelGGL(dot_partials_kernel<1>, dim3(PNH_RED_BLOCKS), dim3(256), 0, s,
                       (const double2*)a, (const double2*)b, n_cells, nb, (double2*)part, skip);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_cg_init(const void* part, int nb, CgScalars* sc, int maxiter, double rtol,
                double atol, hipStream_t s) {
  hipLaunchKernelGGL(cg_init_kernel, dim3(nb), dim3(256), 0, s, (const double2*)part,
                     PNH_RED_BLOCKS, nb, sc, maxiter, rtol, atol);
  PNH_LAUNCH_CHECK();
  if (nb > 1) {
    hipLaunchKernelGGL(cg_all_done_kernel, dim3(1), dim3(64), 0, s, sc, nb);
    PNH_LAUNCH_CHECK();
  }
  return 0;
}

int pnh_cg_p_update(void* p, const void* r, int64_t n_cells, int nb, CgScalars* sc,
                    const void* rr_part, int it, hipStream_t s) {
  hipLaunchKernelGGL(cg_p_update_kernel, dim3(grid_for(n_cells * nb)), dim3(256), 0, s,
                     (double2*)p, (const double2*)r, n_cells, nb, sc,
                     (const double2*)rr_part, it);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_cg_xr_update(void* x, void* r, const void* p, const void* q, int64_t n_cells,
                     int nb, const CgScalars* sc, int it, const void* pq_part, void* part,
                     hipStream_t s) {
  hipLaunchKernelGGL(cg_xr_update_kernel, dim3(PNH_RED_BLOCKS), dim3(256), 0, s,
                     (double2*)x, (double2*)r, (const double2*)p, (const double2*)q,
                     n_cells, nb, sc, it & 1, (const double2*)pq_part, PNH_RED_BLOCKS,
                     (double2*)part);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_cg_resid(const void* rr_part, int nb, const CgScalars* sc, double* out,
                 hipStream_t s) {
  hipLaunchKernelGGL(cg_resid_kernel, dim3(nb), dim3(256), 0, s, (const double2*)rr_part,
                     PNH_RED_BLOCKS, nb, sc, out);
  PNH_LAUNCH_CHECK();
  return 0;
}

// ------------------------------------------------------------- batches ----
// CG over many independent problems at once.  Packed vectors hold the
// problems back to back; block b works on chunk (b - first[p]) of problem
// p = tab[b]; dot products reduce per block and then per problem (one block
// per problem, fixed order), so results are deterministic.
#include "batch.h"

namespace {

__device__ __forceinline__ void seg_range(const BatchProblem& pr, int op, int64_t lb,
                                          int64_t chunk, int64_t& off, int64_t& lo,
                                          int64_t& hi) {
  int64_t len;
  if (op == BVEC_CAST_Y) { off = pr.y_off; len = pr.ia.M; }
  else if (op == BVEC_CAST_DIV) { off = pr.x_off; len = pr.n_img; }
  else { off = pr.k_off; len = pr.n_grid; }
  lo = lb * chunk;
  hi = min(len, lo + chunk);
}

__device__ __forceinline__ void block_sum2(double sx, double sy, double2* out) {
  __shared__ double2 red[256];
  red[threadIdx.x] = make_double2(sx, sy);
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      red[threadIdx.x].x += red[threadIdx.x + st].x;
      red[threadIdx.x].y += red[threadIdx.x + st].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// Sum of problem p's block partials (grid table blocks first[p] ..
// first[p+1]) in one fixed order and tree; every thread gets the total, and
// every workgroup of the problem the same bits.
__device__ __forceinline__ double2 problem_sum(const double2* __restrict__ part,
                                               const BatchTable& t, int p, double2* red) {
  const int64_t b0 = t.first[p], b1 = t.first[p + 1];
  double sx = 0.0, sy = 0.0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
    const double2 v = part[i];
    sx += v.x; sy += v.y;
  }
  red[threadIdx.x] = make_double2(sx, sy);
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      red[threadIdx.x].x += red[threadIdx.x + st].x;
      red[threadIdx.x].y += red[threadIdx.x + st].y;
    }
    __syncthreads();
  }
  const double2 tot = red[0];
  __syncthreads();
  return tot;
}

// blockDim.x == 256.  CG steps read the state of iteration `it` (see
// BatchCg); the p update of iteration it also finishes the previous
// iteration's |r|^2 (or, at it == 0, |b|^2: tolerance and first state),
// applies the stopping test and writes buffer it&1 from its first block.
__global__ void __launch_bounds__(256)
batch_vec_kernel(int op, const BatchProblem* __restrict__ probs, BatchTable t,
                 int64_t chunk, void* a_, void* b_, void* c_, void* d_, BatchCg cg, int it) {
  __shared__ double2 red[256];
  const int p = t.tab[blockIdx.x];
  const int64_t lb = (int64_t)blockIdx.x - t.first[p];
  const BatchProblem& pr = probs[p];
  int64_t off, lo, hi;
  seg_range(pr, op, lb, chunk, off, lo, hi);
  const int cur = it & 1, prev = (it + 1) & 1;
  const bool writer = lb == 0 && threadIdx.x == 0;
  double sx = 0.0, sy = 0.0;
  if (op == BVEC_P_UPDATE) {
    double beta = 0.0;
    if (it == 0) {
      const double bb = problem_sum(cg.part_rr, t, p, red).x;
      const double bn = sqrt(bb);
      const double tol = fmax(cg.atol, cg.rtol * bn);
      const int dn = (bn == 0.0) || (bn < tol);
      if (writer) {
        cg.tol[p] = tol;
        cg.rr[p] = bb;
        cg.done[p] = dn;
        cg.iters[p] = dn ? 0 : cg.maxiter;
        cg.stopped[p] = dn;
        cg.res[p] = bn;
      }
      if (dn) return;
    } else {
      const double rr_old = cg.rr[prev * cg.n + p];
      if (cg.done[prev * cg.n + p]) {
        if (writer) {
          cg.rr[cur * cg.n + p] = rr_old;
          cg.done[cur * cg.n + p] = 1;
        }
        return;
      }
      const double rr_new = problem_sum(cg.part_rr, t, p, red).x;
      const int dn = sqrt(rr_new) < cg.tol[p];
      if (writer) {
        cg.rr[cur * cg.n + p] = rr_new;
        cg.done[cur * cg.n + p] = dn;
        if (dn) {
          cg.iters[p] = it;
          cg.stopped[p] = 1;
          cg.res[p] = sqrt(rr_new);
        }
      }
      if (dn) return;
      beta = rr_new / rr_old;
    }
    double2* pv = (double2*)a_;
    const double2* rv = (const double2*)b_;
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
      const int64_t e = off + i;
      const double2 r = rv[e];
      if (it == 0) { pv[e] = r; continue; }
      double2 q = pv[e];
      q.x = r.x + beta * q.x;
      q.y = r.y + beta * q.y;
      pv[e] = q;
    }
    return;
  }
  if ((op == BVEC_DOT_PQ || op == BVEC_XR_UPDATE) && cg.done[cur * cg.n + p]) return;
  double ax = 0.0, ay = 0.0;
  if (op == BVEC_XR_UPDATE) {
    const double2 pq = problem_sum(cg.part_pq, t, p, red);
    const double rho = cg.rr[cur * cg.n + p];
    const double den = pq.x * pq.x + pq.y * pq.y;
    ax = rho * pq.x / den;
    ay = -rho * pq.y / den;
  }
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const int64_t e = off + i;
    switch (op) {
      case BVEC_CAST_Y: {
        const float2 v = ((const float2*)a_)[e];
        ((double2*)b_)[e] = make_double2(v.x, v.y);
        break;
      }
      case BVEC_ZERO_X:
        ((double2*)a_)[e] = make_double2(0.0, 0.0);
        break;
      case BVEC_DOT_RR: {
        const double2 v = ((const double2*)a_)[e];
        sx += v.x * v.x + v.y * v.y;
        break;
      }
      case BVEC_DOT_PQ: {
        const double2 u = ((const double2*)a_)[e];
        const double2 v = ((const double2*)b_)[e];
        sx += u.x * v.x + u.y * v.y;
        sy += u.x * v.y - u.y * v.x;
        break;
      }
      case BVEC_XR_UPDATE: {
        double2* x = (double2*)a_;
        double2* r = (double2*)b_;
        const double2 pv = ((const double2*)c_)[e], qv = ((const double2*)d_)[e];
        double2 xv = x[e], rv = r[e];
        xv.x += ax * pv.x - ay * pv.y;
        xv.y += ax * pv.y + ay * pv.x;
        rv.x -= ax * qv.x - ay * qv.y;
        rv.y -= ax * qv.y + ay * qv.x;
        x[e] = xv;
        r[e] = rv;
        sx += rv.x * rv.x + rv.y * rv.y;
        break;
      }
      case BVEC_CAST_DIV: {
        const double2 v = ((const double2*)a_)[e];
        const float inv = 1.0f / pr.sn32[i];
        ((float2*)b_)[e] = make_float2((float)v.x * inv, (float)v.y * inv);
        break;
      }
    }
  }
  if (op == BVEC_DOT_RR || op == BVEC_XR_UPDATE)
    block_sum2(sx, sy, &cg.part_rr[blockIdx.x]);
  else if (op == BVEC_DOT_PQ)
    block_sum2(sx, sy, &cg.part_pq[blockIdx.x]);
}

// One block per problem: iterations and final ||r|| (the stopping test's
// value, else the norm from the last |r|^2 partials; ||b|| at maxiter 0).
__global__ void __launch_bounds__(256)
batch_cg_out_kernel(BatchTable t, BatchCg cg, int* __restrict__ iters_out,
                    double* __restrict__ resid_out) {
  __shared__ double2 red[256];
  const int p = blockIdx.x;
  const bool stopped = cg.maxiter > 0 && cg.stopped[p];
  double res = 0.0;
  if (stopped) res = cg.res[p];
  else res = sqrt(problem_sum(cg.part_rr, t, p, red).x);
  if (threadIdx.x != 0) return;
  if (iters_out) iters_out[p] = cg.maxiter > 0 ? cg.iters[p] : 0;
  if (resid_out) resid_out[p] = res;
}

}  // namespace

int pnh_batch_vec_launch(int op, const BatchProblem* probs, const BatchTable& t,
                         int64_t chunk, void* a, void* b, void* c, void* d,
                         const BatchCg& cg, int it, hipStream_t s) {
  if (t.n_blocks == 0) return 0;
  hipLaunchKernelGGL(batch_vec_kernel, dim3((unsigned)t.n_blocks), dim3(256), 0, s, op, probs,
                     t, chunk, a, b, c, d, cg, it);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_batch_cg_out_launch(const BatchTable& t, const BatchCg& cg, int32_t* iters_out,
                            double* resid_out, hipStream_t s) {
  if (!iters_out && !resid_out) return 0;
  hipLaunchKernelGGL(batch_cg_out_kernel, dim3((unsigned)cg.n), dim3(256), 0, s, t, cg,
                     iters_out, resid_out);
  PNH_LAUNCH_CHECK();
  return 0;
}
