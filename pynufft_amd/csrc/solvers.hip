// Element-wise kernels of the Toeplitz self-adjoint, the pipe density
// compensation ('dc') and the L1-TV split-Bregman solver ('L1TVOLS'), in the
// precision and operation order numpy runs the reference
// (linalg/solve_cpu.py, nufft/_nufft_class_methods_cpu.py): complex64, except
// the density weights of 'dc', which the reference carries in complex128.
#include "common.h"
#include "solvers.h"

namespace {

inline unsigned blocks_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65535) b = 65535;
  return (unsigned)(b < 1 ? 1 : b);
}

#define GRID_STRIDE(i, n)                                                   \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
       i += (int64_t)gridDim.x * blockDim.x)

__global__ void fill_kernel(float2* __restrict__ p, int64_t n, float2 v) {
  GRID_STRIDE(i, n) p[i] = v;
}

// a / b with Smith's scaling (numpy's complex64 division)
__device__ __forceinline__ float2 cdivf(float2 a, float2 b) {
  if (fabsf(b.x) >= fabsf(b.y)) {
    if (b.x == 0.f && b.y == 0.f) return make_float2(a.x / fabsf(b.x), a.y / fabsf(b.y));
    const float r = b.y / b.x, d = b.x + b.y * r;
    return make_float2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  }
  const float r = b.x / b.y, d = b.x * r + b.y;
  return make_float2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}

// a / b in FP64, same scaling (numpy's complex128 division)
__device__ __forceinline__ double2 cdivd(double2 a, double2 b) {
  if (fabs(b.x) >= fabs(b.y)) {
    if (b.x == 0.0 && b.y == 0.0) return make_double2(a.x / fabs(b.x), a.y / fabs(b.y));
    const double r = b.y / b.x, d = b.x + b.y * r;
    return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  }
  const double r = b.x / b.y, d = b.x * r + b.y;
  return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}

__global__ void cdiv_kernel(float2* __restrict__ w, const float2* __restrict__ e, int64_t n) {
  GRID_STRIDE(i, n) w[i] = cdivf(w[i], e[i]);
}

__global__ void cdiv64_kernel(double2* __restrict__ w, const double2* __restrict__ e,
                              int64_t n) {
  GRID_STRIDE(i, n) w[i] = cdivd(w[i], e[i]);
}

__global__ void fill64_kernel(double2* __restrict__ p, int64_t n, double2 v) {
  GRID_STRIDE(i, n) p[i] = v;
}

// out = a (complex128) * b (complex64 widened)
__global__ void cmul64_kernel(const double2* __restrict__ a, const float2* __restrict__ b,
                              double2* __restrict__ out, int64_t n) {
  GRID_STRIDE(i, n) {
    const double2 u = a[i];
    const double vx = b[i].x, vy = b[i].y;
    out[i] = make_double2(u.x * vx - u.y * vy, u.x * vy + u.y * vx);
  }
}

// out[i] = complex64(in[i]) * sn[i / nb]: a complex128 image rounded to
// complex64, then apodized in complex64 (the reference's k2xx then xx2x)
__global__ void cast_mul_sn_kernel(const double2* __restrict__ in, float2* __restrict__ out,
                                   const float* __restrict__ sn, int64_t n_img, int nb) {
  GRID_STRIDE(i, n_img * nb) {
    const float s = sn[i / nb];
    const double2 v = in[i];
    out[i] = make_float2((float)v.x * s, (float)v.y * s);
  }
}

__global__ void cmul_kernel(const float2* __restrict__ a, const float2* __restrict__ b,
                            float2* __restrict__ out, int64_t n) {
  GRID_STRIDE(i, n) {
    const float2 u = a[i], v = b[i];
    out[i] = make_float2(u.x * v.x - u.y * v.y, u.x * v.y + u.y * v.x);
  }
}

__global__ void abs_kernel(const float2* __restrict__ k, float* __restrict__ w, int64_t n,
                           int nb) {
  GRID_STRIDE(i, n) {
    const float2 v = k[i * nb];
    w[i] = hypotf(v.x, v.y);
  }
}

// uker = mu w - lambda Lap in complex64, rounded step by step as numpy does
// it for the reference's arrays (float32 w, complex64 Lap): lambda Lap, then
// the difference.  Lap: the Laplacian spectrum passed in (the reference's
// complex64 FFT of the periodic stencil, whose rounding the solve is
// sensitive to), or, with lap null, its closed form
// sum_d 2 cos(2 pi k_d / K_d) - 2 evaluated in FP64.
__global__ void uker_kernel(const float* __restrict__ w, const float2* __restrict__ lap,
                            float2* __restrict__ uker, ShapeArgs s, float mu, float lam) {
  GRID_STRIDE(i, s.n) {
    float2 l;
    if (lap) {
      l = lap[i];
    } else {
      int64_t rest = i;
      double acc = 0.0;
      for (int d = s.D - 1; d >= 0; --d) {
        const int64_t c = rest % s.dim[d];
        rest /= s.dim[d];
        acc += 2.0 * cospi(2.0 * (double)c / (double)s.dim[d]) - 2.0;
      }
      l = make_float2((float)acc, 0.0f);
    }
    const float tr = __fmul_rn(lam, l.x), ti = __fmul_rn(lam, l.y);
    uker[i] = make_float2(__fsub_rn(__fmul_rn(mu, w[i]), tr), -ti);
  }
}

// k = (k + 1e-7) / (uker + 1e-7): the k-space solve of (mu W - lambda Lap);
// the complex quotient is formed in FP64 and rounded once.
__global__ void kdiv_kernel(float2* __restrict__ k, const float2* __restrict__ uker,
                            int64_t n) {
  GRID_STRIDE(i, n) {
    const float2 u = uker[i];
    const float2 v = k[i];
    const double ar = (double)__fadd_rn(v.x, 1e-7f), ai = v.y;
    const double br = (double)__fadd_rn(u.x, 1e-7f), bi = u.y;
    const double den = br * br + bi * bi;
    k[i] = make_float2((float)((ar * br + ai * bi) / den), (float)((ai * br - ar * bi) / den));
  }
}

// linear index of the periodic neighbour at offset `step` (+1 / -1) along axis p
__device__ __forceinline__ int64_t neighbour(const ShapeArgs& s, int64_t i, int p, int step) {
  int64_t stride = 1;
  for (int d = s.D - 1; d > p; --d) stride *= s.dim[d];
  const int64_t c = (i / stride) % s.dim[p];
  int64_t cn = c + step;
  if (cn < 0) cn += s.dim[p];
  if (cn >= s.dim[p]) cn -= s.dim[p];
  return i + (cn - c) * stride;
}

__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cadd2(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }

// rhs = mu AHyk + sum_p lambda ((dd_p - bb_p)[i + e_p] - (dd_p - bb_p)[i])
__global__ void l1tv_rhs_kernel(const float2* __restrict__ AHyk, const float2* __restrict__ dd,
                                const float2* __restrict__ bb, float2* __restrict__ rhs,
                                ShapeArgs s, float mu, float lam) {
  GRID_STRIDE(i, s.n) {
    const float2 a = AHyk[i];
    float2 r = make_float2(mu * a.x, mu * a.y);
    for (int p = 0; p < s.D; ++p) {
      const int64_t j = neighbour(s, i, p, +1);
      const float2* dp = dd + p * s.n;
      const float2* bp = bb + p * s.n;
      const float2 g = csub(csub(dp[j], bp[j]), csub(dp[i], bp[i]));
      r = cadd2(r, make_float2(lam * g.x, lam * g.y));
    }
    rhs[i] = r;
  }
}

// zz_p = x[i - e_p] - x[i]; t_p = zz_p + bb_p; s = |(t_0, .., t_{D-1})|
// (nested hypot) + 1e-5; r = max(s - 1/lambda, 0) / s; dd_p = t_p r;
// bb_p += zz_p - dd_p; AHyk -= AHA x - AHy
__global__ void l1tv_shrink_kernel(const float2* __restrict__ x, const float2* __restrict__ q,
                                   const float2* __restrict__ AHy, float2* __restrict__ AHyk,
                                   float2* __restrict__ dd, float2* __restrict__ bb,
                                   ShapeArgs s, float thr) {
  GRID_STRIDE(i, s.n) {
    const float2 xi = x[i];
    float2 zz[3], t[3];
    float mag = 0.f;
    for (int p = 0; p < s.D; ++p) {
      zz[p] = csub(x[neighbour(s, i, p, -1)], xi);
      t[p] = cadd2(zz[p], bb[p * s.n + i]);
      mag = hypotf(hypotf(t[p].x, t[p].y), mag);
    }
    mag += 1e-5f;
    const float r = mag > thr ? (mag - thr) / mag : 0.f;
    for (int p = 0; p < s.D; ++p) {
      const float2 d = make_float2(t[p].x * r, t[p].y * r);
      dd[p * s.n + i] = d;
      bb[p * s.n + i] = cadd2(bb[p * s.n + i], csub(zz[p], d));
    }
    AHyk[i] = csub(AHyk[i], csub(q[i], AHy[i]));
  }
}

}  // namespace

#define LAUNCH(kernel, n, ...)                                                      \
  do {                                                                              \
    if ((n) > 0) hipLaunchKernelGGL(kernel, dim3(blocks_for(n)), dim3(256), 0, s,   \
                                    __VA_ARGS__);                                   \
    PNH_LAUNCH_CHECK();                                                             \
  } while (0)

int pnh_fill_c64(void* p, int64_t n, float re, float im, hipStream_t s) {
  LAUNCH(fill_kernel, n, (float2*)p, n, make_float2(re, im));
  return 0;
}

int pnh_cdiv_c64(void* w, const void* e, int64_t n, hipStream_t s) {
  LAUNCH(cdiv_kernel, n, (float2*)w, (const float2*)e, n);
  return 0;
}

int pnh_fill_c128(void* p, int64_t n, double re, double im, hipStream_t s) {
  LAUNCH(fill64_kernel, n, (double2*)p, n, make_double2(re, im));
  return 0;
}

int pnh_cdiv_c128(void* w, const void* e, int64_t n, hipStream_t s) {
  LAUNCH(cdiv64_kernel, n, (double2*)w, (const double2*)e, n);
  return 0;
}

int pnh_cmul_c128_c64(const void* a, const void* b, void* out, int64_t n, hipStream_t s) {
  LAUNCH(cmul64_kernel, n, (const double2*)a, (const float2*)b, (double2*)out, n);
  return 0;
}

int pnh_cast_mul_sn(const void* in, void* out, const float* sn, int64_t n_img, int nb,
                    hipStream_t s) {
  LAUNCH(cast_mul_sn_kernel, n_img * nb, (const double2*)in, (float2*)out, sn, n_img, nb);
  return 0;
}

int pnh_cmul_c64(const void* a, const void* b, void* out, int64_t n, hipStream_t s) {
  LAUNCH(cmul_kernel, n, (const float2*)a, (const float2*)b, (float2*)out, n);
  return 0;
}

int pnh_abs_first_coil(const void* k, float* w, int64_t n, int nb, hipStream_t s) {
  LAUNCH(abs_kernel, n, (const float2*)k, w, n, nb);
  return 0;
}

int pnh_l1tv_uker(const float* w, const void* lap, void* uker, const ShapeArgs& sh, float mu,
                  float lam, hipStream_t s) {
  LAUNCH(uker_kernel, sh.n, w, (const float2*)lap, (float2*)uker, sh, mu, lam);
  return 0;
}

int pnh_l1tv_kdiv(void* k, const void* uker, int64_t n, hipStream_t s) {
  LAUNCH(kdiv_kernel, n, (float2*)k, (const float2*)uker, n);
  return 0;
}

int pnh_l1tv_rhs(const void* AHyk, const void* dd, const void* bb, void* rhs,
                 const ShapeArgs& sh, float mu, float lam, hipStream_t s) {
  LAUNCH(l1tv_rhs_kernel, sh.n, (const float2*)AHyk, (const float2*)dd, (const float2*)bb,
         (float2*)rhs, sh, mu, lam);
  return 0;
}

int pnh_l1tv_shrink(const void* x, const void* q, const void* AHy, void* AHyk, void* dd,
                    void* bb, const ShapeArgs& sh, float thr, hipStream_t s) {
  LAUNCH(l1tv_shrink_kernel, sh.n, (const float2*)x, (const float2*)q, (const float2*)AHy,
         (float2*)AHyk, (float2*)dd, (float2*)bb, sh, thr);
  return 0;
}

// ------------------------------------------------------------------ LSQR --
// scipy.sparse.linalg.lsqr (Paige & Saunders) on A = P (k2y), A^H = y2k, in
// FP64 on the device.  Every scalar of the recurrence lives in LsqrState and
// is updated by one thread; the vector kernels read what they need, so the
// host only launches.  Reductions are deterministic (fixed grid, fixed tree).
namespace {

__device__ double sum_partials(const double* __restrict__ part) {
  double s = 0.0;
  for (int b = 0; b < PNH_LSQR_BLOCKS; ++b) s += part[b];
  return s;
}

// per-block partial of sum |scale * a|^2 (scale == 0 means 1)
__device__ void block_sum_store(double v, double* __restrict__ part) {
  __shared__ double red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256)
norm2_partials_kernel(const double2* __restrict__ a, int64_t n, double* __restrict__ part,
                      const int* __restrict__ skip) {
  if (skip && *skip) return;
  double acc = 0.0;
  GRID_STRIDE(i, n) acc += a[i].x * a[i].x + a[i].y * a[i].y;
  block_sum_store(acc, part);
}

// a = b - c * a (c = *coef), partial |a|^2
__global__ void __launch_bounds__(256)
lsqr_axpby_kernel(double2* __restrict__ a, const double2* __restrict__ b,
                  const double* __restrict__ coef, int64_t n, double* __restrict__ part,
                  const int* __restrict__ skip) {
  if (*skip) return;
  const double c = *coef;
  double acc = 0.0;
  GRID_STRIDE(i, n) {
    const double2 v = make_double2(b[i].x - c * a[i].x, b[i].y - c * a[i].y);
    a[i] = v;
    acc += v.x * v.x + v.y * v.y;
  }
  block_sum_store(acc, part);
}

// a *= 1 / *norm when *norm > 0; optionally copy the result to w
__global__ void lsqr_scale_kernel(double2* __restrict__ a, const double* __restrict__ norm,
                                  int64_t n, double2* __restrict__ w,
                                  const int* __restrict__ skip) {
  if (*skip) return;
  const double nv = *norm;
  const double inv = nv > 0.0 ? 1.0 / nv : 1.0;
  GRID_STRIDE(i, n) {
    double2 v = a[i];
    if (nv > 0.0) { v.x *= inv; v.y *= inv; a[i] = v; }
    if (w) w[i] = v;
  }
}

// partial |w / rho|^2 (old w), x += t1 w, w = v + t2 w
__global__ void __launch_bounds__(256)
lsqr_xw_kernel(double2* __restrict__ x, double2* __restrict__ w, const double2* __restrict__ v,
               const LsqrState* __restrict__ st, int64_t n, double* __restrict__ part) {
  if (st->done) return;
  const double t1 = st->t1, t2 = st->t2, ir = 1.0 / st->rho;
  double acc = 0.0;
  GRID_STRIDE(i, n) {
    const double2 wo = w[i];
    const double dx = ir * wo.x, dy = ir * wo.y;
    acc += dx * dx + dy * dy;
    double2 xv = x[i];
    xv.x += t1 * wo.x;
    xv.y += t1 * wo.y;
    x[i] = xv;
    w[i] = make_double2(v[i].x + t2 * wo.x, v[i].y + t2 * wo.y);
  }
  block_sum_store(acc, part);
}

__device__ double sgn(double a) { return a > 0.0 ? 1.0 : (a < 0.0 ? -1.0 : 0.0); }

// stable Givens rotation (scipy _sym_ortho)
__device__ void sym_ortho(double a, double b, double& c, double& s, double& r) {
  if (b == 0.0) { c = sgn(a); s = 0.0; r = fabs(a); }
  else if (a == 0.0) { c = 0.0; s = sgn(b); r = fabs(b); }
  else if (fabs(b) > fabs(a)) {
    const double tau = a / b;
    s = sgn(b) / sqrt(1.0 + tau * tau);
    c = s * tau;
    r = b / s;
  } else {
    const double tau = b / a;
    c = sgn(a) / sqrt(1.0 + tau * tau);
    s = c * tau;
    r = a / c;
  }
}

__global__ void lsqr_init_kernel(LsqrState* __restrict__ st, double atol, double btol,
                                 double conlim, double damp, int iter_lim) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  LsqrState z = {};
  z.atol = atol;
  z.btol = btol;
  z.ctol = conlim > 0.0 ? 1.0 / conlim : 0.0;
  z.damp = damp;
  z.iter_lim = iter_lim;
  z.cs2 = -1.0;
  *st = z;
}

// phase: 0 = bnorm/beta from |b|^2; 1 = alfa of the start vector (v = A^H u);
// 2 = beta of an iteration; 3 = alfa of an iteration, then the rotation;
// 4 = after the x/w update: norm estimates and the stopping tests.
__global__ void lsqr_scalar_kernel(LsqrState* __restrict__ st, const double* __restrict__ part,
                                   int phase) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (st->done) return;
  const double S = sum_partials(part);
  if (phase == 0) {
    st->bnorm = st->beta = sqrt(S);
    if (!(st->beta > 0.0)) { st->alfa = 0.0; st->done = 1; st->skip_v = 1; }
    else st->skip_v = 0;
    return;
  }
  if (phase == 1) {
    st->alfa = sqrt(S);
    st->rhobar = st->alfa;
    st->phibar = st->beta;
    if (st->alfa * st->beta == 0.0) st->done = 1;
    return;
  }
  if (phase == 2) {
    st->itn += 1;
    st->beta = sqrt(S);
    if (st->beta > 0.0) {
      st->anorm = sqrt(st->anorm * st->anorm + st->alfa * st->alfa + st->beta * st->beta +
                       st->damp * st->damp);
      st->skip_v = 0;
    } else {
      st->skip_v = 1;
    }
    return;
  }
  if (phase == 3) {
    if (!st->skip_v) st->alfa = sqrt(S);
    double rhobar1, psi;
    if (st->damp > 0.0) {
      rhobar1 = sqrt(st->rhobar * st->rhobar + st->damp * st->damp);
      const double cs1 = st->rhobar / rhobar1, sn1 = st->damp / rhobar1;
      psi = sn1 * st->phibar;
      st->phibar = cs1 * st->phibar;
    } else {
      rhobar1 = st->rhobar;
      psi = 0.0;
    }
    double cs, sn, rho;
    sym_ortho(rhobar1, st->beta, cs, sn, rho);
    const double theta = sn * st->alfa;
    st->rhobar = -cs * st->alfa;
    const double phi = cs * st->phibar;
    st->phibar = sn * st->phibar;
    st->tau = sn * phi;
    st->t1 = phi / rho;
    st->t2 = -theta / rho;
    st->rho = rho;
    st->theta = theta;
    st->phi = phi;
    st->psi = psi;
    return;
  }
  // phase 4
  const double eps = 2.220446049250313e-16;
  st->ddnorm += S;
  const double rho = st->rho, theta = st->theta, phi = st->phi;
  const double delta = st->sn2 * rho;
  const double gambar = -st->cs2 * rho;
  const double rhs = phi - delta * st->z;
  const double zbar = rhs / gambar;
  st->xnorm = sqrt(st->xxnorm + zbar * zbar);
  const double gamma = sqrt(gambar * gambar + theta * theta);
  st->cs2 = gambar / gamma;
  st->sn2 = theta / gamma;
  st->z = rhs / gamma;
  st->xxnorm += st->z * st->z;
  st->acond = st->anorm * sqrt(st->ddnorm);
  const double res1 = st->phibar * st->phibar;
  st->res2 += st->psi * st->psi;
  const double rnorm = sqrt(res1 + st->res2);
  const double arnorm = st->alfa * fabs(st->tau);
  const double test1 = rnorm / st->bnorm;
  const double test2 = arnorm / (st->anorm * rnorm + eps);
  const double test3 = 1.0 / (st->acond + eps);
  const double t1 = test1 / (1.0 + st->anorm * st->xnorm / st->bnorm);
  const double rtol = st->btol + st->atol * st->anorm * st->xnorm / st->bnorm;
  int istop = 0;
  if (st->itn >= st->iter_lim) istop = 7;
  if (1.0 + test3 <= 1.0) istop = 6;
  if (1.0 + test2 <= 1.0) istop = 5;
  if (1.0 + t1 <= 1.0) istop = 4;
  if (test3 <= st->ctol) istop = 3;
  if (test2 <= st->atol) istop = 2;
  if (test1 <= rtol) istop = 1;
  st->istop = istop;
  if (istop) { st->done = 1; st->skip_v = 1; }
}

}  // namespace

int pnh_lsqr_norm2(const void* a, int64_t n, double* part, const int* skip, hipStream_t s) {
  hipLaunchKernelGGL(norm2_partials_kernel, dim3(PNH_LSQR_BLOCKS), dim3(256), 0, s,
                     (const double2*)a, n, part, skip);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_lsqr_axpby(void* a, const void* b, const double* coef, int64_t n, double* part,
                   const int* skip, hipStream_t s) {
  hipLaunchKernelGGL(lsqr_axpby_kernel, dim3(PNH_LSQR_BLOCKS), dim3(256), 0, s, (double2*)a,
                     (const double2*)b, coef, n, part, skip);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_lsqr_scale(void* a, const double* norm, int64_t n, void* copy, const int* skip,
                   hipStream_t s) {
  LAUNCH(lsqr_scale_kernel, n, (double2*)a, norm, n, (double2*)copy, skip);
  return 0;
}

int pnh_lsqr_xw(void* x, void* w, const void* v, const LsqrState* st, int64_t n, double* part,
                hipStream_t s) {
  hipLaunchKernelGGL(lsqr_xw_kernel, dim3(PNH_LSQR_BLOCKS), dim3(256), 0, s, (double2*)x,
                     (double2*)w, (const double2*)v, st, n, part);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_lsqr_init(LsqrState* st, double atol, double btol, double conlim, double damp,
                  int iter_lim, hipStream_t s) {
  hipLaunchKernelGGL(lsqr_init_kernel, dim3(1), dim3(64), 0, s, st, atol, btol, conlim, damp,
                     iter_lim);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_lsqr_scalar(LsqrState* st, const double* part, int phase, hipStream_t s) {
  hipLaunchKernelGGL(lsqr_scalar_kernel, dim3(1), dim3(64), 0, s, st, part, phase);
  PNH_LAUNCH_CHECK();
  return 0;
}

// ------------------------------------------------------------------ LSMR --
// scipy.sparse.linalg.lsmr (Fong & Saunders) on the same bidiagonalization;
// scalars in LsmrState, one thread.
namespace {

__global__ void lsmr_init_kernel(LsmrState* __restrict__ st, double atol, double btol,
                                 double conlim, double damp, int maxiter) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  LsmrState z = {};
  z.atol = atol;
  z.btol = btol;
  z.ctol = conlim > 0.0 ? 1.0 / conlim : 0.0;
  z.damp = damp;
  z.maxiter = maxiter;
  *st = z;
}

// hbar = h + c1 hbar; x += c2 hbar; h = v + c3 h; partials of |x|^2
__global__ void __launch_bounds__(256)
lsmr_hx_kernel(double2* __restrict__ x, double2* __restrict__ h, double2* __restrict__ hbar,
               const double2* __restrict__ v, const LsmrState* __restrict__ st, int64_t n,
               double* __restrict__ part) {
  if (st->done) return;
  const double c1 = st->c1, c2 = st->c2, c3 = st->c3;
  double acc = 0.0;
  GRID_STRIDE(i, n) {
    const double2 hv = h[i];
    double2 hb = hbar[i];
    hb = make_double2(hv.x + c1 * hb.x, hv.y + c1 * hb.y);
    hbar[i] = hb;
    double2 xv = x[i];
    xv.x += c2 * hb.x;
    xv.y += c2 * hb.y;
    x[i] = xv;
    acc += xv.x * xv.x + xv.y * xv.y;
    h[i] = make_double2(v[i].x + c3 * hv.x, v[i].y + c3 * hv.y);
  }
  block_sum_store(acc, part);
}

// phase 0: normb, beta; 1: alpha of the start vector and the initial
// scalars; 2: beta of an iteration; 3: alpha of an iteration and every
// rotation and estimate except ||x||; 4: ||x|| and the stopping tests.
__global__ void lsmr_scalar_kernel(LsmrState* __restrict__ st, const double* __restrict__ part,
                                   int phase) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (st->done) return;
  const double S = sum_partials(part);
  if (phase == 0) {
    st->normb = st->beta = sqrt(S);
    if (!(st->beta > 0.0)) { st->alpha = 0.0; st->done = 1; st->skip_v = 1; }
    return;
  }
  if (phase == 1) {
    st->alpha = sqrt(S);
    st->zetabar = st->alpha * st->beta;
    st->alphabar = st->alpha;
    st->rho = 1.0;
    st->rhobar = 1.0;
    st->cbar = 1.0;
    st->sbar = 0.0;
    st->betadd = st->beta;
    st->betad = 0.0;
    st->rhodold = 1.0;
    st->tautildeold = 0.0;
    st->thetatilde = 0.0;
    st->zeta = 0.0;
    st->d = 0.0;
    st->normA2 = st->alpha * st->alpha;
    st->maxrbar = 0.0;
    st->minrbar = 1e100;
    st->normA = sqrt(st->normA2);
    st->condA = 1.0;
    st->normx = 0.0;
    st->normr = st->beta;
    if (st->alpha * st->beta == 0.0) st->done = 1;
    return;
  }
  if (phase == 2) {
    st->itn += 1;
    st->beta = sqrt(S);
    st->skip_v = st->beta > 0.0 ? 0 : 1;
    return;
  }
  if (phase == 3) {
    if (!st->skip_v) st->alpha = sqrt(S);
    const double alpha = st->alpha, beta = st->beta;
    double chat, shat, alphahat;
    sym_ortho(st->alphabar, st->damp, chat, shat, alphahat);
    const double rhoold = st->rho;
    double c, s, rho;
    sym_ortho(alphahat, beta, c, s, rho);
    const double thetanew = s * alpha;
    st->alphabar = c * alpha;
    const double rhobarold = st->rhobar;
    const double zetaold = st->zeta;
    const double thetabar = st->sbar * rho;
    const double rhotemp = st->cbar * rho;
    double cbar, sbar, rhobar;
    sym_ortho(st->cbar * rho, thetanew, cbar, sbar, rhobar);
    st->cbar = cbar;
    st->sbar = sbar;
    st->rhobar = rhobar;
    st->zeta = cbar * st->zetabar;
    st->zetabar = -sbar * st->zetabar;
    st->c1 = -(thetabar * rho / (rhoold * rhobarold));
    st->c2 = st->zeta / (rho * rhobar);
    st->c3 = -(thetanew / rho);
    st->rho = rho;
    // estimate of ||r||
    const double betaacute = chat * st->betadd;
    const double betacheck = -shat * st->betadd;
    const double betahat = c * betaacute;
    st->betadd = -s * betaacute;
    const double thetatildeold = st->thetatilde;
    double ctildeold, stildeold, rhotildeold;
    sym_ortho(st->rhodold, thetabar, ctildeold, stildeold, rhotildeold);
    st->thetatilde = stildeold * rhobar;
    st->rhodold = ctildeold * rhobar;
    st->betad = -stildeold * st->betad + ctildeold * betahat;
    st->tautildeold = (zetaold - thetatildeold * st->tautildeold) / rhotildeold;
    const double taud = (st->zeta - st->thetatilde * st->tautildeold) / st->rhodold;
    st->d = st->d + betacheck * betacheck;
    const double bt = st->betad - taud;
    st->normr = sqrt(st->d + bt * bt + st->betadd * st->betadd);
    // estimates of ||A|| and cond(A)
    st->normA2 = st->normA2 + beta * beta;
    st->normA = sqrt(st->normA2);
    st->normA2 = st->normA2 + alpha * alpha;
    st->maxrbar = fmax(st->maxrbar, rhobarold);
    if (st->itn > 1) st->minrbar = fmin(st->minrbar, rhobarold);
    st->condA = fmax(st->maxrbar, rhotemp) / fmin(st->minrbar, rhotemp);
    return;
  }
  // phase 4
  st->normx = sqrt(S);
  const double normar = fabs(st->zetabar);
  const double test1 = st->normr / st->normb;
  const double test2 = (st->normA * st->normr) != 0.0 ? normar / (st->normA * st->normr)
                                                      : INFINITY;
  const double test3 = 1.0 / st->condA;
  const double t1 = test1 / (1.0 + st->normA * st->normx / st->normb);
  const double rtol = st->btol + st->atol * st->normA * st->normx / st->normb;
  int istop = 0;
  if (st->itn >= st->maxiter) istop = 7;
  if (1.0 + test3 <= 1.0) istop = 6;
  if (1.0 + test2 <= 1.0) istop = 5;
  if (1.0 + t1 <= 1.0) istop = 4;
  if (test3 <= st->ctol) istop = 3;
  if (test2 <= st->atol) istop = 2;
  if (test1 <= rtol) istop = 1;
  st->istop = istop;
  if (istop) { st->done = 1; st->skip_v = 1; }
}

}  // namespace

int pnh_lsmr_init(LsmrState* st, double atol, double btol, double conlim, double damp,
                  int maxiter, hipStream_t s) {
  hipLaunchKernelGGL(lsmr_init_kernel, dim3(1), dim3(64), 0, s, st, atol, btol, conlim, damp,
                     maxiter);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_lsmr_hx(void* x, void* h, void* hbar, const void* v, const LsmrState* st, int64_t n,
                double* part, hipStream_t s) {
  hipLaunchKernelGGL(lsmr_hx_kernel, dim3(PNH_LSQR_BLOCKS), dim3(256), 0, s, (double2*)x,
                     (double2*)h, (double2*)hbar, (const double2*)v, st, n, part);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_lsmr_scalar(LsmrState* st, const double* part, int phase, hipStream_t s) {
  hipLaunchKernelGGL(lsmr_scalar_kernel, dim3(1), dim3(64), 0, s, st, part, phase);
  PNH_LAUNCH_CHECK();
  return 0;
}

// --------------------------------------------------------- BiCG(STAB) --
// scipy.sparse.linalg.bicgstab / bicg on the k-space normal operator P^H P
// (no preconditioner), FP64, x0 = 0; complex scalars in KrylovState.
namespace {

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double zabs(double2 a) { return hypot(a.x, a.y); }

__device__ void block_sum2_store(double2 v, double2* __restrict__ part) {
  __shared__ double2 red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      red[threadIdx.x].x += red[threadIdx.x + st].x;
      red[threadIdx.x].y += red[threadIdx.x + st].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// part[0][b] = block partial of vdot(a0, b0); part[1][b] of vdot(a1, b1)
// (second pair optional)
__global__ void __launch_bounds__(256)
vdot2_kernel(const double2* __restrict__ a0, const double2* __restrict__ b0,
             const double2* __restrict__ a1, const double2* __restrict__ b1, int64_t n,
             double2* __restrict__ part, const int* __restrict__ skip) {
  if (*skip) return;
  double2 s0 = make_double2(0.0, 0.0), s1 = make_double2(0.0, 0.0);
  GRID_STRIDE(i, n) {
    const double2 u = a0[i], v = b0[i];
    s0.x += u.x * v.x + u.y * v.y;
    s0.y += u.x * v.y - u.y * v.x;
    if (a1) {
      const double2 p = a1[i], q = b1[i];
      s1.x += p.x * q.x + p.y * q.y;
      s1.y += p.x * q.y - p.y * q.x;
    }
  }
  block_sum2_store(s0, part);
  if (a1) {
    __syncthreads();
    block_sum2_store(s1, part + PNH_LSQR_BLOCKS);
  }
}

__device__ double2 sum_partials2(const double2* __restrict__ part) {
  double2 s = make_double2(0.0, 0.0);
  for (int b = 0; b < PNH_LSQR_BLOCKS; ++b) { s.x += part[b].x; s.y += part[b].y; }
  return s;
}

__global__ void krylov_init_kernel(KrylovState* __restrict__ st, int stab, double rtol,
                                   double atol, int maxiter) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  KrylovState z = {};
  z.stab = stab;
  z.rtol = rtol;
  z.atol = atol;
  z.maxiter = maxiter;
  z.rhotol = 2.220446049250313e-16 * 2.220446049250313e-16;
  *st = z;
}

// Scalar steps.  BiCGSTAB: 0 = |b| (tolerance); 1 = |r| test, rho, breakdown
// tests, beta; 2 = alpha = rho / vdot(rt, v); 3 = |s| test; 4 = omega;
// 5 = end of iteration.  BiCG: 0, 1 (|r| test, rho, beta), 2 (alpha), 5.
__device__ void krylov_step(KrylovState* __restrict__ st, const double2* __restrict__ part,
                            int phase);

// skip_t (the second matvec of BiCGSTAB and its dots) is set whenever the
// solve is done or the half-step test converged
__global__ void krylov_scalar_kernel(KrylovState* __restrict__ st,
                                     const double2* __restrict__ part, int phase) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (!st->done) krylov_step(st, part, phase);
  st->skip_t = (st->done || st->half) ? 1 : 0;
}

__device__ void krylov_step(KrylovState* __restrict__ st, const double2* __restrict__ part,
                            int phase) {
  const double2 S0 = sum_partials2(part);
  if (phase == 0) {
    const double bnrm2 = sqrt(S0.x);
    st->tol = fmax(st->atol, st->rtol * bnrm2);
    if (bnrm2 == 0.0) { st->done = 1; st->info = 0; }
    return;
  }
  if (phase == 1) {
    st->half = 0;
    if (sqrt(S0.x) < st->tol) { st->done = 1; st->info = 0; return; }
    const double2 rho = sum_partials2(part + PNH_LSQR_BLOCKS);
    if (zabs(rho) < st->rhotol) { st->done = 1; st->info = -10; return; }
    if (st->it > 0) {
      if (st->stab) {
        if (zabs(st->omega) < st->rhotol) { st->done = 1; st->info = -11; return; }
        st->beta = zmul(cdivd(rho, st->rho_prev), cdivd(st->alpha, st->omega));
      } else {
        st->beta = cdivd(rho, st->rho_prev);
      }
    }
    st->rho = rho;
    return;
  }
  if (phase == 2) {
    if (S0.x == 0.0 && S0.y == 0.0) { st->done = 1; st->info = -11; return; }
    st->alpha = cdivd(st->rho, S0);
    return;
  }
  if (phase == 3) {
    st->half = sqrt(S0.x) < st->tol ? 1 : 0;
    return;
  }
  if (phase == 4) {
    if (!st->half) st->omega = cdivd(S0, sum_partials2(part + PNH_LSQR_BLOCKS));
    return;
  }
  // phase 5
  st->rho_prev = st->rho;
  st->it += 1;
  if (st->half) { st->done = 1; st->info = 0; }
  else if (st->it >= st->maxiter) { st->done = 1; st->info = st->maxiter; }
}

// BiCGSTAB p: first iteration p = r, else p = (p - omega v) beta + r
// BiCG p, pt: p = p beta + r, pt = pt conj(beta) + rt (first: copies)
__global__ void krylov_p_kernel(double2* __restrict__ p, const double2* __restrict__ r,
                                const double2* __restrict__ v, double2* __restrict__ pt,
                                const double2* __restrict__ rt,
                                const KrylovState* __restrict__ st, int64_t n) {
  if (st->done) return;
  const bool first = st->it == 0;
  const double2 beta = st->beta, om = st->omega;
  GRID_STRIDE(i, n) {
    if (first) {
      p[i] = r[i];
      if (pt) pt[i] = rt[i];
      continue;
    }
    double2 pv = p[i];
    if (st->stab) {
      const double2 ov = zmul(om, v[i]);
      pv = make_double2(pv.x - ov.x, pv.y - ov.y);
    }
    pv = zmul(pv, beta);
    p[i] = make_double2(pv.x + r[i].x, pv.y + r[i].y);
    if (pt) {
      const double2 q = zmul(pt[i], zconj(beta));
      pt[i] = make_double2(q.x + rt[i].x, q.y + rt[i].y);
    }
  }
}

// r -= alpha v; partial |r|^2 (BiCGSTAB s = r)
__global__ void __launch_bounds__(256)
krylov_r_kernel(double2* __restrict__ r, const double2* __restrict__ v,
                const KrylovState* __restrict__ st, int64_t n, double2* __restrict__ part) {
  if (st->done) return;
  const double2 a = st->alpha;
  double acc = 0.0;
  GRID_STRIDE(i, n) {
    const double2 av = zmul(a, v[i]);
    const double2 rv = make_double2(r[i].x - av.x, r[i].y - av.y);
    r[i] = rv;
    acc += rv.x * rv.x + rv.y * rv.y;
  }
  block_sum2_store(make_double2(acc, 0.0), part);
}

// BiCGSTAB end of iteration: x += alpha p; unless |s| converged: x += omega s,
// r -= omega t.  BiCG: x += alpha p; r -= alpha q; rt -= conj(alpha) qt.
__global__ void krylov_x_kernel(double2* __restrict__ x, const double2* __restrict__ p,
                                double2* __restrict__ r, const double2* __restrict__ t,
                                double2* __restrict__ rt, const double2* __restrict__ qt,
                                const KrylovState* __restrict__ st, int64_t n) {
  if (st->done) return;
  const double2 a = st->alpha, om = st->omega;
  const bool stab = st->stab, half = st->half;
  GRID_STRIDE(i, n) {
    double2 xv = x[i];
    const double2 ap = zmul(a, p[i]);
    xv = make_double2(xv.x + ap.x, xv.y + ap.y);
    if (stab) {
      if (!half) {
        const double2 os = zmul(om, r[i]);
        xv = make_double2(xv.x + os.x, xv.y + os.y);
        const double2 ot = zmul(om, t[i]);
        r[i] = make_double2(r[i].x - ot.x, r[i].y - ot.y);
      }
    } else {
      const double2 aq = zmul(a, t[i]);
      r[i] = make_double2(r[i].x - aq.x, r[i].y - aq.y);
      const double2 aqt = zmul(zconj(a), qt[i]);
      rt[i] = make_double2(rt[i].x - aqt.x, rt[i].y - aqt.y);
    }
    x[i] = xv;
  }
}

}  // namespace

int pnh_krylov_init(KrylovState* st, int stab, double rtol, double atol, int maxiter,
                    hipStream_t s) {
  hipLaunchKernelGGL(krylov_init_kernel, dim3(1), dim3(64), 0, s, st, stab, rtol, atol,
                     maxiter);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_krylov_vdot2(const void* a0, const void* b0, const void* a1, const void* b1, int64_t n,
                     void* part, const int* skip, hipStream_t s) {
  hipLaunchKernelGGL(vdot2_kernel, dim3(PNH_LSQR_BLOCKS), dim3(256), 0, s, (const double2*)a0,
                     (const double2*)b0, (const double2*)a1, (const double2*)b1, n,
                     (double2*)part, skip);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_krylov_scalar(KrylovState* st, const void* part, int phase, hipStream_t s) {
  hipLaunchKernelGGL(krylov_scalar_kernel, dim3(1), dim3(64), 0, s, st, (const double2*)part,
                     phase);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_krylov_p(void* p, const void* r, const void* v, void* pt, const void* rt,
                 const KrylovState* st, int64_t n, hipStream_t s) {
  LAUNCH(krylov_p_kernel, n, (double2*)p, (const double2*)r, (const double2*)v, (double2*)pt,
         (const double2*)rt, st, n);
  return 0;
}

int pnh_krylov_r(void* r, const void* v, const KrylovState* st, int64_t n, void* part,
                 hipStream_t s) {
  hipLaunchKernelGGL(krylov_r_kernel, dim3(PNH_LSQR_BLOCKS), dim3(256), 0, s, (double2*)r,
                     (const double2*)v, st, n, (double2*)part);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_krylov_x(void* x, const void* p, void* r, const void* t, void* rt, const void* qt,
                 const KrylovState* st, int64_t n, hipStream_t s) {
  LAUNCH(krylov_x_kernel, n, (double2*)x, (const double2*)p, (double2*)r, (const double2*)t,
         (double2*)rt, (const double2*)qt, st, n);
  return 0;
}

// ----------------------------------------------------------------- GMRES --
// scipy.sparse.linalg.gmres (restarted, modified Gram-Schmidt, Givens
// rotations, no preconditioner) on P^H P in FP64.  The Hessenberg matrix,
// rotations and right-hand side live in GmresState; the host launches every
// column of every restart cycle and the kernels skip once a cycle or the
// solve has ended.
namespace {

// complex Givens rotation (c real): [c s; -conj(s) c] [f; g] = [r; 0]
__device__ void zlartg(double2 f, double2 g, double& c, double2& s, double2& r) {
  if (g.x == 0.0 && g.y == 0.0) {
    c = 1.0; s = make_double2(0.0, 0.0); r = f;
    return;
  }
  if (f.x == 0.0 && f.y == 0.0) {
    const double d = hypot(g.x, g.y);
    c = 0.0;
    s = make_double2(g.x / d, -g.y / d);
    r = make_double2(d, 0.0);
    return;
  }
  const double f2 = f.x * f.x + f.y * f.y, g2 = g.x * g.x + g.y * g.y;
  const double h2 = f2 + g2;
  c = sqrt(f2 / h2);
  r = make_double2(f.x / c, f.y / c);
  const double d = sqrt(f2 * h2);
  const double2 fd = make_double2(f.x / d, f.y / d);
  s = zmul(zconj(g), fd);
}

__global__ void gmres_init_kernel(GmresState* __restrict__ st, int restart, double rtol,
                                  double atol, int maxiter, int flex) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  st->flex = flex;
  st->restart = restart;
  st->rtol = rtol;
  st->atol = atol;
  st->maxiter = maxiter;
  st->ptol_max_factor = 1.0;
  st->done = 0;
  st->info = 0;
  st->outer = 0;
  st->inner_iter = 0;
  st->skip = 0;
}

// phase 0: |b| (tolerances); 1: start of a cycle from |r|; 2: h0 = |w|;
// 3 (col, k): H[col][k] = vdot(v_k, w); 4 (col): end of a column
// (h1, rotations, residual estimate); 6: |r| after the cycle, restart
// tolerance.  skip is set whenever the cycle or the solve has ended.
__global__ void gmres_scalar_kernel(GmresState* __restrict__ st,
                                    const double2* __restrict__ part, int phase, int col,
                                    int k) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (st->done) { st->skip = 1; return; }
  const double eps = 2.220446049250313e-16;
  const double2 S0 = sum_partials2(part);
  if (phase == 0) {
    st->bnrm2 = sqrt(S0.x);
    st->tol = fmax(st->atol, st->rtol * st->bnrm2);
    if (st->bnrm2 == 0.0) { st->done = 1; st->skip = 1; st->info = 0; return; }
    st->ptol = st->bnrm2 * fmin(st->ptol_max_factor, st->tol / st->bnrm2);
    return;
  }
  if (phase == 1) {
    const double nr = sqrt(S0.x);
    if (st->outer == 0 && nr < st->tol) { st->done = 1; st->skip = 1; st->info = 0; return; }
    st->vnorm = nr;
    for (int i = 0; i <= st->restart; ++i) st->S[i] = make_double2(0.0, 0.0);
    st->S[0] = make_double2(nr, 0.0);
    st->breakdown = 0;
    st->col_end = 0;
    st->skip = 0;
    return;
  }
  if (st->skip && phase >= 2 && phase <= 4) return;
  if (phase == 2) { st->h0 = sqrt(S0.x); return; }
  if (phase == 3) { st->H[col][k] = S0; st->H0[col][k] = S0; return; }
  if (phase == 4) {
    const double h1 = sqrt(S0.x);
    st->H[col][col + 1] = make_double2(h1, 0.0);
    st->H0[col][col + 1] = make_double2(h1, 0.0);
    st->vnorm = h1;
    if (st->flex) {
      if (!(h1 > eps * st->h0)) st->breakdown = 1;
    } else if (h1 <= eps * st->h0) {
      st->H[col][col + 1] = make_double2(0.0, 0.0);
      st->breakdown = 1;
    }
    for (int j = 0; j < col; ++j) {
      const double c = st->gc[j];
      const double2 s = st->gs[j];
      const double2 n0 = st->H[col][j], n1 = st->H[col][j + 1];
      const double2 sn1 = zmul(s, n1), csn0 = zmul(zconj(s), n0);
      st->H[col][j] = make_double2(c * n0.x + sn1.x, c * n0.y + sn1.y);
      st->H[col][j + 1] = make_double2(-csn0.x + c * n1.x, -csn0.y + c * n1.y);
    }
    double c;
    double2 s, mag;
    zlartg(st->H[col][col], st->H[col][col + 1], c, s, mag);
    st->gc[col] = c;
    st->gs[col] = s;
    st->H[col][col] = mag;
    st->H[col][col + 1] = make_double2(0.0, 0.0);
    const double2 sc = zmul(zconj(s), st->S[col]);
    const double2 tmp = make_double2(-sc.x, -sc.y);
    st->S[col] = make_double2(c * st->S[col].x, c * st->S[col].y);
    st->S[col + 1] = tmp;
    st->presid = hypot(tmp.x, tmp.y);
    st->inner_iter += 1;
    st->col_end = col;
    if (st->flex ? (st->presid < st->ptol || st->breakdown)
                 : (st->presid <= st->ptol || st->breakdown))
      st->skip = 1;
    return;
  }
  if (phase == 7) {
    // LGMRES outer start: S0 = |b - A x|^2 (the cycle's first vector)
    const double rn = sqrt(S0.x);
    st->rnorm = rn;
    if (rn <= st->tol) { st->done = 1; st->skip = 1; st->info = 0; return; }
    st->vnorm = rn;
    st->ptol = fmin(st->ptol_max_factor, st->tol / rn);
    // the inner least-squares problem starts from the unit vector (its
    // residual estimate is relative, as ptol); y is scaled by |r| after
    for (int i = 0; i <= st->restart; ++i) st->S[i] = make_double2(0.0, 0.0);
    st->S[0] = make_double2(1.0, 0.0);
    st->breakdown = 0;
    st->col_end = 0;
    st->skip = 0;
    return;
  }
  if (phase == 8) {
    // LGMRES: |dx| and q = H0 y (H0 as built, before the rotations)
    st->nx = sqrt(S0.x);
    const int c = st->col_end;
    for (int i = 0; i <= c + 1; ++i) {
      double2 acc = make_double2(0.0, 0.0);
      for (int j = 0; j <= c; ++j) {
        if (i > j + 1) continue;
        const double2 t = zmul(st->H0[j][i], st->y[j]);
        acc.x += t.x;
        acc.y += t.y;
      }
      st->q[i] = acc;
    }
    return;
  }
  if (phase == 9) {
    // LGMRES end of an outer iteration: inner tolerance schedule, limit
    if (st->presid > st->ptol) st->ptol_max_factor = fmin(1.0, 1.5 * st->ptol_max_factor);
    else st->ptol_max_factor = fmax(1e-16, 0.25 * st->ptol_max_factor);
    st->outer += 1;
    if (st->outer >= st->maxiter) { st->done = 1; st->info = st->maxiter; }
    st->skip = st->done;
    return;
  }
  // phase 6: rnorm after the cycle
  const double rnorm = sqrt(S0.x);
  st->rnorm = rnorm;
  st->outer += 1;
  if (rnorm <= st->tol || st->breakdown) {
    st->done = 1;
  } else {
    if (st->presid <= st->ptol) st->ptol_max_factor = fmax(eps, 0.25 * st->ptol_max_factor);
    else st->ptol_max_factor = fmin(1.0, 1.5 * st->ptol_max_factor);
    st->ptol = st->presid * fmin(st->ptol_max_factor, st->tol / rnorm);
    if (st->outer >= st->maxiter) st->done = 1;
  }
  if (st->done) st->info = rnorm <= st->tol ? 0 : st->maxiter;
  st->skip = st->done;
}

// the cycle's triangular solve (one thread); runs whether or not the last
// columns were skipped
__global__ void gmres_solve_kernel(GmresState* __restrict__ st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (st->done) return;
  const int col = st->col_end;
  if (st->H[col][col].x == 0.0 && st->H[col][col].y == 0.0)
    st->S[col] = make_double2(0.0, 0.0);
  for (int i = 0; i <= col; ++i) st->y[i] = st->S[i];
  for (int kk = col; kk > 0; --kk) {
    if (st->y[kk].x != 0.0 || st->y[kk].y != 0.0) {
      st->y[kk] = cdivd(st->y[kk], st->H[kk][kk]);
      const double2 t = st->y[kk];
      for (int i = 0; i < kk; ++i) {
        const double2 th = zmul(t, st->H[kk][i]);
        st->y[i] = make_double2(st->y[i].x - th.x, st->y[i].y - th.y);
      }
    }
  }
  if (st->y[0].x != 0.0 || st->y[0].y != 0.0) st->y[0] = cdivd(st->y[0], st->H[0][0]);
  if (st->flex)
    for (int i = 0; i <= col; ++i) st->y[i] = make_double2(st->y[i].x * st->rnorm, st->y[i].y * st->rnorm);
}

// w -= H[col][k] v_k
__global__ void gmres_mgs_kernel(double2* __restrict__ w, const double2* __restrict__ vk,
                                 const GmresState* __restrict__ st, int col, int k,
                                 int64_t n) {
  if (st->done || st->skip) return;
  const double2 t = st->H[col][k];
  GRID_STRIDE(i, n) {
    const double2 tv = zmul(t, vk[i]);
    w[i] = make_double2(w[i].x - tv.x, w[i].y - tv.y);
  }
}

// out = in / vnorm: the next basis vector (not needed, so skipped, once the
// cycle or the solve has ended)
__global__ void gmres_normalize_kernel(double2* __restrict__ out, const double2* __restrict__ in,
                                       const GmresState* __restrict__ st, int64_t n, int col) {
  // LGMRES keeps the basis vector of the column that ended the cycle (its
  // A dx uses it); GMRES does not need it
  if (st->done || (st->skip && !(st->flex && col >= 0 && col == st->col_end))) return;
  const double nv = st->vnorm;
  const double inv = nv > 0.0 ? 1.0 / nv : 1.0;
  GRID_STRIDE(i, n) {
    double2 v = in[i];
    if (nv > 0.0) { v.x *= inv; v.y *= inv; }
    out[i] = v;
  }
}

// x += sum_{k <= col_end} y_k v_k
__global__ void gmres_x_kernel(double2* __restrict__ x, const double2* __restrict__ V,
                               int64_t ldv, const GmresState* __restrict__ st, int64_t n) {
  if (st->done) return;
  const int col = st->col_end;
  GRID_STRIDE(i, n) {
    double2 acc = make_double2(0.0, 0.0);
    for (int kk = 0; kk <= col; ++kk) {
      const double2 t = zmul(st->y[kk], V[kk * ldv + i]);
      acc.x += t.x;
      acc.y += t.y;
    }
    x[i] = make_double2(x[i].x + acc.x, x[i].y + acc.y);
  }
}

// r = b - t; partials of |r|^2
__global__ void __launch_bounds__(256)
gmres_residual_kernel(double2* __restrict__ r, const double2* __restrict__ b,
                      const double2* __restrict__ t, const GmresState* __restrict__ st,
                      int64_t n, double2* __restrict__ part) {
  if (st->done) return;
  double acc = 0.0;
  GRID_STRIDE(i, n) {
    const double2 v = make_double2(b[i].x - t[i].x, b[i].y - t[i].y);
    r[i] = v;
    acc += v.x * v.x + v.y * v.y;
  }
  block_sum2_store(make_double2(acc, 0.0), part);
}

}  // namespace

namespace {

__global__ void __launch_bounds__(256)
lgmres_dx_kernel(double2* __restrict__ dx, const double2* __restrict__ V, int64_t ldv,
                 int inner_m, LgmresZ zo, const GmresState* __restrict__ st, int64_t n,
                 double2* __restrict__ part) {
  if (st->done) return;
  const int col = st->col_end;
  double acc2 = 0.0;
  GRID_STRIDE(i, n) {
    double2 acc = make_double2(0.0, 0.0);
    for (int c = 0; c <= col; ++c) {
      const double2* z = c < inner_m ? V + c * ldv : (const double2*)zo.z[c - inner_m];
      const double2 t = zmul(st->y[c], z[i]);
      acc.x += t.x;
      acc.y += t.y;
    }
    dx[i] = acc;
    acc2 += acc.x * acc.x + acc.y * acc.y;
  }
  block_sum2_store(make_double2(acc2, 0.0), part);
}

__global__ void lgmres_store_kernel(double2* __restrict__ x, const double2* __restrict__ dx,
                                    const double2* __restrict__ V, int64_t ldv,
                                    double2* __restrict__ z_new, double2* __restrict__ az_new,
                                    const GmresState* __restrict__ st, int64_t n) {
  if (st->done) return;
  const int col = st->col_end;
  const double nx = st->nx;
  GRID_STRIDE(i, n) {
    const double2 d = dx[i];
    if (nx > 0.0 && z_new) {      // no augmentation vectors kept when outer_k == 0
      double2 ax = make_double2(0.0, 0.0);
      for (int c = 0; c <= col + 1; ++c) {
        const double2 t = zmul(st->q[c], V[c * ldv + i]);
        ax.x += t.x;
        ax.y += t.y;
      }
      z_new[i] = make_double2(d.x / nx, d.y / nx);
      az_new[i] = make_double2(ax.x / nx, ax.y / nx);
    }
    x[i] = make_double2(x[i].x + d.x, x[i].y + d.y);
  }
}

}  // namespace

int pnh_lgmres_dx(void* dx, const void* V, int64_t ldv, int inner_m, const LgmresZ& zo,
                  const GmresState* st, int64_t n, void* part, hipStream_t s) {
  hipLaunchKernelGGL(lgmres_dx_kernel, dim3(PNH_LSQR_BLOCKS), dim3(256), 0, s, (double2*)dx,
                     (const double2*)V, ldv, inner_m, zo, st, n, (double2*)part);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_lgmres_store(void* x, const void* dx, const void* V, int64_t ldv, void* z_new,
                     void* az_new, const GmresState* st, int64_t n, hipStream_t s) {
  LAUNCH(lgmres_store_kernel, n, (double2*)x, (const double2*)dx, (const double2*)V, ldv,
         (double2*)z_new, (double2*)az_new, st, n);
  return 0;
}

int pnh_gmres_init(GmresState* st, int restart, double rtol, double atol, int maxiter,
                   int flex, hipStream_t s) {
  hipLaunchKernelGGL(gmres_init_kernel, dim3(1), dim3(64), 0, s, st, restart, rtol, atol,
                     maxiter, flex);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_gmres_scalar(GmresState* st, const void* part, int phase, int col, int k,
                     hipStream_t s) {
  hipLaunchKernelGGL(gmres_scalar_kernel, dim3(1), dim3(64), 0, s, st, (const double2*)part,
                     phase, col, k);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_gmres_solve(GmresState* st, hipStream_t s) {
  hipLaunchKernelGGL(gmres_solve_kernel, dim3(1), dim3(64), 0, s, st);
  PNH_LAUNCH_CHECK();
  return 0;
}

int pnh_gmres_mgs(void* w, const void* vk, const GmresState* st, int col, int k, int64_t n,
                  hipStream_t s) {
  LAUNCH(gmres_mgs_kernel, n, (double2*)w, (const double2*)vk, st, col, k, n);
  return 0;
}

int pnh_gmres_normalize(void* out, const void* in, const GmresState* st, int64_t n, int col,
                        hipStream_t s) {
  LAUNCH(gmres_normalize_kernel, n, (double2*)out, (const double2*)in, st, n, col);
  return 0;
}

int pnh_gmres_x(void* x, const void* V, int64_t ldv, const GmresState* st, int64_t n,
                hipStream_t s) {
  LAUNCH(gmres_x_kernel, n, (double2*)x, (const double2*)V, ldv, st, n);
  return 0;
}

int pnh_gmres_residual(void* r, const void* b, const void* t, const GmresState* st, int64_t n,
                       void* part, hipStream_t s) {
  hipLaunchKernelGGL(gmres_residual_kernel, dim3(PNH_LSQR_BLOCKS), dim3(256), 0, s,
                     (double2*)r, (const double2*)b, (const double2*)t, st, n, (double2*)part);
  PNH_LAUNCH_CHECK();
  return 0;
}
