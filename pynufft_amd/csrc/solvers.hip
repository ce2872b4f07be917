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

// uker = mu w - lambda Lap, Lap = sum_d 2 cos(2 pi k_d / K_d) - 2 (the FFT
// of the periodic Laplacian stencil), evaluated in FP64 and rounded once.
__global__ void uker_kernel(const float* __restrict__ w, float* __restrict__ uker,
                            ShapeArgs s, float mu, float lam) {
  GRID_STRIDE(i, s.n) {
    int64_t rest = i;
    double lap = 0.0;
    for (int d = s.D - 1; d >= 0; --d) {
      const int64_t c = rest % s.dim[d];
      rest /= s.dim[d];
      lap += 2.0 * cospi(2.0 * (double)c / (double)s.dim[d]) - 2.0;
    }
    uker[i] = mu * w[i] - lam * (float)lap;
  }
}

// k = (k + 1e-7) / (uker + 1e-7): the k-space solve of (mu W - lambda Lap)
__global__ void kdiv_kernel(float2* __restrict__ k, const float* __restrict__ uker,
                            int64_t n) {
  GRID_STRIDE(i, n) {
    const float d = uker[i] + 1e-7f;
    const float2 v = k[i];
    k[i] = make_float2((v.x + 1e-7f) / d, v.y / d);
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

int pnh_l1tv_uker(const float* w, float* uker, const ShapeArgs& sh, float mu, float lam,
                  hipStream_t s) {
  LAUNCH(uker_kernel, sh.n, w, uker, sh, mu, lam);
  return 0;
}

int pnh_l1tv_kdiv(void* k, const float* uker, int64_t n, hipStream_t s) {
  LAUNCH(kdiv_kernel, n, (float2*)k, uker, n);
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
