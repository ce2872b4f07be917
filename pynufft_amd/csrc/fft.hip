// LDS-resident Stockham FFT over lines of a strided N-d array (gfx950).
//
// One launch transforms every line along one axis.  A workgroup loads LB lines
// into LDS (zero-filling past n_in, so zero-padding is fused), runs all radix
// passes in LDS (ping-pong buffers, radices 8/4/2/3/5 and a generic small odd
// radix), and writes the first n_out outputs of each line (so cropping is
// fused).  The real apodization factor sn and a scalar may be applied on the
// load side (forward: x2xx fused) or on the store side (adjoint: xx2x fused).
//
// Semantics follow numpy.fft: the forward kernel is exp(-2 pi i t/n), the
// inverse exp(+2 pi i t/n) with the 1/n factors folded into `scale`.
#include "common.h"
#include "fft.h"

namespace {

template <typename C, typename T>
__device__ __forceinline__ C mul_neg_i(C a, int sign) {
  // multiply by (sign * i): sign=-1 -> -i, sign=+1 -> +i
  C z;
  if (sign < 0) { z.x = a.y; z.y = -a.x; } else { z.x = -a.y; z.y = a.x; }
  return z;
}

template <typename T>
__device__ __forceinline__ void dft2(typename cplx<T>::t* v) {
  typename cplx<T>::t a = v[0], b = v[1];
  v[0] = cadd(a, b); v[1] = csub(a, b);
}

template <typename T>
__device__ __forceinline__ void dft4(typename cplx<T>::t* v, int sign) {
  typedef typename cplx<T>::t C;
  C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  C t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
  C t3r = mul_neg_i<C, T>(t3, sign);  // sign*i*(v1-v3)
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3r);
  v[3] = csub(t1, t3r);
}

template <typename T>
__device__ __forceinline__ void dft8(typename cplx<T>::t* v, int sign) {
  typedef typename cplx<T>::t C;
  const T h = (T)0.70710678118654752440;
  // even/odd split: E = DFT4(v0,v2,v4,v6), O = DFT4(v1,v3,v5,v7)
  C e[4] = {v[0], v[2], v[4], v[6]};
  C o[4] = {v[1], v[3], v[5], v[7]};
  dft4<T>(e, sign);
  dft4<T>(o, sign);
  // twiddles w8^q = exp(sign*2*pi*i*q/8), q = 0..3
  C o1, o2, o3;
  T s = (T)sign;
  // q=1: (1 + s i) h
  o1.x = (o[1].x - s * o[1].y) * h;
  o1.y = (o[1].y + s * o[1].x) * h;
  // q=2: s i
  o2 = mul_neg_i<C, T>(o[2], sign);
  // q=3: (-1 + s i) h
  o3.x = (-o[3].x - s * o[3].y) * h;
  o3.y = (-o[3].y + s * o[3].x) * h;
  v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
}

// 16-point DFT as 4 x 4: DFTs over n1 of x[4 n1 + n2], twiddles
// W16^(n2 k1), DFTs over n2; W = exp(sign 2 pi i / 16).
template <typename T>
__device__ __forceinline__ void dft16(typename cplx<T>::t* v, int sign) {
  typedef typename cplx<T>::t C;
  const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173;
  const T h = (T)0.70710678118654752440;
  const T sg = (T)sign;
  C a[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    C t[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
    dft4<T>(t, sign);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = t[k1];
  }
  // W16^e for e = n2 * k1 in {1, 2, 3, 4, 6, 9}
  const C w1 = mk<T>(c1, sg * s1), w2 = mk<T>(h, sg * h), w3 = mk<T>(s1, sg * c1);
  const C w4 = mk<T>((T)0, sg), w6 = mk<T>(-h, sg * h), w9 = mk<T>(-c1, -sg * s1);
  a[1][1] = cmul(a[1][1], w1); a[1][2] = cmul(a[1][2], w2); a[1][3] = cmul(a[1][3], w3);
  a[2][1] = cmul(a[2][1], w2); a[2][2] = cmul(a[2][2], w4); a[2][3] = cmul(a[2][3], w6);
  a[3][1] = cmul(a[3][1], w3); a[3][2] = cmul(a[3][2], w6); a[3][3] = cmul(a[3][3], w9);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    C t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
    dft4<T>(t, sign);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
  }
}

template <typename T>
__device__ __forceinline__ void dft3(typename cplx<T>::t* v, int sign) {
  typedef typename cplx<T>::t C;
  const T c = (T)-0.5, s = (T)0.86602540378443864676 * (T)sign;
  C t1 = cadd(v[1], v[2]);
  C t2 = csub(v[1], v[2]);
  C m = cadd(v[0], cscale(t1, c));
  // s*i*t2
  C r; r.x = -s * t2.y; r.y = s * t2.x;
  v[0] = cadd(v[0], t1);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <typename T>
__device__ __forceinline__ void dft5(typename cplx<T>::t* v, int sign) {
  typedef typename cplx<T>::t C;
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)0.95105651629515357212 * (T)sign;
  const T s2 = (T)0.58778525229247312917 * (T)sign;
  C a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
  C a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
  C m1 = cadd(v[0], cadd(cscale(a1, c1), cscale(a2, c2)));
  C m2 = cadd(v[0], cadd(cscale(a1, c2), cscale(a2, c1)));
  // i*(s1 b1 + s2 b2), i*(s2 b1 - s1 b2)
  C n1 = cadd(cscale(b1, s1), cscale(b2, s2));
  C n2 = csub(cscale(b1, s2), cscale(b2, s1));
  C r1; r1.x = -n1.y; r1.y = n1.x;
  C r2; r2.x = -n2.y; r2.y = n2.x;
  v[0] = cadd(v[0], cadd(a1, a2));
  v[1] = cadd(m1, r1); v[4] = csub(m1, r1);
  v[2] = cadd(m2, r2); v[3] = csub(m2, r2);
}

// Generic small odd radix R via the length-n twiddle table: w_R^k = tw[k*(n/R)].
template <typename T, int R>
__device__ __forceinline__ void dftg(typename cplx<T>::t* v, int n,
                                     const typename cplx<T>::t* tw, int sign) {
  typedef typename cplx<T>::t C;
  C out[R];
  const int step = n / R;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    C acc = v[0];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      C w = tw[((r * q) % R) * step];
      if (sign > 0) w.y = -w.y;
      acc = cadd(acc, cmul(v[r], w));
    }
    out[q] = acc;
  }
#pragma unroll
  for (int q = 0; q < R; ++q) v[q] = out[q];
}

template <typename T, int R>
__device__ __forceinline__ void dft_r(typename cplx<T>::t* v, int n,
                                      const typename cplx<T>::t* tw, int sign) {
  if (R == 2) dft2<T>(v);
  else if (R == 3) dft3<T>(v, sign);
  else if (R == 4) dft4<T>(v, sign);
  else if (R == 5) dft5<T>(v, sign);
  else if (R == 8) dft8<T>(v, sign);
  else if (R == 16) dft16<T>(v, sign);
  else dftg<T, R>(v, n, tw, sign);
}

// q = i / n and r = i % n for 0 <= i < 2^22 and 1 <= n, with rcp = 1.0f / n
// (the float estimate is off by at most one; one correction step fixes it)
__device__ __forceinline__ int fdivmod(int i, int n, float rcp, int& r) {
  int q = (int)((float)i * rcp);
  r = i - q * n;
  if (r < 0) { --q; r += n; }
  else if (r >= n) { ++q; r -= n; }
  return q;
}

// i / n, i % n: shifts for powers of two, else the float reciprocal
template <bool POW2>
__device__ __forceinline__ int divmod_n(int i, int n, int log2n, float rcp, int& r) {
  if (POW2) { r = i & (n - 1); return i >> log2n; }
  return fdivmod(i, n, rcp, r);
}

// One Stockham radix-R pass over LB lines.  Butterfly (l, j), j < n/R,
// reads in(l, j + r n/R) for r < R, applies the twiddles of a stage after
// ns points, and writes out(l, (j / ns) ns R + j % ns + r ns).  With
// lines_fast, consecutive threads take consecutive lines (coalesced global
// access along a strided axis); else consecutive j.
template <typename T, int R, bool POW2, typename LD, typename ST>
__device__ __forceinline__ void stockham_pass(int n, int LB, int ns, bool lines_fast,
                                              const typename cplx<T>::t* tw, int sign,
                                              int tid, int nthr, LD load, ST store) {
  typedef typename cplx<T>::t C;
  const int nb = n / R;
  const int twstep = n / (ns * R);
  const int total = LB * nb;
  const int lg_nb = POW2 ? __builtin_ctz(nb) : 0;
  const int lg_ns = POW2 ? __builtin_ctz(ns) : 0;
  const bool lb_pow2 = (LB & (LB - 1)) == 0;
  const int lg_lb = lb_pow2 ? __builtin_ctz(LB) : 0;
  const float r_nb = 1.0f / (float)nb, r_ns = 1.0f / (float)ns, r_lb = 1.0f / (float)LB;
  for (int idx = tid; idx < total; idx += nthr) {
    int l, j, k;
    if (lines_fast) {
      if (lb_pow2) { l = idx & (LB - 1); j = idx >> lg_lb; }
      else j = fdivmod(idx, LB, r_lb, l);
    } else {
      l = divmod_n<POW2>(idx, nb, lg_nb, r_nb, j);
    }
    const int jq = divmod_n<POW2>(j, ns, lg_ns, r_ns, k);
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = load(l, j + r * nb);
    if (ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        C w = tw[r * k * twstep];
        if (sign > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    dft_r<T, R>(v, n, tw, sign);
    const int o = jq * ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) store(l, o + r * ns, v[r]);
  }
}

template <typename T, bool POW2, typename LD, typename ST>
__device__ __forceinline__ void run_pass(int R, int n, int LB, int ns, bool lines_fast,
                                         const typename cplx<T>::t* tw, int sign, int tid,
                                         int nthr, LD load, ST store) {
  if (POW2) {
    if (R == 16) stockham_pass<T, 16, true>(n, LB, ns, lines_fast, tw, sign, tid, nthr, load, store);
    else if (R == 8) stockham_pass<T, 8, true>(n, LB, ns, lines_fast, tw, sign, tid, nthr, load, store);
    else if (R == 4) stockham_pass<T, 4, true>(n, LB, ns, lines_fast, tw, sign, tid, nthr, load, store);
    else stockham_pass<T, 2, true>(n, LB, ns, lines_fast, tw, sign, tid, nthr, load, store);
    return;
  }
  switch (R) {
    case 2: stockham_pass<T, 2, false>(n, LB, ns, lines_fast, tw, sign, tid, nthr, load, store); break;
    case 3: stockham_pass<T, 3, false>(n, LB, ns, lines_fast, tw, sign, tid, nthr, load, store); break;
    case 4: stockham_pass<T, 4, false>(n, LB, ns, lines_fast, tw, sign, tid, nthr, load, store); break;
    case 5: stockham_pass<T, 5, false>(n, LB, ns, lines_fast, tw, sign, tid, nthr, load, store); break;
    case 7: stockham_pass<T, 7, false>(n, LB, ns, lines_fast, tw, sign, tid, nthr, load, store); break;
    case 8: stockham_pass<T, 8, false>(n, LB, ns, lines_fast, tw, sign, tid, nthr, load, store); break;
    case 11: stockham_pass<T, 11, false>(n, LB, ns, lines_fast, tw, sign, tid, nthr, load, store); break;
    case 13: stockham_pass<T, 13, false>(n, LB, ns, lines_fast, tw, sign, tid, nthr, load, store); break;
    default: stockham_pass<T, 16, false>(n, LB, ns, lines_fast, tw, sign, tid, nthr, load, store); break;
  }
}

template <typename T>
__device__ __forceinline__ T ld_sn(const void* p, int64_t i) {
  return (T)((const T*)p)[i];
}

// Offsets of line `l` (flattened over outer x inner) in a buffer.
__device__ __forceinline__ int64_t line_base(const FftSide& s, int64_t o0,
                                             int64_t o1, int64_t b) {
  return o0 * s.o_stride[0] + o1 * s.o_stride[1] + b * s.b_stride;
}

// POW2: every radix is 2, 4, 8 or 16 (shift/mask indexing).  The first
// pass reads its inputs from global memory (zero padding and load-side
// apodization fused) and the last pass writes its outputs there (crop,
// scale and store-side apodization fused); the passes in between go through
// LDS.  in_ and out_ may alias (in-place passes): a workgroup reads all its
// lines before its first store, and other workgroups own other lines.
// LDS: one line buffer (two with three or more passes), lines of
// n + n/16 + 1 elements (one pad element per 16 against bank conflicts),
// then per-line source/destination/apodization offsets.
template <typename T, bool POW2>
__device__ __forceinline__ void fft_body(const FftDesc& d, const void* in_, void* out_,
                                         const void* __restrict__ tw_, int64_t blk,
                                         unsigned char* smem_raw) {
  typedef typename cplx<T>::t C;
  const int n = d.n;
  const int LB = d.lines_per_block;
  const int ld = pnh_fft_lds_line(n);
  const int nbuf = d.npass >= 3 ? 2 : (d.npass == 2 ? 1 : 0);
  C* buf0 = reinterpret_cast<C*>(smem_raw);
  C* buf1 = buf0 + (size_t)LB * ld;
  int64_t* loff = reinterpret_cast<int64_t*>(buf0 + (size_t)nbuf * LB * ld);  // [3][LB]
  const C* in = (const C*)in_;
  C* out = (C*)out_;
  const C* __restrict__ tw = (const C*)tw_;

  const int64_t line0 = blk * LB;
  const int64_t n_lines = d.n_outer * d.n_inner;
  const int nthr = blockDim.x;
  const int tid = threadIdx.x;
  const int nl = (int)(n_lines - line0 < LB ? n_lines - line0 : LB);  // lines present

  if (tid < LB) {
    const int64_t L = line0 + (tid < nl ? tid : 0);
    const int64_t o = L / d.n_inner, b = L - o * d.n_inner;
    const int64_t o0 = o / d.o_split, o1 = o - o0 * d.o_split;
    loff[tid] = line_base(d.src, o0, o1, b);
    loff[LB + tid] = line_base(d.dst, o0, o1, b);
    loff[2 * LB + tid] = o0 * d.sn_o_stride[0] + o1 * d.sn_o_stride[1];
  }
  __syncthreads();

  const int64_t st = d.src.t_stride, dt = d.dst.t_stride;
  const int n_in = d.n_in, n_out = d.n_out;
  const int sn_side = d.sn_side;
  const T scale = (T)d.scale;
  auto gload = [&](int l, int t) -> C {
    C v = mk<T>((T)0, (T)0);
    if (l < nl && t < n_in) {
      v = in[loff[l] + (int64_t)t * st];
      if (sn_side == 1) {
        const T s = ld_sn<T>(d.sn, loff[2 * LB + l] + (int64_t)t * d.sn_t_stride);
        v.x *= s; v.y *= s;
      }
    }
    return v;
  };
  auto gstore = [&](int l, int t, C v) {
    if (l < nl && t < n_out) {
      T s = scale;
      if (sn_side == 2) s *= ld_sn<T>(d.sn, loff[2 * LB + l] + (int64_t)t * d.sn_t_stride);
      v.x *= s; v.y *= s;
      out[loff[LB + l] + (int64_t)t * dt] = v;
    }
  };
  auto at = [&](C* b, int l, int t) -> C& { return b[l * ld + t + (t >> 4)]; };
  const int sign = d.sign;
  const bool fast = d.inner_major != 0;

  if (d.npass == 1) {
    run_pass<T, POW2>(d.radix[0], n, LB, 1, fast, tw, sign, tid, nthr, gload, gstore);
    return;
  }
  run_pass<T, POW2>(d.radix[0], n, LB, d.ns[0], fast, tw, sign, tid, nthr, gload,
                    [&](int l, int t, C v) { at(buf0, l, t) = v; });
  __syncthreads();
  C* src = buf0;
  C* dst = buf1;
  for (int p = 1; p + 1 < d.npass; ++p) {
    run_pass<T, POW2>(d.radix[p], n, LB, d.ns[p], false, tw, sign, tid, nthr,
                      [&](int l, int t) -> C { return at(src, l, t); },
                      [&](int l, int t, C v) { at(dst, l, t) = v; });
    __syncthreads();
    C* tmp = src; src = dst; dst = tmp;
  }
  const int pl = d.npass - 1;
  run_pass<T, POW2>(d.radix[pl], n, LB, d.ns[pl], fast, tw, sign, tid, nthr,
                    [&](int l, int t) -> C { return at(src, l, t); }, gstore);
}

template <typename T, bool POW2>
__global__ void __launch_bounds__(256)
fft_lines_kernel(FftDesc d, const void* in_, void* out_,
                 const void* __restrict__ tw_, const int* __restrict__ skip) {
  if (skip && *skip) return;
  extern __shared__ unsigned char smem_raw[];
  fft_body<T, POW2>(d, in_, out_, tw_, blockIdx.x, smem_raw);
}

}  // namespace

int pnh_fft_launch(const FftDesc& d, int fp64, const void* in, void* out,
                   const void* tw, const int* skip, hipStream_t stream) {
  const int64_t n_lines = d.n_outer * d.n_inner;
  if (n_lines <= 0) return 0;
  const int64_t nblk = (n_lines + d.lines_per_block - 1) / d.lines_per_block;
  const size_t esz = fp64 ? 16 : 8;
  const size_t lds = pnh_fft_lds_bytes(d, fp64);
  dim3 grid((unsigned)nblk), block(d.threads);
  bool pow2 = true;
  for (int p = 0; p < d.npass; ++p)
    pow2 = pow2 && (d.radix[p] == 2 || d.radix[p] == 4 || d.radix[p] == 8 || d.radix[p] == 16);
  if (fp64) {
    if (pow2)
      hipLaunchKernelGGL((fft_lines_kernel<double, true>), grid, block, lds, stream, d,
                         in, out, tw, skip);
    else
      hipLaunchKernelGGL((fft_lines_kernel<double, false>), grid, block, lds, stream, d,
                         in, out, tw, skip);
  } else {
    if (pow2)
      hipLaunchKernelGGL((fft_lines_kernel<float, true>), grid, block, lds, stream, d,
                         in, out, tw, skip);
    else
      hipLaunchKernelGGL((fft_lines_kernel<float, false>), grid, block, lds, stream, d,
                         in, out, tw, skip);
  }
  PNH_LAUNCH_CHECK();
  return 0;
}

size_t pnh_fft_lds_bytes(const FftDesc& d, int fp64) {
  const size_t esz = fp64 ? 16 : 8;
  const int nbuf = d.npass >= 3 ? 2 : (d.npass == 2 ? 1 : 0);
  return (size_t)nbuf * d.lines_per_block * pnh_fft_lds_line(d.n) * esz +
         3 * (size_t)d.lines_per_block * sizeof(int64_t);
}

// Factor n into radices: 16 first, then 8, 4, 2, then 3, 5, then the other
// small odd primes.
int pnh_fft_factor(int n, int* radix, int* ns, int* npass) {
  int m = n, cnt = 0;
  int prod = 1;
  while (m > 1) {
    int R;
    if (m % 16 == 0) R = 16;
    else if (m % 8 == 0) R = 8;
    else if (m % 4 == 0) R = 4;
    else if (m % 2 == 0) R = 2;
    else if (m % 3 == 0) R = 3;
    else if (m % 5 == 0) R = 5;
    else {
      R = 0;
      const int generic[3] = {7, 11, 13};
      for (int q = 0; q < 3; ++q)
        if (m % generic[q] == 0) { R = generic[q]; break; }
      if (R == 0) {
        pnh_set_error("FFT length %d has a prime factor other than 2, 3, 5, 7, 11, 13", n);
        return -1;
      }
    }
    if (cnt >= PNH_FFT_MAX_PASS) {
      pnh_set_error("FFT length %d needs too many passes", n);
      return -1;
    }
    radix[cnt] = R;
    ns[cnt] = prod;
    prod *= R;
    m /= R;
    ++cnt;
  }
  *npass = cnt;
  return 0;
}

// ------------------------------------------------------------- batches ----
#include "batch.h"

namespace {

template <typename T, bool POW2>
__global__ void __launch_bounds__(256)
fft_batch_kernel(const BatchProblem* __restrict__ probs, BatchTable t, int axis, int kind,
                 const void* in_, int in_image, void* out_, int out_image,
                 const int* __restrict__ done, const int* __restrict__ all_done) {
  typedef typename cplx<T>::t C;
  if (all_done && *all_done) return;
  const int p = t.tab[blockIdx.x];
  if (done && done[p]) return;
  const int64_t lb = (int64_t)blockIdx.x - t.first[p];
  extern __shared__ unsigned char smem_raw[];
  const BatchProblem& pr = probs[p];
  // a reference into the problem table (uniform loads); a by-value copy
  // selected at run time went to per-thread scratch memory
  const FftDesc& d = kind == BFFT_FWD ? pr.fwd[axis]
                                      : (kind == BFFT_INV ? pr.inv[axis] : pr.inv64[axis]);
  const void* tw = kind == BFFT_INV64 ? pr.tw64[axis] : pr.tw32[axis];
  const C* in = (const C*)in_ + (in_image ? pr.x_off : pr.k_off);
  C* out = (C*)out_ + (out_image ? pr.x_off : pr.k_off);
  fft_body<T, POW2>(d, in, out, tw, lb, smem_raw);
}

}  // namespace

int pnh_batch_fft_launch(const BatchProblem* probs, const BatchTable& t, int axis,
                         int kind, int fp64, int pow2, size_t lds, int threads,
                         const void* in, int in_image, void* out, int out_image,
                         const int* done, const int* all_done, hipStream_t s) {
  if (t.n_blocks == 0) return 0;
  dim3 g((unsigned)t.n_blocks), b(threads);
  if (fp64) {
    if (pow2) hipLaunchKernelGGL((fft_batch_kernel<double, true>), g, b, lds, s, probs, t, axis, kind, in, in_image, out, out_image, done, all_done);
    else hipLaunchKernelGGL((fft_batch_kernel<double, false>), g, b, lds, s, probs, t, axis, kind, in, in_image, out, out_image, done, all_done);
  } else {
    if (pow2) hipLaunchKernelGGL((fft_batch_kernel<float, true>), g, b, lds, s, probs, t, axis, kind, in, in_image, out, out_image, done, all_done);
    else hipLaunchKernelGGL((fft_batch_kernel<float, false>), g, b, lds, s, probs, t, axis, kind, in, in_image, out, out_image, done, all_done);
  }
  PNH_LAUNCH_CHECK();
  return 0;
}
