// Interpolation (k2y) and deterministic gridding (y2k) for the min-max NUFFT.
//
// Rows are the samples sorted by the gridding tile that holds their first
// support cell (row r -> sample perm[r]).  The coefficient of row r at
// support offset (j0, j1, j2) is the product of the per-dimension min-max
// factors u0[j0] u1[j1] u2[j2] (the reference's Khatri-Rao construction),
// either formed on the fly from the factors or read from a stored complex64
// table (CSR format: rows interleaved in slices of S = 64 / lanes-per-sample
// so that a wavefront reads consecutive entries).  The support cell along
// dim d is (base_d + j_d) mod K_d.
//
// Both directions work tile by tile over the same work list.  k2y copies a
// tile's padded box of the grid into LDS once and gathers every row of the
// tile from it.  y2k grids every row into an LDS accumulator padded to hold
// whole support boxes, with exact 64-bit fixed-point integer atomics
// (order-independent, hence bit-reproducible); a gather pass then sums the
// padded partial tiles that cover each grid cell in a fixed order.  No float
// atomics touch global memory.
#include <string.h>

#include <type_traits>

#include "common.h"
#include "interp.h"

namespace {

__device__ __forceinline__ int wrap_add(int b, int j, int K) {
  int c = b + j;
  return c >= K ? c - K : c;
}

template <typename C>
__device__ __forceinline__ C shfl_down_c(C v, int off) {
  C z;
  z.x = __shfl_down(v.x, off, 64);
  z.y = __shfl_down(v.y, off, 64);
  return z;
}

__device__ __forceinline__ int rl_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float rl_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rl_f(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <typename C>
__device__ __forceinline__ C rl_c(C v, int l) {
  C z; z.x = rl_f(v.x, l); z.y = rl_f(v.y, l); return z;
}

// Offset of CSR entry (row r, flat support index j = g*JL + jl, g the index
// over all but the last axis) in the slice-interleaved table: slices of SPW
// rows, then g, then row-in-slice, then jl.
__device__ __forceinline__ int64_t csr_index(int64_t r, int g, int jl, int G,
                                             int JL, int SPW) {
  const int64_t slice = r / SPW;
  const int rl = (int)(r - slice * SPW);
  return ((slice * G + g) * SPW + rl) * JL + jl;
}

// ------------------------------------------------------------- dy/dom ----
// Same lane layout as k2y.  For every sample, the D partial derivatives of
// y with respect to its coordinates, d_d y = sum_j (prod_{a != d} u_a)
// du_d * grid, contracted with the incoming gradient g as
// out[m, d] = sum_c Re(conj(g[m, c]) d_d y[m, c]).
template <typename T, int D, int LPS, int NBC>
__global__ void __launch_bounds__(256)
k2y_dom_kernel(InterpArgs a, const void* __restrict__ grid_,
               const void* __restrict__ g_, T* __restrict__ out, int c0,
               int accumulate) {
  typedef typename cplx<T>::t C;
  constexpr int SPW = 64 / LPS;
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int sl = lane / LPS;
  const int jl = lane - sl * LPS;
  const int JL = a.J[D - 1];
  const int64_t r = wave * SPW + sl;
  const bool valid = sl < SPW && jl < JL && r < a.M;
  const C* __restrict__ grid = (const C*)grid_;
  const int nb = a.nb;
  C acc[D][NBC];
#pragma unroll
  for (int d = 0; d < D; ++d)
#pragma unroll
    for (int c = 0; c < NBC; ++c) acc[d][c] = mk<T>((T)0, (T)0);
  if (valid) {
    const C* __restrict__ u = (const C*)a.u;
    const C* __restrict__ du = (const C*)a.du;
    const int J0 = a.J[0];
    const int J1 = D > 1 ? a.J[1] : 1;
    const int64_t ol = a.u_off[D - 1] + r * JL + jl;
    const C ul = u[ol], dul = du[ol];
    const int64_t cl = wrap_add(a.base[(D - 1) * a.M + r], jl, a.K[D - 1]);
    if constexpr (D == 1) {
#pragma unroll
      for (int c = 0; c < NBC; ++c) acc[0][c] = cmul(dul, grid[cl * nb + c0 + c]);
    } else {
      const int b0 = a.base[r];
      const int64_t o0 = a.u_off[0] + r * J0;
      if constexpr (D == 2) {
        for (int j0 = 0; j0 < J0; ++j0) {
          const int64_t cell = (int64_t)wrap_add(b0, j0, a.K[0]) * a.Kstride[0] + cl;
          const C w0 = cmul(du[o0 + j0], ul);
          const C w1 = cmul(u[o0 + j0], dul);
#pragma unroll
          for (int c = 0; c < NBC; ++c) {
            const C gv = grid[cell * nb + c0 + c];
            acc[0][c] = cadd(acc[0][c], cmul(w0, gv));
            acc[D - 1][c] = cadd(acc[D - 1][c], cmul(w1, gv));
          }
        }
      } else {
        const int b1 = a.base[a.M + r];
        const int64_t o1 = a.u_off[1] + r * J1;
        for (int j0 = 0; j0 < J0; ++j0) {
          const int64_t i0 = (int64_t)wrap_add(b0, j0, a.K[0]) * a.Kstride[0] + cl;
          const C u0 = u[o0 + j0], du0 = du[o0 + j0];
          for (int j1 = 0; j1 < J1; ++j1) {
            const int64_t cell = i0 + (int64_t)wrap_add(b1, j1, a.K[1]) * a.Kstride[1];
            const C u1 = u[o1 + j1], du1 = du[o1 + j1];
            const C w0 = cmul(cmul(du0, u1), ul);
            const C w1 = cmul(cmul(u0, du1), ul);
            const C w2 = cmul(cmul(u0, u1), dul);
#pragma unroll
            for (int c = 0; c < NBC; ++c) {
              const C gv = grid[cell * nb + c0 + c];
              acc[0][c] = cadd(acc[0][c], cmul(w0, gv));
              acc[1][c] = cadd(acc[1][c], cmul(w1, gv));
              acc[D - 1][c] = cadd(acc[D - 1][c], cmul(w2, gv));
            }
          }
        }
      }
    }
  }
#pragma unroll
  for (int off = 1; off < LPS; off <<= 1) {
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
      for (int c = 0; c < NBC; ++c) {
        const C o = shfl_down_c(acc[d][c], off);
        if (jl + off < LPS) acc[d][c] = cadd(acc[d][c], o);
      }
  }
  if (valid && jl == 0) {
    const int64_t m = a.perm[r];
    const C* __restrict__ g = (const C*)g_;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      T s = 0;
#pragma unroll
      for (int c = 0; c < NBC; ++c) {
        const C gv = g[m * nb + c0 + c];
        s += gv.x * acc[d][c].x + gv.y * acc[d][c].y;
      }
      if (accumulate) out[m * D + d] += s;
      else out[m * D + d] = s;
    }
  }
}

// ---------------------------------------------------------------- y2k ----
// Gridding by tile ownership.  A row belongs to the tile that holds its
// first support cell; rows are sorted by that tile, so each tile owns one
// contiguous run of rows, cut into work items of at most kMaxItemRows.  One
// workgroup grids one item into an LDS accumulator that covers the tile
// plus J_d - 1 cells on the high side of every axis split into several
// tiles (an axis spanned by one tile wraps inside the accumulator instead),
// so every support box lies wholly inside and nothing is clipped.  Lanes
// work on different samples at once: LPS lanes per sample, one per
// last-axis offset, each looping over the other axes.
//
// Contributions are summed in 64-bit fixed point.  With 2^e bounding
// |Re| and |Im| of every contribution of the item, each is rounded to a
// multiple of 2^(e - 50) and added with an integer LDS atomic.  Integer
// addition is exact, so the accumulator does not depend on the order in
// which lanes and waves arrive and the result is bit-reproducible, while
// the rounding (2^-50 of the item's largest contribution) is far below
// complex64 resolution.  An item has at most 4096 rows, so a cell sums at
// most 2^12 terms below 2^50 and cannot overflow.
//
// Packed accumulators (complex64 operators of plans with per-item bounds,
// BinArgs::item_bound).  When 2^e bounds max(|Re y| + |Im y|) over the
// item times the item's largest per-cell sum of |coefficient| (a plan-time
// constant of the trajectory), every cell's real and imaginary sums stay
// below 2^30 in units of 2^(e - 30).  Each contribution is then one 64-bit
// integer re + im * 2^32 (two's complement, modulo 2^64), added with one
// ds_add_u64: the low 32 bits of the total are the real sum exactly and the
// rest, after removing it, the imaginary sum.  Half the atomics and half the
// LDS of two separate accumulators (two 16^3 workgroups per CU in 3-D); the
// rounding, 2^-30 of that bound, stays below complex64 resolution.
//
// The finished accumulator is written, scaled back, to the item's partial
// slot.  The gather pass then forms every grid cell as the sum of the slots
// that cover it -- its own tile's and those of the up to 2^D - 1 lower
// neighbours whose padding reaches it -- in a fixed order.  No float
// atomics touch global memory.

constexpr int kMaxItemRows = 4096;
constexpr int kFixBits = 50;
constexpr int kPackBits = 30;

// Adding 1.5 * 2^52 to a double of magnitude below 2^51 rounds it to the
// nearest integer and leaves that integer in the low mantissa bits; the low
// word of the constant's bits is zero, so subtracting them is one 32-bit
// subtract of the high word.
constexpr double kMagic = 6755399441055744.0;
constexpr long long kMagicBits = 0x4338000000000000LL;

__device__ __forceinline__ void lds_add(long long* p, long long v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ double2 to_d2(double2 v) { return v; }

// a * b in FP64 (fused multiply-adds)
__device__ __forceinline__ double2 dmul(double2 a, double2 b) {
  return make_double2(__builtin_fma(a.x, b.x, -(a.y * b.y)), __builtin_fma(a.x, b.y, a.y * b.x));
}

// Fixed-point contribution u * a, with a already scaled by 2^(50 - e): each
// part is two fused multiply-adds onto 1.5 * 2^52 (the inner one rounds its
// product to an integer, the outer adds the other product and rounds again,
// so the result is within one unit of the exact product), real part at
// cell[0], imaginary part at cell[1] (one instruction offset apart).
// |u * a| < 2^50 by the choice of e.
__device__ __forceinline__ void lds_add_prod(long long* cell, double2 u, double2 a) {
  const double re = __builtin_fma(u.x, a.x, __builtin_fma(-u.y, a.y, kMagic));
  const double im = __builtin_fma(u.x, a.y, __builtin_fma(u.y, a.x, kMagic));
  lds_add(cell, __double_as_longlong(re) - kMagicBits);
  lds_add(cell + 1, __double_as_longlong(im) - kMagicBits);
}

// Fixed-point contribution of an already scaled value v.
__device__ __forceinline__ void lds_add_fixed(long long* cell, double2 v) {
  lds_add(cell, __double_as_longlong(v.x + kMagic) - kMagicBits);
  lds_add(cell + 1, __double_as_longlong(v.y + kMagic) - kMagicBits);
}

// Packed accumulator word of rounded parts re, im (each already offset by
// 1.5 * 2^52, magnitude below 2^31): re + im * 2^32 modulo 2^64.  re's bits
// minus the constant's are re as a 64-bit integer (sign included); im
// enters through its low word only.  The constant's low word is zero, so the
// low word is re's and the high word is one three-operand 32-bit add (written
// out: the compiler otherwise forms a 64-bit add and a second 32-bit add).
__device__ __forceinline__ long long pack_fixed(double re, double im) {
  const unsigned long long rb = (unsigned long long)__double_as_longlong(re);
  const unsigned neg_magic_hi = 0u - (unsigned)((unsigned long long)kMagicBits >> 32);
  unsigned hi;
  asm("v_add3_u32 %0, %1, %2, %3"
      : "=v"(hi)
      : "v"((unsigned)(rb >> 32)), "v"((unsigned)__double_as_longlong(im)), "s"(neg_magic_hi));
  return (long long)(((unsigned long long)hi << 32) | (rb & 0xffffffffull));
}

// Packed contribution u * a (a scaled by 2^(30 - e)), rounding as lds_add_prod.
__device__ __forceinline__ void lds_add_prod_packed(long long* cell, double2 u, double2 a) {
  const double re = __builtin_fma(u.x, a.x, __builtin_fma(-u.y, a.y, kMagic));
  const double im = __builtin_fma(u.x, a.y, __builtin_fma(u.y, a.x, kMagic));
  lds_add(cell, pack_fixed(re, im));
}

__device__ __forceinline__ void lds_add_fixed_packed(long long* cell, double2 v) {
  lds_add(cell, pack_fixed(v.x + kMagic, v.y + kMagic));
}

// Real and imaginary sums of a packed accumulator word.
__device__ __forceinline__ void unpack_fixed(long long v, long long& re, long long& im) {
  re = (long long)(int)(unsigned)(unsigned long long)v;
  im = (long long)((unsigned long long)v - (unsigned long long)re) >> 32;
}

// Complex product with contraction allowed (fused multiply-adds).
template <typename C>
__device__ __forceinline__ C cmulf(C a, C b) {
  C z;
  z.x = fma(a.x, b.x, -(a.y * b.y));
  z.y = fma(a.x, b.y, a.y * b.x);
  return z;
}

// acc + a * b with contraction allowed
template <typename C>
__device__ __forceinline__ C cfma(C a, C b, C acc) {
  C z;
  z.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
  z.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
  return z;
}

// Origin of a tile and, per axis, the index at which the accumulator wraps
// (K_d for an axis spanned by one tile, else never).
struct TileFrame {
  int T0[PNH_MAXD];
  int wrap[PNH_MAXD];
};

template <int D>
__device__ __forceinline__ TileFrame tile_frame(const InterpArgs& a, const BinArgs& b, int tid) {
  TileFrame f;
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    const int tg = b.tgrid[d];
    const int td = tid % tg;
    tid /= tg;
    f.T0[d] = td * b.tile[d];
    f.wrap[d] = tg == 1 ? a.K[d] : 0x40000000;
  }
  return f;
}

__device__ __forceinline__ int wrap_local(int x, int wrap, int K) {
  return x >= wrap ? x - K : x;
}

// Local coordinate in the padded box: FP states that every axis is split
// (the box never wraps), so the coordinate is x itself.
template <bool FP>
__device__ __forceinline__ int box_local(int x, int wrap, int K) {
  return FP ? x : wrap_local(x, wrap, K);
}

// The interpolation kernel keeps run-time box extents: its variant with
// PNH_FIXED_PAD needed 66 registers, above the 64 that let two 1 024-thread
// workgroups share a CU, and ran 43 % slower (the gridding kernel's variant:
// 3-D 128^3 tile kernel 240 -> 215 us).

// Walks the cells i = first, first + step, ... < n of a padded box of
// extents (P0,) P1, P2 in C order, calling f(i, x0, x1, x2): the coordinates
// advance by the step's own decomposition with carries (each stays below
// twice its extent, so one subtraction suffices), with no division per cell.
template <int D, typename F>
__device__ __forceinline__ void for_box_cells(int n, int P1, int P2, int first, int step, F&& f) {
  int x2 = 0, x1 = 0, x0 = first;
  if (D > 2) { x2 = x0 % P2; x0 /= P2; }
  if (D > 1) { x1 = x0 % P1; x0 /= P1; }
  int s2 = 0, s1 = 0, s0 = step;
  if (D > 2) { s2 = s0 % P2; s0 /= P2; }
  if (D > 1) { s1 = s0 % P1; s0 /= P1; }
  for (int i = first; i < n; i += step) {
    f(i, x0, x1, x2);
    if (D > 2) {
      x2 += s2;
      const int c = x2 >= P2;
      x2 -= c ? P2 : 0;
      x1 += s1 + c;
    } else if (D > 1) {
      x1 += s1;
    }
    if (D > 1) {
      const int c = x1 >= P1;
      x1 -= c ? P1 : 0;
      x0 += s0 + c;
    } else {
      x0 += s0;
    }
  }
}

// Max over the workgroup; `red` holds one slot per wave.
template <typename T>
__device__ __forceinline__ T block_max(T v, T* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) red[wave] = v;
  __syncthreads();
  T r = red[0];
  for (int w = 1; w < nw; ++w) r = fmax(r, red[w]);
  return r;
}

// One work item (tile id, first row, rows, slot).  LDS: the accumulator as
// int64, [pad_cells][NBC][real, imaginary], then
// reduction scratch.  JN bounds the
// unrolled support loops of the leading axes; JE > 0 states that they are
// exactly JE long (no per-offset guards).
// Fixed-point exponent of one work item: e with 2^e > (|Re y| + |Im y|) *
// bound over the item's rows and coils c0 .. c0+NBC-1, or kPoisonExp when
// one of those samples is not finite.  bound: u_bound (two accumulators per
// cell) or the item's bound (packed).  Block-wide (every thread calls it);
// red holds one double per wave.
constexpr int kPoisonExp = 0x7fffffff;

// Whether the complex64 gridding of these work items packs its accumulators.
template <typename T>
__host__ __device__ inline bool packed_items(const BinArgs& b) {
  return sizeof(T) == 4 && b.item_bound != nullptr;
}

template <typename T>
__device__ __forceinline__ double exponent_bound(const BinArgs& b, int item) {
  return packed_items<T>(b) ? b.item_bound[item] : b.u_bound;
}

// ys: when non-null, also stores the item's samples in row order there
// (ys[r * NBC + c]).  YR: y_ itself holds the samples in row order (then c0
// and the coil count are not used).
template <typename T, int NBC, bool YR = false>
__device__ __forceinline__ int item_exponent(const InterpArgs& a, const BinArgs& b,
                                             int r_begin, int r_end, const void* __restrict__ y_,
                                             int c0, double bound, double* red,
                                             void* __restrict__ ys_ = nullptr) {
  typedef typename cplx<T>::t C;
  const C* __restrict__ y = (const C*)y_;
  C* __restrict__ ys = (C*)ys_;
  double ymax = 0.0;
  int bad = 0;
  for (int r = r_begin + (int)threadIdx.x; r < r_end; r += blockDim.x) {
    const int64_t m = YR ? 0 : a.perm[r];
#pragma unroll
    for (int c = 0; c < NBC; ++c) {
      const C v = YR ? y[(int64_t)r * NBC + c] : y[m * a.nb + c0 + c];
      if (ys) ys[(int64_t)r * NBC + c] = v;
      const double s = fabs((double)v.x) + fabs((double)v.y);
      if (s <= 1.0e300) ymax = fmax(ymax, s);
      else bad = 1;                       // NaN or infinity
    }
  }
  if (bad) ymax = __builtin_inf();
  ymax = block_max(ymax, red);
  if (!(ymax <= 1.0e300)) return kPoisonExp;
  const double B = ymax * bound;
  int e = 0;
  if (B > 0.0) (void)frexp(B, &e);       // B < 2^e
  return e < -960 ? -960 : e;
}

// Exponents of all work items for coils c0 .. c0+NBC-1, one workgroup per
// item, ahead of the gridding kernel (so its workgroups start adding at once);
// the samples, read through the row permutation anyway, are stored in row
// order (ys) for the gridding kernel to read contiguously.
template <typename T, int NBC>
__global__ void __launch_bounds__(256)
y2k_exponent_kernel(InterpArgs a, BinArgs b, const void* __restrict__ y_, int c0,
                    int* __restrict__ item_e, const int* __restrict__ skip,
                    void* __restrict__ ys) {
  if (skip && *skip) return;
  __shared__ double red[16];
  const int4 it = reinterpret_cast<const int4*>(b.items)[blockIdx.x];
  const int e = item_exponent<T, NBC>(a, b, it.y, it.y + it.z, y_, c0,
                                      exponent_bound<T>(b, blockIdx.x), red, ys);
  if (threadIdx.x == 0) item_e[blockIdx.x] = e;
}

// e_pre: the item's exponent computed ahead (y2k_exponent_kernel), or null
// to compute it here.  item: the item's index in the plan's work list.  PK:
// packed accumulators (one 64-bit word per cell and coil, else two).  ys_:
// the samples in row order (ys_[r * NBC + c]) or null to read y through the
// row permutation; with e_pre null, the exponent is taken from ys_ too.
template <typename T, int D, bool CSRV, int NBC, int JN, int JE, bool PK, int PADC = 0>
__device__ __forceinline__ void y2k_item(const InterpArgs& a, const BinArgs& b, const int4 it,
                                         int item, long long* __restrict__ acc,
                                         const void* __restrict__ y_,
                                         void* __restrict__ part_, int c0,
                                         const int* __restrict__ e_pre,
                                         const void* __restrict__ ys_ = nullptr) {
  typedef typename cplx<T>::t C;
  constexpr int CW = PK ? 1 : 2;         // accumulator words per (cell, coil)
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int pc = (int)b.pad_cells;
  const int plane = pc * NBC;            // (cell, coil) pairs
  double* red = reinterpret_cast<double*>(acc + (int64_t)plane * CW);
  const C* __restrict__ y = (const C*)y_;
  const int nb = a.nb;
  const int r_begin = it.y, r_end = it.y + it.z;

  if (PK) {
    for (int i = tid; i < plane; i += nthr) acc[i] = 0;
  } else {
    longlong2* acc2 = reinterpret_cast<longlong2*>(acc);
    for (int i = tid; i < plane; i += nthr) acc2[i] = make_longlong2(0, 0);
  }
  // scale: 2^e bounds (|Re y| + |Im y|) * |coef| for every row of the item
  int e;
  if (e_pre) {
    e = *e_pre;
    __syncthreads();                      // zeroing before the adds
  } else {
    const double bound = PK ? b.item_bound[item] : b.u_bound;
    e = ys_ ? item_exponent<T, NBC, true>(a, b, r_begin, r_end, ys_, c0, bound, red)
            : item_exponent<T, NBC>(a, b, r_begin, r_end, y_, c0, bound, red);  // also a barrier
  }
  // a non-finite sample poisons the item's whole slot
  const bool poison = e == kPoisonExp;
  if (poison) e = 0;
  constexpr int FB = PK ? kPackBits : kFixBits;
  const double scale = ldexp(1.0, FB - e);
  const double unscale = ldexp(1.0, e - FB);

  const int JL = a.J[D - 1];
  const int LPS = pnh_lanes_per_sample(JL);
  const int SPW = 64 / LPS;
  const int lane = tid & 63, wave = tid >> 6, nw = nthr >> 6;
  const int sl = lane / LPS;
  const int jl = lane - sl * LPS;
  constexpr bool FP = PADC > 0;          // every axis split, extent PADC
  const int P1 = D > 1 ? (FP ? PADC : b.pad[1]) : 1;
  const int P2 = D > 2 ? (FP ? PADC : b.pad[2]) : 1;
  const TileFrame f = tile_frame<D>(a, b, it.x);
  const C* __restrict__ u = (const C*)a.u;
  const int J0 = JE > 0 ? JE : a.J[0];
  const int J1 = D > 2 ? (JE > 0 ? JE : a.J[1]) : 1;

  for (int r0 = r_begin + wave * SPW; !poison && r0 < r_end; r0 += nw * SPW) {
    const int r = r0 + sl;
    if (sl >= SPW || jl >= JL || r >= r_end) continue;
    C yv[NBC];
    if (ys_) {
      // row-ordered samples of the exponent pass (contiguous)
#pragma unroll
      for (int c = 0; c < NBC; ++c) yv[c] = ((const C*)ys_)[(int64_t)r * NBC + c];
    } else {
      const int64_t m = a.perm[r];
#pragma unroll
      for (int c = 0; c < NBC; ++c) yv[c] = y[m * nb + c0 + c];
    }
    // this lane's last-axis accumulator index
    const int xl = box_local<FP>(a.base[(int64_t)(D - 1) * a.M + r] - f.T0[D - 1] + jl,
                              f.wrap[D - 1], a.K[D - 1]);
    // The row's samples times the item's scale, in FP64 (exact: a float
    // times a power of two); every product below is formed in FP64 and
    // rounded once, onto the fixed-point grid.
    double2 ys[NBC];
#pragma unroll
    for (int c = 0; c < NBC; ++c) {
      ys[c] = to_d2(yv[c]);
      ys[c].x *= scale;
      ys[c].y *= scale;
    }
    if (CSRV) {
      const int G = D == 1 ? 1 : (D == 2 ? J0 : J0 * J1);
      const C* __restrict__ val = (const C*)a.csr_val;
      const int64_t i0 = csr_index(r, 0, jl, G, JL, SPW);
      const int64_t gs = (int64_t)SPW * JL;
      const int x0b = D > 1 ? a.base[r] - f.T0[0] : 0;
      const int x1b = D > 2 ? a.base[a.M + r] - f.T0[1] : 0;
      int j0 = 0, j1 = 0;
      for (int g = 0; g < G; ++g) {
        int idx = xl;
        if (D == 2) idx += box_local<FP>(x0b + g, f.wrap[0], a.K[0]) * P1;
        if (D == 3)
          idx += (box_local<FP>(x0b + j0, f.wrap[0], a.K[0]) * P1 +
                  box_local<FP>(x1b + j1, f.wrap[1], a.K[1])) * P2;
        const double2 w = to_d2(cconj(val[i0 + g * gs]));
        long long* cell = acc + CW * idx * NBC;
#pragma unroll
        for (int c = 0; c < NBC; ++c) {
          if (PK) lds_add_prod_packed(cell + c, w, ys[c]);
          else lds_add_prod(cell + 2 * c, w, ys[c]);
        }
        if (D == 3 && ++j1 == J1) { j1 = 0; ++j0; }
      }
      continue;
    }
    // last-axis factor times the scaled samples
    const double2 ul = to_d2(cconj(u[a.u_off[D - 1] + (int64_t)r * JL + jl]));
    double2 wl[NBC];
#pragma unroll
    for (int c = 0; c < NBC; ++c) wl[c] = dmul(ul, ys[c]);
    if (D == 1) {
      long long* cell = acc + CW * xl * NBC;
#pragma unroll
      for (int c = 0; c < NBC; ++c) {
        if (PK) lds_add_fixed_packed(cell + c, wl[c]);
        else lds_add_fixed(cell + 2 * c, wl[c]);
      }
      continue;
    }
    const int x0b = a.base[r] - f.T0[0];
    const C* __restrict__ u0 = u + a.u_off[0] + (int64_t)r * J0;
    // accumulator offsets in bytes (a cell holds NBC coils of CW words)
    constexpr int CB = 8 * CW * NBC;
    char* const accb = reinterpret_cast<char*>(acc);
    double2 u0c[JN];
    int row0[JN];
#pragma unroll
    for (int j = 0; j < JN; ++j)
      if (JE > 0 || j < J0) {
        u0c[j] = to_d2(cconj(u0[j]));
        row0[j] = box_local<FP>(x0b + j, f.wrap[0], a.K[0]) * P1 * P2 * CB;
      }
    if (D == 2) {
      const int xlb = xl * CB;
#pragma unroll
      for (int j0 = 0; j0 < JN; ++j0) {
        if (JE > 0 || j0 < J0) {
          long long* cell = reinterpret_cast<long long*>(accb + row0[j0] + xlb);
#pragma unroll
          for (int c = 0; c < NBC; ++c) {
            if (PK) lds_add_prod_packed(cell + c, u0c[j0], wl[c]);
            else lds_add_prod(cell + 2 * c, u0c[j0], wl[c]);
          }
        }
      }
      continue;
    }
    // D == 3: factors of axes 0 and 1 in registers, both loops unrolled
    const int x1b = a.base[a.M + r] - f.T0[1];
    const C* __restrict__ u1 = u + a.u_off[1] + (int64_t)r * J1;
    double2 u1c[JN];
    int off1[JN];
#pragma unroll
    for (int j = 0; j < JN; ++j)
      if (JE > 0 || j < J1) {
        u1c[j] = to_d2(cconj(u1[j]));
        off1[j] = (box_local<FP>(x1b + j, f.wrap[1], a.K[1]) * P2 + xl) * CB;
      }
#pragma unroll
    for (int j0 = 0; j0 < JN; ++j0) {
      if (JE > 0 || j0 < J0) {
        double2 a0[NBC];
#pragma unroll
        for (int c = 0; c < NBC; ++c) a0[c] = dmul(u0c[j0], wl[c]);
#pragma unroll
        for (int j1 = 0; j1 < JN; ++j1) {
          if (JE > 0 || j1 < J1) {
            long long* cell = reinterpret_cast<long long*>(accb + row0[j0] + off1[j1]);
#pragma unroll
            for (int c = 0; c < NBC; ++c) {
              if (PK) lds_add_prod_packed(cell + c, u1c[j1], a0[c]);
              else lds_add_prod(cell + 2 * c, u1c[j1], a0[c]);
            }
          }
        }
      }
    }
  }
  __syncthreads();
  // partial slot, scaled back
  C* __restrict__ part = (C*)part_ + (int64_t)it.w * pc * NBC;
  const T nan = (T)__builtin_nan("");
  for (int i = tid; i < plane; i += nthr) {
    C v;
    if (poison) {
      v = mk<T>(nan, nan);
    } else if (PK) {
      long long re, im;
      unpack_fixed(acc[i], re, im);
      v = mk<T>((T)((double)re * unscale), (T)((double)im * unscale));
    } else {
      v = mk<T>((T)((double)acc[2 * i] * unscale), (T)((double)acc[2 * i + 1] * unscale));
    }
    part[i] = v;
  }
}

template <typename T, int D, bool CSRV, int NBC, int JN, int JE, bool PK, int PADC>
__global__ void __launch_bounds__(1024)
y2k_tile_kernel(InterpArgs a, BinArgs b, const void* __restrict__ y_,
                void* __restrict__ part_, int c0, const int* __restrict__ item_e,
                const int* __restrict__ skip, const void* __restrict__ ys_) {
  if (skip && *skip) return;
  extern __shared__ long long smem_acc[];
  const int4 it = reinterpret_cast<const int4*>(b.items)[blockIdx.x];
  y2k_item<T, D, CSRV, NBC, JN, JE, PK, PADC>(a, b, it, blockIdx.x, smem_acc, y_, part_, c0,
                                              item_e ? item_e + blockIdx.x : nullptr, ys_);
}

// ------------------------------------------------------------ k2y (tiles) --
// Interpolation by the same tile ownership as y2k: a workgroup copies the
// padded box of its tile (tile + J - 1 cells on split axes, wrapping across
// the grid edge; the whole axis on spanning axes) from the grid into LDS
// once, then every row of the work item gathers its support from LDS.  LPS
// lanes per sample, one per last-axis offset, loop over the other axes with
// the factors in registers; a segmented shuffle sums the lanes of a sample.
// Nothing is accumulated across rows, so the result does not depend on
// scheduling.  LDS: [pad_cells][NBC] of the operator's complex type.
// NW: every axis of the work list is split, so the padded box never wraps
// (no wrap tests on the box coordinates).
template <typename T, int D, bool CSRV, int NBC, int JN, int JE, bool NW = false>
__device__ __forceinline__ void k2y_item(const InterpArgs& a, const BinArgs& b, const int4 it,
                                         typename cplx<T>::t* __restrict__ box,
                                         const void* __restrict__ grid_,
                                         void* __restrict__ y_, int c0,
                                         bool y_rows = false) {
  typedef typename cplx<T>::t C;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int nb = a.nb;
  const TileFrame f = tile_frame<D>(a, b, it.x);
  const int P1 = D > 1 ? b.pad[1] : 1;
  const int P2 = D > 2 ? b.pad[2] : 1;
  const C* __restrict__ grid = (const C*)grid_;
  // padded box -> LDS (last axis fastest for coalescing)
  for_box_cells<D>((int)b.pad_cells, P1, P2, tid, nthr, [&](int i, int x0, int x1, int x2) {
    int c = f.T0[0] + x0;
    int g = (c >= a.K[0] ? c - a.K[0] : c) * (int)a.Kstride[0];
    if (D > 1) {
      c = f.T0[1] + x1;
      g += (c >= a.K[1] ? c - a.K[1] : c) * (int)a.Kstride[1];
    }
    if (D > 2) {
      c = f.T0[2] + x2;
      g += c >= a.K[2] ? c - a.K[2] : c;
    }
#pragma unroll
    for (int cc = 0; cc < NBC; ++cc) box[i * NBC + cc] = grid[(int64_t)g * nb + c0 + cc];
  });
  __syncthreads();

  const int JL = a.J[D - 1];
  const int LPS = pnh_lanes_per_sample(JL);
  const int SPW = 64 / LPS;
  const int lane = tid & 63, wave = tid >> 6, nw = nthr >> 6;
  const int sl = lane / LPS;
  const int jl = lane - sl * LPS;
  const C* __restrict__ u = (const C*)a.u;
  const int J0 = JE > 0 ? JE : a.J[0];
  const int J1 = D > 2 ? (JE > 0 ? JE : a.J[1]) : 1;
  const int r_end = it.y + it.z;
  C* __restrict__ y = (C*)y_;

  for (int r0 = it.y + wave * SPW; r0 < r_end; r0 += nw * SPW) {
    const int r = r0 + sl;
    const bool valid = sl < SPW && jl < JL && r < r_end;
    C acc[NBC];
#pragma unroll
    for (int c = 0; c < NBC; ++c) acc[c] = mk<T>((T)0, (T)0);
    if (valid) {
      const int xl = box_local<NW>(a.base[(int64_t)(D - 1) * a.M + r] - f.T0[D - 1] + jl,
                                f.wrap[D - 1], a.K[D - 1]);
      if (CSRV) {
        const int G = D == 1 ? 1 : (D == 2 ? J0 : J0 * J1);
        const C* __restrict__ val = (const C*)a.csr_val;
        const int64_t i0 = csr_index(r, 0, jl, G, JL, SPW);
        const int64_t gs = (int64_t)SPW * JL;
        const int x0b = D > 1 ? a.base[r] - f.T0[0] : 0;
        const int x1b = D > 2 ? a.base[a.M + r] - f.T0[1] : 0;
        int j0 = 0, j1 = 0;
        for (int g = 0; g < G; ++g) {
          int idx = xl;
          if (D == 2) idx += box_local<NW>(x0b + g, f.wrap[0], a.K[0]) * P1;
          if (D == 3)
            idx += (box_local<NW>(x0b + j0, f.wrap[0], a.K[0]) * P1 +
                    box_local<NW>(x1b + j1, f.wrap[1], a.K[1])) * P2;
          const C w = val[i0 + g * gs];
#pragma unroll
          for (int c = 0; c < NBC; ++c) acc[c] = cfma(w, box[idx * NBC + c], acc[c]);
          if (D == 3 && ++j1 == J1) { j1 = 0; ++j0; }
        }
      } else {
        const C ul = u[a.u_off[D - 1] + (int64_t)r * JL + jl];
        if (D == 1) {
#pragma unroll
          for (int c = 0; c < NBC; ++c) acc[c] = cmulf(ul, box[xl * NBC + c]);
        } else {
          const int x0b = a.base[r] - f.T0[0];
          const C* __restrict__ u0 = u + a.u_off[0] + (int64_t)r * J0;
          C u0c[JN];
          int row0[JN];
#pragma unroll
          for (int j = 0; j < JN; ++j)
            if (JE > 0 || j < J0) {
              u0c[j] = u0[j];
              row0[j] = box_local<NW>(x0b + j, f.wrap[0], a.K[0]) * P1 * P2;
            }
          if (D == 2) {
            C s0[NBC];
#pragma unroll
            for (int c = 0; c < NBC; ++c) s0[c] = mk<T>((T)0, (T)0);
#pragma unroll
            for (int j0 = 0; j0 < JN; ++j0) {
              if (JE > 0 || j0 < J0) {
                const int idx = row0[j0] + xl;
#pragma unroll
                for (int c = 0; c < NBC; ++c) s0[c] = cfma(u0c[j0], box[idx * NBC + c], s0[c]);
              }
            }
#pragma unroll
            for (int c = 0; c < NBC; ++c) acc[c] = cmulf(ul, s0[c]);
          } else {
            const int x1b = a.base[a.M + r] - f.T0[1];
            const C* __restrict__ u1 = u + a.u_off[1] + (int64_t)r * J1;
            C u1c[JN];
            int off1[JN];
#pragma unroll
            for (int j = 0; j < JN; ++j)
              if (JE > 0 || j < J1) {
                u1c[j] = u1[j];
                off1[j] = box_local<NW>(x1b + j, f.wrap[1], a.K[1]) * P2 + xl;
              }
            // acc = ul * sum_j0 u0[j0] * (sum_j1 u1[j1] * box[j0, j1]): one
            // complex multiply-add per support cell
            C s0[NBC];
#pragma unroll
            for (int c = 0; c < NBC; ++c) s0[c] = mk<T>((T)0, (T)0);
#pragma unroll
            for (int j0 = 0; j0 < JN; ++j0) {
              if (JE > 0 || j0 < J0) {
                C s1[NBC];
#pragma unroll
                for (int c = 0; c < NBC; ++c) s1[c] = mk<T>((T)0, (T)0);
#pragma unroll
                for (int j1 = 0; j1 < JN; ++j1) {
                  if (JE > 0 || j1 < J1) {
                    const int idx = row0[j0] + off1[j1];
#pragma unroll
                    for (int c = 0; c < NBC; ++c) s1[c] = cfma(u1c[j1], box[idx * NBC + c], s1[c]);
                  }
                }
#pragma unroll
                for (int c = 0; c < NBC; ++c) s0[c] = cfma(u0c[j0], s1[c], s0[c]);
              }
            }
#pragma unroll
            for (int c = 0; c < NBC; ++c) acc[c] = cmulf(ul, s0[c]);
          }
        }
      }
    }
    // segmented sum over the LPS lanes of each sample (all lanes take part)
    for (int off = 1; off < LPS; off <<= 1) {
#pragma unroll
      for (int c = 0; c < NBC; ++c) {
        const C o = shfl_down_c(acc[c], off);
        if (jl + off < LPS) acc[c] = cadd(acc[c], o);
      }
    }
    if (valid && jl == 0) {
      if (y_rows) {
#pragma unroll
        for (int c = 0; c < NBC; ++c) y[(int64_t)r * NBC + c] = acc[c];
      } else {
        const int64_t m = a.perm[r];
#pragma unroll
        for (int c = 0; c < NBC; ++c) y[m * nb + c0 + c] = acc[c];
      }
    }
  }
}

template <typename T, int D, bool CSRV, int NBC, int JN, int JE, bool NW>
__global__ void __launch_bounds__(1024)
k2y_tile_kernel(InterpArgs a, BinArgs b, const void* __restrict__ grid_,
                void* __restrict__ y_, int c0, const int* __restrict__ skip) {
  typedef typename cplx<T>::t C;
  if (skip && *skip) return;
  extern __shared__ long long smem_box[];
  const int4 it = reinterpret_cast<const int4*>(b.items)[blockIdx.x];
  k2y_item<T, D, CSRV, NBC, JN, JE, NW>(a, b, it, reinterpret_cast<C*>(smem_box), grid_, y_,
                                        c0);
}

// Grid cells from the partial slots that cover them, one workgroup per
// tile.  A cell of tile t sums, in a fixed order, the slots of the tiles
// whose padded box holds it: neighbour masks 0 .. 2^D - 1 (bit d: the tile
// below along axis d, whose padding reaches the J_d - 1 lowest cells of t),
// and within a mask the tile's slots (several when the tile's rows were split
// into several work items: NQ running sums over the slots, combined in a
// fixed tree).  The tile's own slot is read contiguously and every slot cell
// is read exactly once over the launch.
template <typename T, int D, int NBC>
__device__ __forceinline__ void gather_tile(const InterpArgs& a, const BinArgs& b, int tid,
                                            int first, int count,
                                            const typename cplx<T>::t* __restrict__ part,
                                            typename cplx<T>::t* __restrict__ grid, int c0) {
  typedef typename cplx<T>::t C;
  int td[PNH_MAXD], T0[PNH_MAXD], E[PNH_MAXD], below[PNH_MAXD], tdb[PNH_MAXD];
  {
    int rest = tid;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      const int tg = b.tgrid[d];
      td[d] = rest % tg;
      rest /= tg;
      T0[d] = td[d] * b.tile[d];
      E[d] = min(b.tile[d], a.K[d] - T0[d]);
      tdb[d] = td[d] == 0 ? tg - 1 : td[d] - 1;
      // cells of the tile below reached by its padding (none on a spanning axis)
      below[d] = tg == 1 ? 0 : a.J[d] - 1;
    }
  }
  // slots and accumulator offsets of the 2^D source tiles (uniform)
  const int2* __restrict__ ts = reinterpret_cast<const int2*>(b.tile_slots);
  int2 src[1 << D];
  int off[1 << D][PNH_MAXD];
#pragma unroll
  for (int mask = 0; mask < (1 << D); ++mask) {
    int tix = 0;
    bool ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const bool bit = (mask >> d) & 1;
      ok &= !bit || below[d] > 0;
      const int t = bit ? tdb[d] : td[d];
      tix = tix * b.tgrid[d] + t;
      off[mask][d] = bit ? min(b.tile[d], a.K[d] - t * b.tile[d]) : 0;
    }
    src[mask] = ok ? ts[tix] : make_int2(0, 0);
  }
  const int E1 = D > 1 ? E[1] : 1, E2 = D > 2 ? E[2] : 1;
  const int P1 = D > 1 ? b.pad[1] : 1, P2 = D > 2 ? b.pad[2] : 1;
  const int n = E[0] * E1 * E2;
  const int end = first + count < n ? first + count : n;
  // per source: the padded-box cell offset of the tile's cells in the
  // source's slot (uniform); a cell's slot index is its own padded-box index
  // plus this
  int offc[1 << D];
#pragma unroll
  for (int mask = 0; mask < (1 << D); ++mask) {
    int o = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) o = o * (d == 0 ? 1 : (d == 1 ? P1 : P2)) + off[mask][d];
    offc[mask] = o;
  }
  int64_t g0 = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) g0 += (int64_t)T0[d] * a.Kstride[d];
  // the cell's coordinates in the tile, advanced by blockDim.x cells per
  // step (no division per cell)
  int l[PNH_MAXD];
  {
    int rest = first + (int)threadIdx.x;
    if (D > 2) { l[2] = rest % E2; rest /= E2; }
    if (D > 1) { l[1] = rest % E1; rest /= E1; }
    l[0] = rest;
  }
  int sl[PNH_MAXD];
  {
    int rest = (int)blockDim.x;
    if (D > 2) { sl[2] = rest % E2; rest /= E2; }
    if (D > 1) { sl[1] = rest % E1; rest /= E1; }
    sl[0] = rest;
  }
  for (int i = first + threadIdx.x; i < end; i += blockDim.x) {
    int pc = l[0];
    if (D > 1) pc = pc * P1 + l[1];
    if (D > 2) pc = pc * P2 + l[2];
    C s[NBC];
#pragma unroll
    for (int k = 0; k < NBC; ++k) s[k] = mk<T>((T)0, (T)0);
    if (b.max_tile_slots == 1) {
      // one slot per source at most: the loads issued before any sum, so
      // they overlap; the sum selects the sources that apply, in mask order.
      // 3-D loads only those (on average 2.3 of 8 at 16^3); 1-D and 2-D
      // load every source (the ones that do not apply read the slot array's
      // first cell), which measured faster there
      constexpr bool CL = D == 3;
      C v[1 << D][NBC];
      bool use[1 << D];
#pragma unroll
      for (int mask = 0; mask < (1 << D); ++mask) {
        bool ok = src[mask].y > 0;
#pragma unroll
        for (int d = 0; d < D; ++d)
          if ((mask >> d) & 1) ok &= l[d] < below[d];
        use[mask] = ok;
        const C* p = part + (ok ? ((int64_t)src[mask].x * b.pad_cells + pc + offc[mask]) * NBC : 0);
#pragma unroll
        for (int cc = 0; cc < NBC; ++cc) v[mask][cc] = (!CL || ok) ? p[cc] : mk<T>((T)0, (T)0);
      }
#pragma unroll
      for (int mask = 0; mask < (1 << D); ++mask)
        if (use[mask]) {
#pragma unroll
          for (int cc = 0; cc < NBC; ++cc) s[cc] = cadd(s[cc], v[mask][cc]);
        }
    } else {
      // The first slot of every source is loaded before any sum (one
      // round trip when the sources have one slot each, the common case);
      // a heavy source's further slots follow in rounds of NQ loads into NQ
      // running sums (slots 1 + k with k = j mod NQ, in row order each),
      // combined pairwise in a fixed tree, so the result stays
      // reproducible.  Eight for one coil, four for several (register
      // budget).
      constexpr int NQ = NBC == 1 ? 8 : 4;
      const int64_t stride = b.pad_cells * NBC;
      C v0[1 << D][NBC];
      const C* ps[1 << D];
      int nsl[1 << D];
#pragma unroll
      for (int mask = 0; mask < (1 << D); ++mask) {
        bool ok = src[mask].y > 0;
#pragma unroll
        for (int d = 0; d < D; ++d)
          if ((mask >> d) & 1) ok &= l[d] < below[d];
        nsl[mask] = ok ? src[mask].y : 0;
        ps[mask] = part + (ok ? ((int64_t)src[mask].x * b.pad_cells + pc + offc[mask]) * NBC : 0);
#pragma unroll
        for (int cc = 0; cc < NBC; ++cc)
          v0[mask][cc] = ok ? ps[mask][cc] : mk<T>((T)0, (T)0);
      }
#pragma unroll
      for (int mask = 0; mask < (1 << D); ++mask) {
        const int ns = nsl[mask];
        if (ns == 0) continue;
        C t[NBC];
#pragma unroll
        for (int cc = 0; cc < NBC; ++cc) t[cc] = v0[mask][cc];
        if (ns > 1) {
          const C* p = ps[mask] + stride;
          C q[NQ][NBC];
#pragma unroll
          for (int j = 0; j < NQ; ++j)
#pragma unroll
            for (int cc = 0; cc < NBC; ++cc) q[j][cc] = mk<T>((T)0, (T)0);
          const int nr = ns - 1;
          for (int k = 0; k < nr; k += NQ) {
            // NQ loads in flight, the last round's beyond nr skipped
            C v[NQ][NBC];
#pragma unroll
            for (int j = 0; j < NQ; ++j)
#pragma unroll
              for (int cc = 0; cc < NBC; ++cc)
                v[j][cc] = k + j < nr ? p[(int64_t)(k + j) * stride + cc] : mk<T>((T)0, (T)0);
#pragma unroll
            for (int j = 0; j < NQ; ++j)
#pragma unroll
              for (int cc = 0; cc < NBC; ++cc) q[j][cc] = cadd(q[j][cc], v[j][cc]);
          }
#pragma unroll
          for (int w = NQ / 2; w >= 1; w /= 2)
#pragma unroll
            for (int j = 0; j < w; ++j)
#pragma unroll
              for (int cc = 0; cc < NBC; ++cc) q[j][cc] = cadd(q[2 * j][cc], q[2 * j + 1][cc]);
#pragma unroll
          for (int cc = 0; cc < NBC; ++cc) t[cc] = cadd(t[cc], q[0][cc]);
        }
#pragma unroll
        for (int cc = 0; cc < NBC; ++cc) s[cc] = cadd(s[cc], t[cc]);
      }
    }
    int64_t g = g0 + l[0] * a.Kstride[0];
    if (D > 1) g += l[1] * a.Kstride[1];
    if (D > 2) g += l[2] * a.Kstride[2];
#pragma unroll
    for (int cc = 0; cc < NBC; ++cc) grid[g * a.nb + c0 + cc] = s[cc];
    // next cell: add the step with carries (each coordinate stays below
    // twice its extent, so one subtraction suffices)
    if (D > 2) {
      l[2] += sl[2];
      const int c2 = l[2] >= E2;
      l[2] -= c2 ? E2 : 0;
      l[1] += sl[1] + c2;
    } else if (D > 1) {
      l[1] += sl[1];
    }
    if (D > 1) {
      const int c1 = l[1] >= E1;
      l[1] -= c1 ? E1 : 0;
      l[0] += sl[0] + c1;
    } else {
      l[0] += sl[0];
    }
  }
}

// blockIdx.x = tile * parts + part: one workgroup per tile, or (parts > 1)
// one per 256 consecutive cells of a tile.
template <typename T, int D, int NBC>
__global__ void __launch_bounds__(256)
y2k_gather_kernel(InterpArgs a, BinArgs b, int parts, const void* __restrict__ part_,
                  void* __restrict__ grid_, int c0, const int* __restrict__ skip) {
  typedef typename cplx<T>::t C;
  if (skip && *skip) return;
  const int t = blockIdx.x / parts, q = blockIdx.x - t * parts;
  const int span = parts == 1 ? (1 << 30) : 256;
  gather_tile<T, D, NBC>(a, b, t, q * 256, span, (const C*)part_, (C*)grid_, c0);
}

// accumulator (one or two 64-bit words per cell and coil) + reduction scratch
size_t y2k_lds_bytes(int64_t pad_cells, int nbc, bool packed) {
  return (size_t)pad_cells * nbc * (packed ? 8 : 16) + 256;
}

// LDS of the one-pass FP64 P^H P: the box and the two-word accumulator.
size_t k2y2k_1p_lds_bytes(int64_t pad_cells) { return (size_t)pad_cells * 32 + 256; }

// Calls f(std::integral_constant<bool, PK>) with PK = packed (the packed
// variant exists for complex64 only).
template <typename T, typename F>
int dispatch_pk(bool packed, F&& f) {
  if constexpr (sizeof(T) == 4) {
    if (packed) return f(std::integral_constant<bool, true>());
  }
  return f(std::integral_constant<bool, false>());
}

// Dynamic LDS above 64 KB must be allowed per kernel before the launch.
template <typename K>
void allow_lds(K kernel, size_t bytes) {
  if (bytes > 65536)
    (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes);
}

// Calls f(std::integral_constant<int, PADC>): PNH_FIXED_PAD when every axis
// is split with that padded extent (a support of exactly 6, factor format,
// 2-D or 3-D), else 0 (extents read at run time).
template <int D, bool CSRV, int JE, typename F>
int dispatch_pad(const BinArgs& b, F&& f) {
  if constexpr (!CSRV && JE == 6 && D >= 2) {
    if (pnh_fixed_pad(b, D)) return f(std::integral_constant<int, PNH_FIXED_PAD>());
  }
  return f(std::integral_constant<int, 0>());
}

// Support-loop specialization of the leading axes: exact sizes 4, 6 and 8
// (factor format) unrolled without guards, other sizes up to 8 or 16 with
// guards.  Calls f(JN, JE) with std::integral_constant arguments.
template <int D, bool CSRV, typename F>
int dispatch_j(const InterpArgs& a, F&& f) {
  typedef std::integral_constant<int, 0> E0;
  if (D == 1) return f(std::integral_constant<int, 1>(), E0());
  const int j0 = a.J[0], j1 = D > 2 ? a.J[1] : a.J[0];
  if (!CSRV && j0 == j1) {
    if (j0 == 4) return f(std::integral_constant<int, 4>(), std::integral_constant<int, 4>());
    if (j0 == 6) return f(std::integral_constant<int, 6>(), std::integral_constant<int, 6>());
    if (j0 == 8) return f(std::integral_constant<int, 8>(), std::integral_constant<int, 8>());
  }
  const int jm = j0 > j1 ? j0 : j1;
  if (jm <= 8) return f(std::integral_constant<int, 8>(), E0());
  if (jm <= 16) return f(std::integral_constant<int, 16>(), E0());
  pnh_set_error("interpolator size %d above 16", jm);
  return -2;
}

template <typename T, int D, bool CSRV, int NBC>
int launch_k2y_tile_j(const InterpArgs& a, const BinArgs& b, const void* grid, void* y, int c0,
                      size_t lds, const int* skip, hipStream_t s) {
  bool split = true;                     // every axis split: the box never wraps
  for (int d = 0; d < D; ++d) split = split && b.tgrid[d] > 1;
  return dispatch_j<D, CSRV>(a, [&](auto jn, auto je) {
    auto go = [&](auto nw) {
      auto kern = k2y_tile_kernel<T, D, CSRV, NBC, decltype(jn)::value, decltype(je)::value,
                                  decltype(nw)::value>;
      allow_lds(kern, lds);
      hipLaunchKernelGGL(kern, dim3((unsigned)b.n_items), dim3(b.k2y_threads), lds, s, a, b,
                         grid, y, c0, skip);
      PNH_LAUNCH_CHECK();
      return 0;
    };
    if (split) return go(std::integral_constant<bool, true>());
    return go(std::integral_constant<bool, false>());
  });
}

// k2y over the tile work list: coils in passes of up to 4 that fit in LDS.
template <typename T, int D, bool CSRV>
int launch_k2y_tiles(const InterpArgs& a, const BinArgs& b, const void* grid, void* y,
                     const int* skip, hipStream_t s) {
  typedef typename cplx<T>::t C;
  if (b.n_items == 0) return 0;
  const size_t lim = pnh_lds_limit();
  int c0 = 0;
  while (c0 < a.nb) {
    const int left = a.nb - c0;
    int nbc = left >= 4 ? 4 : (left >= 2 ? 2 : 1);
    while (nbc > 1 && (size_t)b.pad_cells * nbc * sizeof(C) > lim) nbc >>= 1;
    const size_t lds = (size_t)b.pad_cells * nbc * sizeof(C);
    if (lds > lim) {
      pnh_set_error("interpolation tile needs %zu bytes of LDS (device limit %zu)", lds, lim);
      return -2;
    }
    int rc;
    if (nbc == 4) rc = launch_k2y_tile_j<T, D, CSRV, 4>(a, b, grid, y, c0, lds, skip, s);
    else if (nbc == 2) rc = launch_k2y_tile_j<T, D, CSRV, 2>(a, b, grid, y, c0, lds, skip, s);
    else rc = launch_k2y_tile_j<T, D, CSRV, 1>(a, b, grid, y, c0, lds, skip, s);
    if (rc) return rc;
    c0 += nbc;
  }
  return 0;
}

template <typename T, int D, bool CSRV, int NBC>
int launch_tile_j(const InterpArgs& a, const BinArgs& b, const void* y, void* part, int c0,
                  int* item_e, void* ys, size_t lds, const int* skip, hipStream_t s) {
  // With many items per CU, computing every item's exponent in one pass
  // ahead saves each gridding workgroup a dependent load round and a
  // barrier (3-D 128^3: 4 096 items, 11 us of 650); with few items the extra
  // launch costs more than it saves.
  if (b.n_items < PNH_EXPONENT_PASS_ITEMS) item_e = nullptr;
  if (!item_e) ys = nullptr;
  if (item_e) {
    hipLaunchKernelGGL((y2k_exponent_kernel<T, NBC>), dim3((unsigned)b.n_items), dim3(256), 0,
                       s, a, b, y, c0, item_e, skip, ys);
    PNH_LAUNCH_CHECK();
  }
  return dispatch_j<D, CSRV>(a, [&](auto jn, auto je) {
    return dispatch_pk<T>(packed_items<T>(b), [&](auto pk) {
      return dispatch_pad<D, CSRV, decltype(je)::value>(b, [&](auto pad) {
        auto kern = y2k_tile_kernel<T, D, CSRV, NBC, decltype(jn)::value, decltype(je)::value,
                                    decltype(pk)::value, decltype(pad)::value>;
        allow_lds(kern, lds);
        hipLaunchKernelGGL(kern, dim3((unsigned)b.n_items), dim3(b.threads), lds, s, a, b, y,
                           part, c0, item_e, skip, ys);
        PNH_LAUNCH_CHECK();
        return 0;
      });
    });
  });
}

template <typename T, int D, bool CSRV>
int launch_y2k_nb(const InterpArgs& a, const BinArgs& b, const void* y,
                  void* grid, void* part, const int* skip, hipStream_t s) {
  if (b.threads < 64 || b.threads > 1024 || b.threads % 64) {
    pnh_set_error("gridding workgroup size %d is not a multiple of 64 in 64..1024", b.threads);
    return -1;
  }
  // part: the items' exponents, the row-ordered samples, then the slots
  typedef typename cplx<T>::t C;
  int* item_e = (int*)part;
  void* ys = nullptr;
  if (part) {
    part = (char*)part + pnh_item_exponent_bytes(b.n_items);
    const size_t yb = pnh_row_samples_bytes(b, a.M, a.nb, sizeof(C));
    if (yb) {
      ys = part;
      part = (char*)part + yb;
    }
  }
  int c0 = 0;
  while (c0 < a.nb) {
    const bool pk = packed_items<T>(b);
    const int nbc = pnh_y2k_coils_per_pass(b, a.nb - c0, pk);
    if (!nbc) {
      pnh_set_error("gridding tile needs %zu bytes of LDS (device limit %zu)",
                    y2k_lds_bytes(b.pad_cells, 1, pk), pnh_lds_limit());
      return -2;
    }
    const size_t lds = y2k_lds_bytes(b.pad_cells, nbc, pk);
    if (b.n_items > 0) {
      int rc;
      if (nbc == 4) rc = launch_tile_j<T, D, CSRV, 4>(a, b, y, part, c0, item_e, ys, lds, skip, s);
      else if (nbc == 2) rc = launch_tile_j<T, D, CSRV, 2>(a, b, y, part, c0, item_e, ys, lds, skip, s);
      else rc = launch_tile_j<T, D, CSRV, 1>(a, b, y, part, c0, item_e, ys, lds, skip, s);
      if (rc) return rc;
    }
    const int parts = pnh_gather_parts(b);
    dim3 gg((unsigned)(b.n_tiles * parts));
    if (nbc == 4) hipLaunchKernelGGL((y2k_gather_kernel<T, D, 4>), gg, dim3(256), 0, s, a, b, parts, part, grid, c0, skip);
    else if (nbc == 2) hipLaunchKernelGGL((y2k_gather_kernel<T, D, 2>), gg, dim3(256), 0, s, a, b, parts, part, grid, c0, skip);
    else hipLaunchKernelGGL((y2k_gather_kernel<T, D, 1>), gg, dim3(256), 0, s, a, b, parts, part, grid, c0, skip);
    PNH_LAUNCH_CHECK();
    c0 += nbc;
  }
  return 0;
}

template <typename T, int D, int LPS>
int launch_dom_t(const InterpArgs& a, const void* grid, const void* g, void* out,
                 hipStream_t s) {
  constexpr int SPW = 64 / LPS;
  const int64_t waves = (a.M + SPW - 1) / SPW;
  const int64_t nblk = (waves + 3) / 4;
  for (int c0 = 0; c0 < a.nb;) {
    if (a.nb - c0 >= 4) {
      hipLaunchKernelGGL((k2y_dom_kernel<T, D, LPS, 4>), dim3((unsigned)nblk), dim3(256), 0,
                         s, a, grid, g, (T*)out, c0, c0 > 0);
      c0 += 4;
    } else {
      hipLaunchKernelGGL((k2y_dom_kernel<T, D, LPS, 1>), dim3((unsigned)nblk), dim3(256), 0,
                         s, a, grid, g, (T*)out, c0, c0 > 0);
      c0 += 1;
    }
    PNH_LAUNCH_CHECK();
  }
  return 0;
}

template <typename T, int D>
int launch_dom_j(const InterpArgs& a, const void* grid, const void* g, void* out,
                 hipStream_t s) {
  switch (pnh_lanes_per_sample(a.J[D - 1])) {
    case 1: return launch_dom_t<T, D, 1>(a, grid, g, out, s);
    case 2: return launch_dom_t<T, D, 2>(a, grid, g, out, s);
    case 3: return launch_dom_t<T, D, 3>(a, grid, g, out, s);
    case 4: return launch_dom_t<T, D, 4>(a, grid, g, out, s);
    case 6: return launch_dom_t<T, D, 6>(a, grid, g, out, s);
    case 8: return launch_dom_t<T, D, 8>(a, grid, g, out, s);
    case 16: return launch_dom_t<T, D, 16>(a, grid, g, out, s);
    default:
      pnh_set_error("interpolator size %d on the last axis is above 16", a.J[D - 1]);
      return -2;
  }
}

}  // namespace

namespace {

// P^H P on one work item: k2y of the item's rows from the padded box of the
// input grid, then y2k of those rows into the same-sized accumulator (the
// rows' support boxes lie in the tile's padded box for both).  The box and the
// accumulator share the LDS (the box is dead once y is formed); y goes
// through the caller's sample buffer, in row order (y[r], contiguous over the
// item's rows), which the barrier makes visible to the whole workgroup, so
// the factors and samples of the second phase come from L2 rather than HBM.
// One coil, per-axis factors.
template <typename T, int D, int JN, int JE, bool PK>
__device__ __forceinline__ void k2y2k_item(const InterpArgs& a, const BinArgs& b, const int4 it,
                                           int item, long long* smem,
                                           const void* __restrict__ k_, void* __restrict__ y_,
                                           void* __restrict__ part_) {
  typedef typename cplx<T>::t C;
  k2y_item<T, D, false, 1, JN, JE>(a, b, it, reinterpret_cast<C*>(smem), k_, y_, 0, true);
  __syncthreads();
  y2k_item<T, D, false, 1, JN, JE, PK>(a, b, it, item, smem, y_, part_, 0, nullptr, y_);
}

// P^H P on one work item in one pass over its rows (FP64, one coil,
// per-axis factors).  The item's padded box of the input and the accumulator
// are both in LDS; each row's factors are loaded once, its sample
// y_r = sum_j coef(r, j) box[cell(r, j)] is formed by the row's lanes (as
// k2y_item) and at once gridded back as conj(coef(r, j)) y_r (as y2k_item),
// so the samples never leave registers.  The fixed-point exponent is fixed
// before the first add from a bound of every sample of the item:
// |Re y_r| + |Im y_r| <= sqrt(2) |y_r| <= sqrt(2) row_bound m, m the largest
// |Re k| + |Im k| of the box; with 2^e > 1.5 row_bound u_bound m every
// contribution's parts stay below 2^e, as the two-phase pass guarantees
// from the exact samples.  A non-finite sample poisons the item's slot, as
// in the two-phase pass.  LDS: box [pad_cells] complex128, accumulator
// [pad_cells][2] int64, then 16 doubles of reduction scratch and a flag.
// PADC > 0: every axis split with that padded extent (compile-time offsets).
template <int D, int JN, int JE, int PADC = 0>
__device__ __forceinline__ void k2y2k_item_1p(const InterpArgs& a, const BinArgs& b,
                                              const int4 it, long long* smem,
                                              const void* __restrict__ k_,
                                              void* __restrict__ part_) {
  typedef double2 C;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int pc = (int)b.pad_cells;
  C* __restrict__ box = reinterpret_cast<C*>(smem);
  long long* __restrict__ acc = smem + 2 * (int64_t)pc;
  double* red = reinterpret_cast<double*>(acc + 2 * (int64_t)pc);
  int* poisoned = reinterpret_cast<int*>(red + 16);
  const TileFrame f = tile_frame<D>(a, b, it.x);
  constexpr bool FP = PADC > 0;          // every axis split, extent PADC
  const int P1 = D > 1 ? (FP ? PADC : b.pad[1]) : 1;
  const int P2 = D > 2 ? (FP ? PADC : b.pad[2]) : 1;
  const C* __restrict__ grid = (const C*)k_;
  // padded box -> LDS (last axis fastest), accumulator cleared, box maximum
  double kmax = 0.0;
  {
    longlong2* acc2 = reinterpret_cast<longlong2*>(acc);
    for_box_cells<D>(pc, P1, P2, tid, nthr, [&](int i, int x0, int x1, int x2) {
      int c = f.T0[0] + x0;
      int g = (c >= a.K[0] ? c - a.K[0] : c) * (int)a.Kstride[0];
      if (D > 1) {
        c = f.T0[1] + x1;
        g += (c >= a.K[1] ? c - a.K[1] : c) * (int)a.Kstride[1];
      }
      if (D > 2) {
        c = f.T0[2] + x2;
        g += c >= a.K[2] ? c - a.K[2] : c;
      }
      const C v = grid[g];
      box[i] = v;
      acc2[i] = make_longlong2(0, 0);
      // non-finite values are caught in the samples they reach
      const double m = fabs(v.x) + fabs(v.y);
      if (m <= 1.0e300) kmax = fmax(kmax, m);
    });
  }
  if (tid == 0) *poisoned = 0;
  kmax = block_max(kmax, red);            // also the barrier after the box writes
  // 1.5 > sqrt(2), with room for the rounding of the products
  const double B = 1.5 * b.row_bound * b.u_bound * kmax;
  int e = 0;
  if (B > 1.0e300) e = 1000;
  else if (B > 0.0) (void)frexp(B, &e);   // B < 2^e
  if (e < -960) e = -960;
  const double scale = ldexp(1.0, kFixBits - e);
  const double unscale = ldexp(1.0, e - kFixBits);

  const int JL = a.J[D - 1];
  const int LPS = pnh_lanes_per_sample(JL);
  const int SPW = 64 / LPS;
  const int lane = tid & 63, wave = tid >> 6, nw = nthr >> 6;
  const int sl = lane / LPS;
  const int jl = lane - sl * LPS;
  const C* __restrict__ u = (const C*)a.u;
  const int J0 = JE > 0 ? JE : a.J[0];
  const int J1 = D > 2 ? (JE > 0 ? JE : a.J[1]) : 1;
  const int r_end = it.y + it.z;
  constexpr int CB = 16;                  // bytes of one accumulator cell
  char* const accb = reinterpret_cast<char*>(acc);
  bool bad = false;

  for (int r0 = it.y + wave * SPW; r0 < r_end; r0 += nw * SPW) {
    const int r = r0 + sl;
    const bool valid = sl < SPW && jl < JL && r < r_end;
    C ul = make_double2(0.0, 0.0);
    C s = make_double2(0.0, 0.0);
    int xl = 0;
    C u0c[JN], u1c[JN];
    int row0[JN], off1[JN];
    if (valid) {
      xl = box_local<FP>(a.base[(int64_t)(D - 1) * a.M + r] - f.T0[D - 1] + jl, f.wrap[D - 1],
                      a.K[D - 1]);
      ul = u[a.u_off[D - 1] + (int64_t)r * JL + jl];
      if (D == 1) {
        s = cmulf(ul, box[xl]);
      } else {
        const int x0b = a.base[r] - f.T0[0];
        const C* __restrict__ u0 = u + a.u_off[0] + (int64_t)r * J0;
#pragma unroll
        for (int j = 0; j < JN; ++j)
          if (JE > 0 || j < J0) {
            u0c[j] = u0[j];
            row0[j] = box_local<FP>(x0b + j, f.wrap[0], a.K[0]) * P1 * P2;
          }
        C s0 = make_double2(0.0, 0.0);
        if (D == 2) {
#pragma unroll
          for (int j0 = 0; j0 < JN; ++j0)
            if (JE > 0 || j0 < J0) s0 = cfma(u0c[j0], box[row0[j0] + xl], s0);
        } else {
          const int x1b = a.base[a.M + r] - f.T0[1];
          const C* __restrict__ u1 = u + a.u_off[1] + (int64_t)r * J1;
#pragma unroll
          for (int j = 0; j < JN; ++j)
            if (JE > 0 || j < J1) {
              u1c[j] = u1[j];
              off1[j] = box_local<FP>(x1b + j, f.wrap[1], a.K[1]) * P2 + xl;
            }
#pragma unroll
          for (int j0 = 0; j0 < JN; ++j0) {
            if (JE > 0 || j0 < J0) {
              C s1 = make_double2(0.0, 0.0);
#pragma unroll
              for (int j1 = 0; j1 < JN; ++j1)
                if (JE > 0 || j1 < J1) s1 = cfma(u1c[j1], box[row0[j0] + off1[j1]], s1);
              s0 = cfma(u0c[j0], s1, s0);
            }
          }
        }
        s = cmulf(ul, s0);
      }
    }
    // sum over the row's lanes (all lanes take part), then every lane of
    // the row takes the total from the row's first lane
    for (int off = 1; off < LPS; off <<= 1) {
      const C o = shfl_down_c(s, off);
      if (jl + off < LPS) s = cadd(s, o);
    }
    C y;
    y.x = __shfl(s.x, lane - jl, 64);
    y.y = __shfl(s.y, lane - jl, 64);
    if (!valid) continue;
    if (!(fabs(y.x) + fabs(y.y) <= 1.0e300)) {
      bad = true;
      continue;
    }
    const double2 ys = make_double2(y.x * scale, y.y * scale);
    const double2 wl = dmul(cconj(ul), ys);
    if (D == 1) {
      lds_add_fixed(acc + 2 * xl, wl);
      continue;
    }
    if (D == 2) {
#pragma unroll
      for (int j0 = 0; j0 < JN; ++j0)
        if (JE > 0 || j0 < J0)
          lds_add_prod(reinterpret_cast<long long*>(accb + (row0[j0] + xl) * CB),
                       cconj(u0c[j0]), wl);
      continue;
    }
#pragma unroll
    for (int j0 = 0; j0 < JN; ++j0) {
      if (JE > 0 || j0 < J0) {
        const double2 a0 = dmul(cconj(u0c[j0]), wl);
#pragma unroll
        for (int j1 = 0; j1 < JN; ++j1)
          if (JE > 0 || j1 < J1)
            lds_add_prod(reinterpret_cast<long long*>(accb + (row0[j0] + off1[j1]) * CB),
                         cconj(u1c[j1]), a0);
      }
    }
  }
  if (bad) *poisoned = 1;
  __syncthreads();
  const bool poison = *poisoned != 0;
  C* __restrict__ part = (C*)part_ + (int64_t)it.w * pc;
  const double nan = __builtin_nan("");
  for (int i = tid; i < pc; i += nthr)
    part[i] = poison ? make_double2(nan, nan)
                     : make_double2((double)acc[2 * i] * unscale, (double)acc[2 * i + 1] * unscale);
}

template <int D, int JN, int JE, int PADC>
__global__ void __launch_bounds__(1024)
k2y2k_1p_tile_kernel(InterpArgs a, BinArgs b, const void* __restrict__ k_,
                     void* __restrict__ part_, const int* __restrict__ skip) {
  if (skip && *skip) return;
  extern __shared__ long long smem_kk1[];
  const int4 it = reinterpret_cast<const int4*>(b.items)[blockIdx.x];
  k2y2k_item_1p<D, JN, JE, PADC>(a, b, it, smem_kk1, k_, part_);
}

template <typename T, int D, int JN, int JE, bool PK>
__global__ void __launch_bounds__(1024)
k2y2k_tile_kernel(InterpArgs a, BinArgs b, const void* __restrict__ k_, void* __restrict__ y_,
                  void* __restrict__ part_, const int* __restrict__ skip) {
  if (skip && *skip) return;
  extern __shared__ long long smem_kk[];
  const int4 it = reinterpret_cast<const int4*>(b.items)[blockIdx.x];
  k2y2k_item<T, D, JN, JE, PK>(a, b, it, blockIdx.x, smem_kk, k_, y_, part_);
}

template <typename T, int D>
int launch_k2y2k(const InterpArgs& a, const BinArgs& b, const void* k, void* y, void* grid,
                 void* part, const int* skip, hipStream_t s) {
  void* slots = (char*)part + pnh_item_exponent_bytes(b.n_items);
  const size_t lds = y2k_lds_bytes(b.pad_cells, 1, packed_items<T>(b));
  const size_t lds1 = k2y2k_1p_lds_bytes(b.pad_cells);
  if (sizeof(T) == 8 && b.row_bound > 0.0 && lds1 <= pnh_lds_limit()) {
    // FP64 with a row bound: one pass over each item's rows
    if (b.n_items > 0) {
      int rc = dispatch_j<D, false>(a, [&](auto jn, auto je) {
        return dispatch_pad<D, false, decltype(je)::value>(b, [&](auto pad) {
          auto kern = k2y2k_1p_tile_kernel<D, decltype(jn)::value, decltype(je)::value,
                                           decltype(pad)::value>;
          allow_lds(kern, lds1);
          hipLaunchKernelGGL(kern, dim3((unsigned)b.n_items), dim3(b.threads), lds1, s, a, b,
                             k, slots, skip);
          PNH_LAUNCH_CHECK();
          return 0;
        });
      });
      if (rc) return rc;
    }
  } else if (b.n_items > 0) {
    int rc = dispatch_j<D, false>(a, [&](auto jn, auto je) {
      return dispatch_pk<T>(packed_items<T>(b), [&](auto pk) {
        auto kern = k2y2k_tile_kernel<T, D, decltype(jn)::value, decltype(je)::value,
                                      decltype(pk)::value>;
        allow_lds(kern, lds);
        hipLaunchKernelGGL(kern, dim3((unsigned)b.n_items), dim3(b.threads), lds, s, a, b, k,
                           y, slots, skip);
        PNH_LAUNCH_CHECK();
        return 0;
      });
    });
    if (rc) return rc;
  }
  const int parts = pnh_gather_parts(b);
  hipLaunchKernelGGL((y2k_gather_kernel<T, D, 1>), dim3((unsigned)(b.n_tiles * parts)),
                     dim3(256), 0, s, a, b, parts, slots, grid, 0, skip);
  PNH_LAUNCH_CHECK();
  return 0;
}

}  // namespace

int pnh_k2y2k_launch(const InterpArgs& a, const BinArgs& b, int fp64, const void* k, void* y,
                     void* grid, void* part, const int* skip, hipStream_t s) {
  const size_t esz = fp64 ? 16 : 8;
  const bool pk = !fp64 && b.item_bound;
  const bool fused = a.nb == 1 && !a.csr_val && b.threads >= 64 && b.threads <= 1024 &&
                     b.threads % 64 == 0 &&
                     y2k_lds_bytes(b.pad_cells, 1, pk) <= pnh_lds_limit() &&
                     (size_t)b.pad_cells * esz <= y2k_lds_bytes(b.pad_cells, 1, pk);
  if (!fused) {
    int rc = a.M > 0 ? pnh_k2y_launch(a, b, fp64, k, y, skip, s) : 0;
    return rc ? rc : pnh_y2k_launch(a, b, fp64, y, grid, part, skip, s);
  }
  if (b.n_items > 0 && !part) {
    pnh_set_error("workspace required for gridding");
    return -1;
  }
  if (fp64) {
    switch (a.D) {
      case 1: return launch_k2y2k<double, 1>(a, b, k, y, grid, part, skip, s);
      case 2: return launch_k2y2k<double, 2>(a, b, k, y, grid, part, skip, s);
      case 3: return launch_k2y2k<double, 3>(a, b, k, y, grid, part, skip, s);
    }
  } else {
    switch (a.D) {
      case 1: return launch_k2y2k<float, 1>(a, b, k, y, grid, part, skip, s);
      case 2: return launch_k2y2k<float, 2>(a, b, k, y, grid, part, skip, s);
      case 3: return launch_k2y2k<float, 3>(a, b, k, y, grid, part, skip, s);
    }
  }
  pnh_set_error("unsupported ndim %d", a.D);
  return -2;
}

int pnh_k2y_launch(const InterpArgs& a, const BinArgs& b, int fp64, const void* grid, void* y,
                   const int* skip, hipStream_t s) {
  const bool csrv = a.csr_val != nullptr;
  if (b.k2y_threads < 64 || b.k2y_threads > 1024 || b.k2y_threads % 64) {
    pnh_set_error("workgroup size %d is not a multiple of 64 in 64..1024", b.k2y_threads);
    return -1;
  }
#define K2Y_CASE(T, D)                                                       \
  return csrv ? launch_k2y_tiles<T, D, true>(a, b, grid, y, skip, s)         \
              : launch_k2y_tiles<T, D, false>(a, b, grid, y, skip, s);
#define K2Y_CASE64(D) return launch_k2y_tiles<double, D, false>(a, b, grid, y, skip, s);
  if (fp64) {
    if (csrv) { pnh_set_error("FP64 operators use the per-axis factors"); return -2; }
    switch (a.D) { case 1: K2Y_CASE64(1) case 2: K2Y_CASE64(2) case 3: K2Y_CASE64(3) }
  } else {
    switch (a.D) { case 1: K2Y_CASE(float, 1) case 2: K2Y_CASE(float, 2) case 3: K2Y_CASE(float, 3) }
  }
#undef K2Y_CASE
#undef K2Y_CASE64
  pnh_set_error("unsupported ndim %d", a.D);
  return -2;
}

int pnh_y2k_launch(const InterpArgs& a, const BinArgs& b, int fp64,
                   const void* y, void* grid, void* part, const int* skip,
                   hipStream_t s) {
  const bool csrv = a.csr_val != nullptr;
#define Y2K_CASE(T, D)                                                       \
  return csrv ? launch_y2k_nb<T, D, true>(a, b, y, grid, part, skip, s)      \
              : launch_y2k_nb<T, D, false>(a, b, y, grid, part, skip, s);
#define Y2K_CASE64(D) return launch_y2k_nb<double, D, false>(a, b, y, grid, part, skip, s);
  if (fp64) {
    if (csrv) { pnh_set_error("FP64 operators use the per-axis factors"); return -2; }
    switch (a.D) { case 1: Y2K_CASE64(1) case 2: Y2K_CASE64(2) case 3: Y2K_CASE64(3) }
  } else {
    switch (a.D) { case 1: Y2K_CASE(float, 1) case 2: Y2K_CASE(float, 2) case 3: Y2K_CASE(float, 3) }
  }
#undef Y2K_CASE
#undef Y2K_CASE64
  pnh_set_error("unsupported ndim %d", a.D);
  return -2;
}

namespace {

// Per-item bound of the packed accumulators (BinArgs::item_bound): the
// largest over the cells of the item's padded box of the sum over the
// item's rows of |u_0| ... |u_last| (complex64 factors, magnitudes in FP64),
// summed in 64-bit fixed point (so the bound is reproducible), then rounded
// up.  One workgroup per item, one thread per (row, last-axis offset).
template <int D>
__global__ void __launch_bounds__(256)
item_bound_kernel(InterpArgs a, BinArgs b, double* __restrict__ bound) {
  extern __shared__ long long smem_ib[];
  __shared__ double red[16];
  long long* acc = smem_ib;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int pc = (int)b.pad_cells;
  const int4 it = reinterpret_cast<const int4*>(b.items)[blockIdx.x];
  for (int i = tid; i < pc; i += nthr) acc[i] = 0;
  // every product of factor magnitudes is at most u_bound < 2^eu: units of
  // 2^(eu - 50), at most 4096 rows per cell, so no overflow
  int eu = 0;
  (void)frexp(b.u_bound, &eu);
  const double scale = ldexp(1.0, kFixBits - eu);
  __syncthreads();
  const TileFrame f = tile_frame<D>(a, b, it.x);
  const float2* __restrict__ u = (const float2*)a.u;
  const int JL = a.J[D - 1];
  const int J0 = D > 1 ? a.J[0] : 1, J1 = D > 2 ? a.J[1] : 1;
  const int P1 = D > 1 ? b.pad[1] : 1, P2 = D > 2 ? b.pad[2] : 1;
  const int n = it.z * JL;
  for (int i = tid; i < n; i += nthr) {
    const int q = i / JL;
    const int r = it.y + q, jl = i - q * JL;
    const int xl = wrap_local(a.base[(int64_t)(D - 1) * a.M + r] - f.T0[D - 1] + jl,
                              f.wrap[D - 1], a.K[D - 1]);
    const float2 vl = u[a.u_off[D - 1] + (int64_t)r * JL + jl];
    const double wl = hypot((double)vl.x, (double)vl.y) * scale;
    const int x0b = D > 1 ? a.base[r] - f.T0[0] : 0;
    const int x1b = D > 2 ? a.base[a.M + r] - f.T0[1] : 0;
    for (int j0 = 0; j0 < J0; ++j0) {
      double w0 = wl;
      int c0 = xl;
      if (D > 1) {
        const float2 v0 = u[a.u_off[0] + (int64_t)r * J0 + j0];
        w0 *= hypot((double)v0.x, (double)v0.y);
        c0 += wrap_local(x0b + j0, f.wrap[0], a.K[0]) * P1 * P2;
      }
      for (int j1 = 0; j1 < J1; ++j1) {
        double w = w0;
        int c = c0;
        if (D > 2) {
          const float2 v1 = u[a.u_off[1] + (int64_t)r * J1 + j1];
          w *= hypot((double)v1.x, (double)v1.y);
          c += wrap_local(x1b + j1, f.wrap[1], a.K[1]) * P2;
        }
        lds_add(acc + c, __double_as_longlong(w + kMagic) - kMagicBits);
      }
    }
  }
  __syncthreads();
  double m = 0.0;
  for (int i = tid; i < pc; i += nthr) m = fmax(m, (double)acc[i]);
  m = block_max(m, red);
  // a rounded term is at most half a unit below its value; margin for the
  // FP64 products and the conversion
  if (tid == 0)
    bound[blockIdx.x] = (m + 0.5 * n + 1.0) * ldexp(1.0, eu - kFixBits) * (1.0 + 0x1p-20);
}

}  // namespace

int pnh_item_bounds_launch(const InterpArgs& a, const BinArgs& b, double* bound, hipStream_t s) {
  if (b.n_items == 0) return 0;
  if (!a.u || !bound) {
    pnh_set_error("item bounds need the complex64 factors and an output buffer");
    return -1;
  }
  const size_t lds = (size_t)b.pad_cells * 8;
  if (lds > pnh_lds_limit()) {
    pnh_set_error("gridding tile needs %zu bytes of LDS (device limit %zu)", lds,
                  pnh_lds_limit());
    return -2;
  }
  switch (a.D) {
#define IB(DD)                                                                            \
    case DD:                                                                              \
      allow_lds(item_bound_kernel<DD>, lds);                                              \
      hipLaunchKernelGGL(item_bound_kernel<DD>, dim3((unsigned)b.n_items), dim3(256), lds, s, \
                         a, b, bound);                                                    \
      PNH_LAUNCH_CHECK();                                                                 \
      return 0;
    IB(1) IB(2) IB(3)
#undef IB
  }
  pnh_set_error("unsupported ndim %d", a.D);
  return -2;
}

size_t pnh_lds_limit() {
  static int cached[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (!cached[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess || v <= 0)
      v = 65536;
    cached[dev] = v;
  }
  return (size_t)cached[dev];
}

int pnh_y2k_coils_per_pass(const BinArgs& b, int nb, bool packed) {
  const size_t lim = pnh_lds_limit();
  for (int nbc = nb >= 4 ? 4 : (nb >= 2 ? 2 : 1); nbc >= 1; nbc >>= 1)
    if (y2k_lds_bytes(b.pad_cells, nbc, packed) <= lim) return nbc;
  return 0;
}

int pnh_k2y_dom_launch(const InterpArgs& a, int fp64, const void* grid,
                       const void* g, void* out, hipStream_t s) {
  if (!a.du) { pnh_set_error("plan has no factor derivatives"); return -1; }
  if (a.M == 0) return 0;
  if (fp64) {
    switch (a.D) {
      case 1: return launch_dom_j<double, 1>(a, grid, g, out, s);
      case 2: return launch_dom_j<double, 2>(a, grid, g, out, s);
      case 3: return launch_dom_j<double, 3>(a, grid, g, out, s);
    }
  } else {
    switch (a.D) {
      case 1: return launch_dom_j<float, 1>(a, grid, g, out, s);
      case 2: return launch_dom_j<float, 2>(a, grid, g, out, s);
      case 3: return launch_dom_j<float, 3>(a, grid, g, out, s);
    }
  }
  pnh_set_error("unsupported ndim %d", a.D);
  return -2;
}

// ------------------------------------------------------------- batches ----
// The same kernels over many independent problems in one launch: each block
// (k2y) or wavefront (y2k) looks up its problem and runs the single-problem
// body on that problem's slice of the packed vectors.
#include "batch.h"

namespace {

template <typename T, int D, int JN, int JE>
__global__ void __launch_bounds__(1024)
k2y_tile_batch_kernel(const BatchProblem* __restrict__ probs, const int2* __restrict__ items,
                      const void* __restrict__ k_, void* __restrict__ y_,
                      const int* __restrict__ done, const int* __restrict__ all_done) {
  typedef typename cplx<T>::t C;
  if (all_done && *all_done) return;
  const int2 pi = items[blockIdx.x];
  const int p = pi.x;
  if (done && done[p]) return;            // uniform over the workgroup
  extern __shared__ long long smem_box[];
  const InterpArgs a = sizeof(T) == 8 ? probs[p].ia64 : probs[p].ia;
  const BinArgs b = probs[p].bins;
  const int4 it = reinterpret_cast<const int4*>(b.items)[pi.y];
  k2y_item<T, D, false, 1, JN, JE>(a, b, it, reinterpret_cast<C*>(smem_box),
                               (const C*)k_ + probs[p].k_off, (C*)y_ + probs[p].y_off, 0);
}

template <typename T, int D, int JN, int JE, bool PK>
__global__ void __launch_bounds__(1024)
y2k_tile_batch_kernel(const BatchProblem* __restrict__ probs, const int2* __restrict__ items,
                      const void* __restrict__ y_, void* __restrict__ part_,
                      const int* __restrict__ done, const int* __restrict__ all_done) {
  typedef typename cplx<T>::t C;
  if (all_done && *all_done) return;
  const int2 pi = items[blockIdx.x];
  const int p = pi.x;
  if (done && done[p]) return;            // uniform over the workgroup
  extern __shared__ long long smem_acc[];
  const InterpArgs a = sizeof(T) == 8 ? probs[p].ia64 : probs[p].ia;
  const BinArgs b = probs[p].bins;
  const int4 it = reinterpret_cast<const int4*>(b.items)[pi.y];
  y2k_item<T, D, false, 1, JN, JE, PK>(a, b, it, pi.y, smem_acc, (const C*)y_ + probs[p].y_off,
                                   (C*)part_ + probs[p].part_off, 0, nullptr);
}

// P^H P of every problem's work items (k2y then y2k per item, as k2y2k_item;
// FP64 problems with a row bound in one pass, as k2y2k_item_1p, when
// one_pass: the launch's LDS holds that pass's box and accumulator).  PADC:
// every problem's boxes have the extent PNH_FIXED_PAD on every axis.
template <typename T, int D, int JN, int JE, int PADC>
__global__ void __launch_bounds__(1024)
k2y2k_tile_batch_kernel(const BatchProblem* __restrict__ probs, const int2* __restrict__ items,
                        const void* __restrict__ k_, void* __restrict__ y_,
                        void* __restrict__ part_, const int* __restrict__ done,
                        const int* __restrict__ all_done, int one_pass) {
  typedef typename cplx<T>::t C;
  if (all_done && *all_done) return;
  const int2 pi = items[blockIdx.x];
  const int p = pi.x;
  if (done && done[p]) return;            // uniform over the workgroup
  extern __shared__ long long smem_kkb[];
  const InterpArgs a = sizeof(T) == 8 ? probs[p].ia64 : probs[p].ia;
  const BinArgs b = probs[p].bins;
  const int4 it = reinterpret_cast<const int4*>(b.items)[pi.y];
  if constexpr (sizeof(T) == 8) {
    // PADC > 0: every problem has a row bound (the launch checked), so the
    // kernel holds the one-pass variant only (fewer registers)
    if (PADC > 0 || (one_pass && b.row_bound > 0.0)) {  // uniform over the workgroup
      k2y2k_item_1p<D, JN, JE, PADC>(a, b, it, smem_kkb, (const C*)k_ + probs[p].k_off,
                                     (C*)part_ + probs[p].part_off);
      return;
    }
  }
  if constexpr (PADC == 0) {
    k2y2k_item<T, D, JN, JE, false>(a, b, it, pi.y, smem_kkb, (const C*)k_ + probs[p].k_off,
                                    (C*)y_ + probs[p].y_off, (C*)part_ + probs[p].part_off);
  }
}

// Gather of every problem's grid: pnh_gather_parts blocks per tile, blocks by table.
template <typename T, int D>
__global__ void __launch_bounds__(256)
gather_batch_kernel(const BatchProblem* __restrict__ probs, BatchTable t,
                    const void* __restrict__ part_, void* __restrict__ k_,
                    const int* __restrict__ done, const int* __restrict__ all_done) {
  typedef typename cplx<T>::t C;
  if (all_done && *all_done) return;
  const int p = t.tab[blockIdx.x];
  if (done && done[p]) return;
  const int lb = (int)(blockIdx.x - t.first[p]);
  const InterpArgs a = probs[p].ia;
  const BinArgs b = probs[p].bins;
  const int parts = pnh_gather_parts(b);
  const int tile = lb / parts, q = lb - tile * parts;
  const int span = parts == 1 ? (1 << 30) : 256;
  gather_tile<T, D, 1>(a, b, tile, q * 256, span, (const C*)part_ + probs[p].part_off,
                       (C*)k_ + probs[p].k_off, 0);
}

}  // namespace

namespace {

// Host-side argument block holding only the common support sizes (for the
// support-loop dispatch of batch launches).
InterpArgs j_only(const int* Jd, int D) {
  InterpArgs a;
  memset(&a, 0, sizeof(a));
  a.D = D;
  for (int d = 0; d < D; ++d) a.J[d] = Jd[d];
  return a;
}

template <typename T, int D>
int batch_k2y_t(const BatchProblem* probs, const int2* items, int64_t n_items, const int* Jd,
                size_t lds, int threads, const void* k, void* y, const int* done,
                const int* all_done, hipStream_t s) {
  return dispatch_j<D, false>(j_only(Jd, D), [&](auto jn, auto je) {
    auto kern = k2y_tile_batch_kernel<T, D, decltype(jn)::value, decltype(je)::value>;
    allow_lds(kern, lds);
    hipLaunchKernelGGL(kern, dim3((unsigned)n_items), dim3(threads), lds, s, probs, items, k, y,
                       done, all_done);
    PNH_LAUNCH_CHECK();
    return 0;
  });
}

template <typename T, int D>
int batch_y2k_t(const BatchProblem* probs, const int2* items, int64_t n_items, const int* Jd,
                size_t lds, int threads, bool packed, const void* y, void* part,
                const int* done, const int* all_done, hipStream_t s) {
  return dispatch_j<D, false>(j_only(Jd, D), [&](auto jn, auto je) {
    return dispatch_pk<T>(packed, [&](auto pk) {
      auto kern = y2k_tile_batch_kernel<T, D, decltype(jn)::value, decltype(je)::value,
                                        decltype(pk)::value>;
      allow_lds(kern, lds);
      hipLaunchKernelGGL(kern, dim3((unsigned)n_items), dim3(threads), lds, s, probs, items, y,
                         part, done, all_done);
      PNH_LAUNCH_CHECK();
      return 0;
    });
  });
}

}  // namespace

int pnh_batch_k2y_launch(const BatchProblem* probs, const int2* items, int64_t n_items, int D,
                         const int* Jd, int fp64, int64_t pad_cells, int threads, const void* k,
                         void* y, const int* done, const int* all_done, hipStream_t s) {
  if (n_items == 0) return 0;
  const size_t lds = (size_t)pad_cells * (fp64 ? 16 : 8);
  if (lds > pnh_lds_limit()) {
    pnh_set_error("interpolation tile needs %zu bytes of LDS (device limit %zu)", lds,
                  pnh_lds_limit());
    return -2;
  }
#define KB(T, DD) return batch_k2y_t<T, DD>(probs, items, n_items, Jd, lds, threads, k, y, done, all_done, s)
  if (fp64) { if (D == 1) KB(double, 1); if (D == 2) KB(double, 2); KB(double, 3); }
  if (D == 1) KB(float, 1);
  if (D == 2) KB(float, 2);
  KB(float, 3);
#undef KB
}

namespace {
template <typename T, int D>
int batch_k2y2k_t(const BatchProblem* probs, const int2* items, int64_t n_items, const int* Jd,
                  size_t lds, int threads, const void* k, void* y, void* part, const int* done,
                  const int* all_done, int one_pass, int fixed_pad, hipStream_t s) {
  return dispatch_j<D, false>(j_only(Jd, D), [&](auto jn, auto je) {
    auto go = [&](auto pad) {
      auto kern = k2y2k_tile_batch_kernel<T, D, decltype(jn)::value, decltype(je)::value,
                                          decltype(pad)::value>;
      allow_lds(kern, lds);
      hipLaunchKernelGGL(kern, dim3((unsigned)n_items), dim3(threads), lds, s, probs, items, k,
                         y, part, done, all_done, one_pass);
      PNH_LAUNCH_CHECK();
      return 0;
    };
    if constexpr (sizeof(T) == 8 && decltype(je)::value == 6 && D >= 2) {
      if (fixed_pad && one_pass) return go(std::integral_constant<int, PNH_FIXED_PAD>());
    }
    return go(std::integral_constant<int, 0>());
  });
}
}  // namespace

int pnh_batch_k2y2k_launch(const BatchProblem* probs, const int2* items, int64_t n_items,
                           const BatchTable& gather, int D, const int* Jd, int fp64,
                           int64_t pad_cells, int threads, int fixed_pad, const void* k, void* y,
                           void* k_out, void* part, const int* done, const int* all_done,
                           hipStream_t s) {
  size_t lds = y2k_lds_bytes(pad_cells, 1, false);
  if (lds > pnh_lds_limit()) {
    pnh_set_error("gridding tile needs %zu bytes of LDS (device limit %zu)", lds,
                  pnh_lds_limit());
    return -2;
  }
  const int one_pass = fp64 && k2y2k_1p_lds_bytes(pad_cells) <= pnh_lds_limit();
  if (one_pass) lds = k2y2k_1p_lds_bytes(pad_cells);
  int rc = 0;
#define KK(T, DD) rc = batch_k2y2k_t<T, DD>(probs, items, n_items, Jd, lds, threads, k, y, part, done, all_done, one_pass, fixed_pad, s)
  if (n_items > 0) {
    if (fp64) { if (D == 1) KK(double, 1); else if (D == 2) KK(double, 2); else KK(double, 3); }
    else { if (D == 1) KK(float, 1); else if (D == 2) KK(float, 2); else KK(float, 3); }
    if (rc) return rc;
  }
#undef KK
#define GB(T, DD) hipLaunchKernelGGL((gather_batch_kernel<T, DD>), dim3((unsigned)gather.n_blocks), \
                                     dim3(256), 0, s, probs, gather, part, k_out, done, all_done)
  if (gather.n_blocks > 0) {
    if (fp64) { if (D == 1) GB(double, 1); else if (D == 2) GB(double, 2); else GB(double, 3); }
    else { if (D == 1) GB(float, 1); else if (D == 2) GB(float, 2); else GB(float, 3); }
    PNH_LAUNCH_CHECK();
  }
#undef GB
  return 0;
}

int pnh_batch_y2k_launch(const BatchProblem* probs, const int2* items, int64_t n_items,
                         const BatchTable& gather, int D, const int* Jd, int fp64, int packed,
                         int64_t pad_cells, int threads, const void* y, void* k, void* part,
                         const int* done, const int* all_done, hipStream_t s) {
  const bool pk = packed && !fp64;
  const size_t lds = y2k_lds_bytes(pad_cells, 1, pk);
  if (lds > pnh_lds_limit()) {
    pnh_set_error("gridding tile needs %zu bytes of LDS (device limit %zu)", lds,
                  pnh_lds_limit());
    return -2;
  }
  int rc = 0;
#define YB(T, DD) rc = batch_y2k_t<T, DD>(probs, items, n_items, Jd, lds, threads, pk, y, part, done, all_done, s)
  if (n_items > 0) {
    if (fp64) { if (D == 1) YB(double, 1); else if (D == 2) YB(double, 2); else YB(double, 3); }
    else { if (D == 1) YB(float, 1); else if (D == 2) YB(float, 2); else YB(float, 3); }
    if (rc) return rc;
  }
#undef YB
#define GB(T, DD) hipLaunchKernelGGL((gather_batch_kernel<T, DD>), dim3((unsigned)gather.n_blocks), \
                                     dim3(256), 0, s, probs, gather, part, k, done, all_done)
  if (gather.n_blocks > 0) {
    if (fp64) { if (D == 1) GB(double, 1); else if (D == 2) GB(double, 2); else GB(double, 3); }
    else { if (D == 1) GB(float, 1); else if (D == 2) GB(float, 2); else GB(float, 3); }
    PNH_LAUNCH_CHECK();
  }
#undef GB
  return 0;
}
