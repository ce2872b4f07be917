// Interpolation (k2y) and deterministic gridding (y2k) for the min-max NUFFT.
//
// Samples are stored in a trajectory-sorted order (row r -> sample perm[r]),
// so neighbouring lanes gather neighbouring grid cells.  The coefficient of
// row r at support offset (j0, j1, j2) is the Kronecker product of the FP64
// per-dimension min-max factors, ((u0[j0] * u1[j1]) * u2[j2]) -- the product
// order of the reference's Khatri-Rao construction -- either computed on the
// fly from the factors (FACT) or read from a stored complex64 table (CSR,
// rows interleaved in slices of 64 so that lane r%64 reads entry j of its row
// with unit stride).  The support cell along dim d is (base_d + j_d) mod K_d.
//
// y2k is a gather-free deterministic gridding: the K grid is cut into tiles,
// each tile has the sorted list of rows whose support meets it, and one
// wavefront accumulates one tile (or one chunk of a heavy tile) in its own
// LDS region, one sample per step, so every cell receives its contributions
// in the same order on every run.  Heavy tiles are split into chunks whose
// partial tiles are summed in chunk order by a merge kernel.  No float
// atomics touch global memory.
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

// ---------------------------------------------------------------- k2y ----
// JLC lanes per sample (one per last-axis support offset), SPW = 64/JLC
// samples per wavefront; per-lane partial sums are combined with a segmented
// shuffle reduction.  JLC == 0: generic, 16 lanes per sample (J_last <= 16).
// y[perm[r]*nb + c0 + c] = sum_j coef(r, j) * grid[cell(r, j)*nb + c0 + c]
template <typename T, int D, int JLC, bool CSRV, int NBC>
__device__ __forceinline__ void k2y_body(const InterpArgs& a, int64_t wave,
                                         const void* __restrict__ grid_,
                                         void* __restrict__ y_, int c0) {
  typedef typename cplx<T>::t C;
  constexpr int LPS = JLC > 0 ? JLC : 16;
  constexpr int SPW = 64 / LPS;
  const int lane = threadIdx.x & 63;
  const int sl = lane / LPS;
  const int jl = lane - sl * LPS;
  const int JL = a.J[D - 1];
  const int64_t r = wave * SPW + sl;
  const bool valid = sl < SPW && jl < JL && r < a.M;
  const C* __restrict__ grid = (const C*)grid_;
  const int nb = a.nb;

  C acc[NBC];
#pragma unroll
  for (int c = 0; c < NBC; ++c) acc[c] = mk<T>((T)0, (T)0);

  if (valid) {
    const int J0 = a.J[0];
    const int J1 = D > 1 ? a.J[1] : 1;
    if (CSRV) {
      const int G = D == 1 ? 1 : (D == 2 ? J0 : J0 * J1);
      const C* __restrict__ val = (const C*)a.csr_val;
      const int* __restrict__ col = a.csr_col;
      const int64_t i0 = csr_index(r, 0, jl, G, JL, SPW);
      const int64_t gs = (int64_t)SPW * JL;
      for (int g = 0; g < G; ++g) {
        const C w = val[i0 + g * gs];
        const int64_t cell = col[i0 + g * gs];
#pragma unroll
        for (int c = 0; c < NBC; ++c)
          acc[c] = cadd(acc[c], cmul(w, grid[cell * nb + c0 + c]));
      }
    } else {
      const C* __restrict__ u = (const C*)a.u;
      const C ul = u[a.u_off[D - 1] + r * JL + jl];
      const int bl = a.base[(D - 1) * a.M + r];
      const int64_t cl = wrap_add(bl, jl, a.K[D - 1]);
      if (D == 1) {
#pragma unroll
        for (int c = 0; c < NBC; ++c)
          acc[c] = cmul(ul, grid[cl * nb + c0 + c]);
      } else {
        const int b0 = a.base[r];
        const C* u0 = u + a.u_off[0] + r * J0;
        if (D == 2) {
          for (int j0 = 0; j0 < J0; ++j0) {
            const int64_t cell = (int64_t)wrap_add(b0, j0, a.K[0]) * a.Kstride[0] + cl;
            const C w = cmul(u0[j0], ul);
#pragma unroll
            for (int c = 0; c < NBC; ++c)
              acc[c] = cadd(acc[c], cmul(w, grid[cell * nb + c0 + c]));
          }
        } else {
          const int b1 = a.base[a.M + r];
          const C* u1 = u + a.u_off[1] + r * J1;
          for (int j0 = 0; j0 < J0; ++j0) {
            const int64_t i0 = (int64_t)wrap_add(b0, j0, a.K[0]) * a.Kstride[0] + cl;
            const C w0 = u0[j0];
            for (int j1 = 0; j1 < J1; ++j1) {
              const int64_t cell = i0 + (int64_t)wrap_add(b1, j1, a.K[1]) * a.Kstride[1];
              const C w = cmul(cmul(w0, u1[j1]), ul);
#pragma unroll
              for (int c = 0; c < NBC; ++c)
                acc[c] = cadd(acc[c], cmul(w, grid[cell * nb + c0 + c]));
            }
          }
        }
      }
    }
  }
  // segmented reduction over the LPS lanes of each sample (lane jl == 0 ends
  // with the sum of lanes jl < JL of its group)
#pragma unroll
  for (int off = 1; off < LPS; off <<= 1) {
#pragma unroll
    for (int c = 0; c < NBC; ++c) {
      const C o = shfl_down_c(acc[c], off);
      if (jl + off < LPS) acc[c] = cadd(acc[c], o);
    }
  }
  if (valid && jl == 0) {
    const int64_t m = a.perm[r];
    C* __restrict__ y = (C*)y_;
#pragma unroll
    for (int c = 0; c < NBC; ++c) y[m * nb + c0 + c] = acc[c];
  }
}


template <typename T, int D, int JLC, bool CSRV, int NBC>
__global__ void __launch_bounds__(256)
k2y_kernel(InterpArgs a, const void* __restrict__ grid_, void* __restrict__ y_,
           int c0, const int* __restrict__ skip) {
  if (skip && *skip) return;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  k2y_body<T, D, JLC, CSRV, NBC>(a, wave, grid_, y_, c0);
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
// Support offsets j in [0, J) whose cell (b + j) mod K lies in the tile
// [T0, T0 + E).  Returns 1 and one interval [lo, lo + n) when the hit cells
// are one contiguous run (the common case), 0 when they are two runs (a
// wrapped support meeting both ends of a tile that spans or nearly spans the
// axis).
__device__ __forceinline__ int clip_axis(int b, int J, int K, int T0, int E,
                                         int& lo, int& n) {
  if (E == K) {
    // the tile spans the axis: the hit cells are contiguous from b unless
    // the support wraps past K - 1
    lo = 0; n = J;
    return b + J <= K ? 1 : 0;
  }
  const int T1 = T0 + E;
  const int l1 = max(0, T0 - b), h1 = min(min(J, K - b), T1 - b);
  const int l2 = max(K - b, T0 - b + K), h2 = min(J, T1 - b + K);
  const bool a1 = h1 > l1, a2 = h2 > l2;
  if (a1 && a2) { lo = 0; n = 0; return 0; }
  if (a1) { lo = l1; n = h1 - l1; }
  else if (a2) { lo = l2; n = h2 - l2; }
  else { lo = 0; n = 0; }
  return 1;
}

// q -> (q / n, q % n) for q < 64, n <= 64 via a 16-bit reciprocal
__device__ __forceinline__ void divmod_small(int q, int n, int magic, int& d, int& m) {
  d = (q * magic) >> 16;
  m = q - d * n;
}

// Gridding of one work item (tile id, entry start, entry count, slot) by one
// wavefront with its own LDS accumulator.
// Entries are taken in batches: lanes load one entry's row, sample value and
// first cells each (and stage its per-axis factors in LDS); the entries of
// the batch are then applied one after another, lanes spread over the
// support cells that fall inside the tile.
template <typename T, int D, bool CSRV, int NBC>
__device__ __forceinline__ void y2k_body(const InterpArgs& a, const BinArgs& b,
                                         const int4 it, unsigned char* wave_lds,
                                         const void* __restrict__ y_,
                                         void* __restrict__ grid_,
                                         void* __restrict__ part_, int c0) {
  typedef typename cplx<T>::t C;
  const int lane = threadIdx.x & 63;
  const int cells = b.tile_cells;      // compact tile (partials, output)
  const int acells = b.acc_cells;      // LDS accumulator (padded rows)
  const int J0 = a.J[0];
  const int J1 = D > 1 ? a.J[1] : 1;
  const int J2 = D > 2 ? a.J[2] : 1;
  const int SJ = J0 + (D > 1 ? J1 : 0) + (D > 2 ? J2 : 0);
  T* acc = reinterpret_cast<T*>(wave_lds);
  C* stage = reinterpret_cast<C*>(acc + (size_t)acells * 2 * NBC);
  int* srow = reinterpret_cast<int*>(stage + (size_t)b.batch * SJ);
  const C* __restrict__ y = (const C*)y_;
  const C* __restrict__ u = (const C*)a.u;
  const int nb = a.nb;

  int tid = it.x;
  int T0[PNH_MAXD], E[PNH_MAXD];
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    const int tg = b.tgrid[d];
    const int td = tid % tg;
    tid /= tg;
    T0[d] = td * b.tile[d];
    E[d] = min(b.tile[d], a.K[d] - T0[d]);
  }
  const int E1 = D > 1 ? E[1] : 1;
  const int E2 = D > 2 ? E[2] : 1;
  // LDS strides: the last axis is contiguous, rows are padded by one cell so
  // that a support box's rows fall on different banks
  const int S2 = 1;
  const int S1 = D == 3 ? E2 + 1 : 1;
  const int S0 = D == 3 ? E1 * (E2 + 1) : (D == 2 ? E1 + 1 : 1);

  for (int i = lane; i < acells * 2 * NBC; i += 64) acc[i] = (T)0;

  const int JL = a.J[D - 1];
  const int G = D == 1 ? 1 : (D == 2 ? J0 : J0 * J1);
  const int LPS = pnh_lanes_per_sample(JL);
  const int SPW = 64 / LPS;
  const int* __restrict__ rows = b.bin_rows + it.y;
  const int cnt = it.z;
  const int BATCH = b.batch;

  for (int e0 = 0; e0 < cnt; e0 += BATCH) {
    const int ne = min(BATCH, cnt - e0);
    // ---- lane-parallel setup: lane l prepares entry e0 + l ----
    int my_r = 0;
    int my_b[PNH_MAXD] = {0, 0, 0};
    C my_y[NBC];
#pragma unroll
    for (int c = 0; c < NBC; ++c) my_y[c] = mk<T>((T)0, (T)0);
    int my_geo = 0, my_start = 0, my_m1 = 0, my_m2 = 0, my_tot = 0;
    if (lane < ne) {
      my_r = rows[e0 + lane];
      const int64_t m = a.perm[my_r];
#pragma unroll
      for (int d = 0; d < D; ++d) my_b[d] = a.base[d * a.M + my_r];
#pragma unroll
      for (int c = 0; c < NBC; ++c) my_y[c] = y[m * nb + c0 + c];
      // clip the support box to the tile: per axis first offset lo, count n
      // and tile-local first cell s; flag 1 << 30 when some axis splits
      int lo[PNH_MAXD] = {0, 0, 0}, n[PNH_MAXD] = {1, 1, 1}, st0[PNH_MAXD] = {0, 0, 0};
      int single = 1;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        single &= clip_axis(my_b[d], a.J[d], a.K[d], T0[d], E[d], lo[d], n[d]);
        st0[d] = wrap_add(my_b[d], lo[d], a.K[d]) - T0[d];
      }
      my_tot = single ? n[0] * n[1] * n[2] : -1;
      my_geo = lo[0] | (lo[1] << 5) | (lo[2] << 10) | (n[1] << 20) | (n[2] << 25);
      my_start = st0[0] * S0 + st0[1] * S1 + st0[2] * S2;
      my_m2 = (65536 + n[2] - 1) / max(n[2], 1);
      my_m1 = (65536 + n[1] - 1) / max(n[1], 1);
    }
    if (!CSRV) {
      // rows of the batch through LDS (a lane may read any entry's row)
      if (lane < ne) srow[lane] = my_r;
      for (int idx = lane; idx < ne * SJ; idx += 64) {
        const int ei = idx / SJ;
        const int q = idx - ei * SJ;
        const int64_t re = srow[ei];
        int64_t src;
        if (q < J0) src = a.u_off[0] + re * J0 + q;
        else if (D > 1 && q < J0 + J1) src = a.u_off[1] + re * J1 + (q - J0);
        else src = a.u_off[2] + re * J2 + (q - J0 - J1);
        stage[idx] = u[src];
      }
    }
    // ---- apply the entries in order ----
    for (int i = 0; i < ne; ++i) {
      const int tot = rl_i(my_tot, i);
      C yv[NBC];
#pragma unroll
      for (int c = 0; c < NBC; ++c) yv[c] = rl_c(my_y[c], i);
      const C* st = stage + i * SJ;
      const int r = rl_i(my_r, i);
      if (tot >= 0) {
        const int geo = rl_i(my_geo, i);
        const int start = rl_i(my_start, i);
        const int m1 = rl_i(my_m1, i), m2 = rl_i(my_m2, i);
        const int lo0 = geo & 31, lo1 = (geo >> 5) & 31, lo2 = (geo >> 10) & 31;
        const int n1 = (geo >> 20) & 31, n2 = (geo >> 25) & 31;
        for (int q = lane; q < tot; q += 64) {
          int t0, t1, t2;
          if (tot <= 256) {
            // exact for q < 256 and n <= 16 (16-bit reciprocals)
            const int t = (q * m2) >> 16;
            t2 = q - t * n2;
            t0 = (t * m1) >> 16;
            t1 = t - t0 * n1;
          } else {
            t2 = q % n2;
            const int t = q / n2;
            t1 = t % n1;
            t0 = t / n1;
          }
          const int cell = start + t0 * S0 + t1 * S1 + t2 * S2;
          C w;
          if (CSRV) {
            const int j0 = lo0 + t0, j1 = lo1 + t1, j2 = lo2 + t2;
            const int g = D == 1 ? 0 : (D == 2 ? j0 : j0 * J1 + j1);
            const int jlast = D == 1 ? j0 : (D == 2 ? j1 : j2);
            w = ((const C*)a.csr_val)[csr_index(r, g, jlast, G, JL, SPW)];
          } else {
            w = st[lo0 + t0];
            if (D > 1) w = cmul(w, st[J0 + lo1 + t1]);
            if (D > 2) w = cmul(w, st[J0 + J1 + lo2 + t2]);
          }
#pragma unroll
          for (int c = 0; c < NBC; ++c) {
            const C v = cmulc(w, yv[c]);
            atomicAdd(&acc[(c * 2 + 0) * acells + cell], v.x);
            atomicAdd(&acc[(c * 2 + 1) * acells + cell], v.y);
          }
        }
      } else {
        // a wrapped support meets both ends of this tile: masked walk over
        // the whole support box
        int bb[PNH_MAXD];
#pragma unroll
        for (int d = 0; d < D; ++d) bb[d] = rl_i(my_b[d], i);
        const int P = J0 * J1 * J2;
        const int J12 = J1 * J2;
        for (int q = lane; q < P; q += 64) {
          const int j0 = q / J12;
          const int rem = q - j0 * J12;
          const int j1 = D > 1 ? rem / J2 : 0;
          const int j2 = D > 2 ? rem - j1 * J2 : 0;
          const int l0 = wrap_add(bb[0], j0, a.K[0]) - T0[0];
          if ((unsigned)l0 >= (unsigned)E[0]) continue;
          int l1 = 0, l2 = 0;
          if (D > 1) {
            l1 = wrap_add(bb[1], j1, a.K[1]) - T0[1];
            if ((unsigned)l1 >= (unsigned)E1) continue;
          }
          if (D > 2) {
            l2 = wrap_add(bb[2], j2, a.K[2]) - T0[2];
            if ((unsigned)l2 >= (unsigned)E2) continue;
          }
          C w;
          if (CSRV) {
            const int g = D == 1 ? 0 : (D == 2 ? j0 : j0 * J1 + j1);
            const int jlast = D == 1 ? j0 : (D == 2 ? j1 : j2);
            w = ((const C*)a.csr_val)[csr_index(r, g, jlast, G, JL, SPW)];
          } else {
            w = st[j0];
            if (D > 1) w = cmul(w, st[J0 + j1]);
            if (D > 2) w = cmul(w, st[J0 + J1 + j2]);
          }
          const int cell = l0 * S0 + l1 * S1 + l2 * S2;
#pragma unroll
          for (int c = 0; c < NBC; ++c) {
            const C v = cmulc(w, yv[c]);
            atomicAdd(&acc[(c * 2 + 0) * acells + cell], v.x);
            atomicAdd(&acc[(c * 2 + 1) * acells + cell], v.y);
          }
        }
      }
    }
  }

  // write the tile (direct) or the partial tile (split item)
  const int ncell = E[0] * E1 * E2;
  C* __restrict__ grid = (C*)grid_;
  C* __restrict__ part = (C*)part_;
  for (int cell = lane; cell < ncell; cell += 64) {
    const int l2 = D > 2 ? cell % E2 : 0;
    const int t = D > 2 ? cell / E2 : cell;
    const int l1 = D > 1 ? t % E1 : 0;
    const int l0 = D > 1 ? t / E1 : t;
    int64_t g = (int64_t)(T0[0] + l0) * a.Kstride[0];
    if (D > 1) g += (int64_t)(T0[1] + l1) * a.Kstride[1];
    if (D > 2) g += (int64_t)(T0[2] + l2);
    const int ac = l0 * S0 + l1 * S1 + l2 * S2;
#pragma unroll
    for (int c = 0; c < NBC; ++c) {
      const C v = mk<T>(acc[(c * 2 + 0) * acells + ac], acc[(c * 2 + 1) * acells + ac]);
      if (it.w < 0) grid[g * nb + c0 + c] = v;
      else part[((int64_t)it.w * cells + cell) * NBC + c] = v;
    }
  }
}

// One wavefront per work item (tile id, entry start, entry count, slot).
template <typename T, int D, bool CSRV, int NBC>
__global__ void __launch_bounds__(256)
y2k_binned_kernel(InterpArgs a, BinArgs b, const void* __restrict__ y_,
                  void* __restrict__ grid_, void* __restrict__ part_, int c0,
                  const int* __restrict__ skip) {
  if (skip && *skip) return;
  extern __shared__ unsigned char smem_raw[];
  const int wave = threadIdx.x >> 6;
  const int64_t item = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
  if (item >= b.n_items) return;  // whole wave leaves together
  const int4 it = reinterpret_cast<const int4*>(b.items)[item];
  y2k_body<T, D, CSRV, NBC>(a, b, it, smem_raw + (size_t)wave * b.wave_lds_bytes,
                            y_, grid_, part_, c0);
}

// Sum the partial tiles of every split tile in chunk order.
template <typename T, int D, int NBC>
__global__ void __launch_bounds__(256)
y2k_merge_kernel(InterpArgs a, BinArgs b, const void* __restrict__ part_,
                 void* __restrict__ grid_, int c0, const int* __restrict__ skip) {
  typedef typename cplx<T>::t C;
  if (skip && *skip) return;
  const int cells = b.tile_cells;
  const C* __restrict__ part = (const C*)part_;
  C* __restrict__ grid = (C*)grid_;
  for (int64_t mi = blockIdx.y; mi < b.n_merge; mi += gridDim.y) {
    const int3 me = reinterpret_cast<const int3*>(b.merge)[mi];
    int tid = me.x;
    int T0[PNH_MAXD], E[PNH_MAXD];
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      const int tg = b.tgrid[d];
      const int td = tid % tg;
      tid /= tg;
      T0[d] = td * b.tile[d];
      E[d] = min(b.tile[d], a.K[d] - T0[d]);
    }
    const int E1 = D > 1 ? E[1] : 1;
    const int E2 = D > 2 ? E[2] : 1;
    const int ncell = E[0] * E1 * E2;
    for (int cell = blockIdx.x * blockDim.x + threadIdx.x; cell < ncell;
         cell += gridDim.x * blockDim.x) {
      const int l2 = D > 2 ? cell % E2 : 0;
      const int t = D > 2 ? cell / E2 : cell;
      const int l1 = D > 1 ? t % E1 : 0;
      const int l0 = D > 1 ? t / E1 : t;
      int64_t g = (int64_t)(T0[0] + l0) * a.Kstride[0];
      if (D > 1) g += (int64_t)(T0[1] + l1) * a.Kstride[1];
      if (D > 2) g += (int64_t)(T0[2] + l2);
#pragma unroll
      for (int c = 0; c < NBC; ++c) {
        C s = part[((int64_t)me.y * cells + cell) * NBC + c];
        for (int k = 1; k < me.z; ++k)
          s = cadd(s, part[((int64_t)(me.y + k) * cells + cell) * NBC + c]);
        grid[g * a.nb + c0 + c] = s;
      }
    }
  }
}

template <typename T, int D, int LPS, bool CSRV, int NBC>
int launch_k2y_t(const InterpArgs& a, const void* grid, void* y, int c0,
                 const int* skip, hipStream_t s) {
  constexpr int SPW = 64 / LPS;
  const int64_t waves = (a.M + SPW - 1) / SPW;
  const int64_t nblk = (waves + 3) / 4;
  hipLaunchKernelGGL((k2y_kernel<T, D, LPS, CSRV, NBC>), dim3((unsigned)nblk),
                     dim3(256), 0, s, a, grid, y, c0, skip);
  PNH_LAUNCH_CHECK();
  return 0;
}

template <typename T, int D, bool CSRV, int NBC>
int launch_k2y_j(const InterpArgs& a, const void* grid, void* y, int c0,
                 const int* skip, hipStream_t s) {
  switch (pnh_lanes_per_sample(a.J[D - 1])) {
    case 1: return launch_k2y_t<T, D, 1, CSRV, NBC>(a, grid, y, c0, skip, s);
    case 2: return launch_k2y_t<T, D, 2, CSRV, NBC>(a, grid, y, c0, skip, s);
    case 3: return launch_k2y_t<T, D, 3, CSRV, NBC>(a, grid, y, c0, skip, s);
    case 4: return launch_k2y_t<T, D, 4, CSRV, NBC>(a, grid, y, c0, skip, s);
    case 6: return launch_k2y_t<T, D, 6, CSRV, NBC>(a, grid, y, c0, skip, s);
    case 8: return launch_k2y_t<T, D, 8, CSRV, NBC>(a, grid, y, c0, skip, s);
    case 16: return launch_k2y_t<T, D, 16, CSRV, NBC>(a, grid, y, c0, skip, s);
    default:
      pnh_set_error("interpolator size %d on the last axis is above 16", a.J[D - 1]);
      return -2;
  }
}

template <typename T, int D, bool CSRV>
int launch_k2y_nb(const InterpArgs& a, const void* grid, void* y,
                  const int* skip, hipStream_t s) {
  int c0 = 0;
  while (c0 < a.nb) {
    int rc;
    if (a.nb - c0 >= 4) { rc = launch_k2y_j<T, D, CSRV, 4>(a, grid, y, c0, skip, s); c0 += 4; }
    else { rc = launch_k2y_j<T, D, CSRV, 1>(a, grid, y, c0, skip, s); c0 += 1; }
    if (rc) return rc;
  }
  return 0;
}

template <typename T, int D, bool CSRV>
int launch_y2k_nb(const InterpArgs& a, const BinArgs& b0, const void* y,
                  void* grid, void* part, const int* skip, hipStream_t s) {
  typedef typename cplx<T>::t C;
  int SJ = 0;
  for (int d = 0; d < D; ++d) SJ += a.J[d];
  int c0 = 0;
  while (c0 < a.nb) {
    const int left = a.nb - c0;
    int nbc = left >= 4 ? 4 : (left >= 2 ? 2 : 1);
    BinArgs b = b0;
    // per-wave LDS: accumulators (2 * nbc * cells reals) + factor staging
    // (batch * SJ complex) + batch row ids; shrink until one wave fits
    int batch = 32;
    size_t wl = 0;
    for (;;) {
      wl = (size_t)b.acc_cells * 2 * nbc * sizeof(T) + (size_t)batch * SJ * sizeof(C) +
           (size_t)batch * sizeof(int);
      wl = (wl + 15) & ~(size_t)15;
      if (wl <= b.lds_budget) break;
      if (nbc > 1) nbc >>= 1;
      else if (batch > 4) batch >>= 1;
      else {
        pnh_set_error("y2k tile needs %zu bytes of LDS (budget %zu)", wl, b.lds_budget);
        return -2;
      }
    }
    int waves = 4;
    while (waves > 1 && (size_t)waves * wl > b.lds_budget) --waves;
    b.batch = batch;
    b.wave_lds_bytes = (int)wl;
    b.waves_per_block = waves;
    const size_t lds = (size_t)waves * wl;
    const int64_t nblk = (b.n_items + waves - 1) / waves;
    dim3 grid_dim((unsigned)nblk), blk(64 * waves);
    switch (nbc) {
      case 4: hipLaunchKernelGGL((y2k_binned_kernel<T, D, CSRV, 4>), grid_dim, blk, lds, s, a, b, y, grid, part, c0, skip); break;
      case 2: hipLaunchKernelGGL((y2k_binned_kernel<T, D, CSRV, 2>), grid_dim, blk, lds, s, a, b, y, grid, part, c0, skip); break;
      default: hipLaunchKernelGGL((y2k_binned_kernel<T, D, CSRV, 1>), grid_dim, blk, lds, s, a, b, y, grid, part, c0, skip); break;
    }
    PNH_LAUNCH_CHECK();
    if (b.n_merge > 0) {
      const unsigned gx = (unsigned)((b.tile_cells + 255) / 256);
      dim3 mg(gx, (unsigned)(b.n_merge < 65535 ? b.n_merge : 65535));
      switch (nbc) {
        case 4: hipLaunchKernelGGL((y2k_merge_kernel<T, D, 4>), mg, dim3(256), 0, s, a, b, part, grid, c0, skip); break;
        case 2: hipLaunchKernelGGL((y2k_merge_kernel<T, D, 2>), mg, dim3(256), 0, s, a, b, part, grid, c0, skip); break;
        default: hipLaunchKernelGGL((y2k_merge_kernel<T, D, 1>), mg, dim3(256), 0, s, a, b, part, grid, c0, skip); break;
      }
      PNH_LAUNCH_CHECK();
    }
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

int pnh_k2y_launch(const InterpArgs& a, int fp64, const void* grid, void* y,
                   const int* skip, hipStream_t s) {
  const bool csrv = a.csr_val != nullptr;
#define K2Y_CASE(T, D)                                                       \
  return csrv ? launch_k2y_nb<T, D, true>(a, grid, y, skip, s)               \
              : launch_k2y_nb<T, D, false>(a, grid, y, skip, s);
#define K2Y_CASE64(D) return launch_k2y_nb<double, D, false>(a, grid, y, skip, s);
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

template <typename T, int D, int LPS>
__global__ void __launch_bounds__(256)
k2y_batch_kernel(const BatchProblem* __restrict__ probs, BatchTable t,
                 const void* __restrict__ k_, void* __restrict__ y_,
                 const int* __restrict__ done, const int* __restrict__ all_done) {
  typedef typename cplx<T>::t C;
  if (all_done && *all_done) return;
  const int p = t.tab[blockIdx.x];
  if (done && done[p]) return;
  const int64_t lb = (int64_t)blockIdx.x - t.first[p];
  const InterpArgs a = sizeof(T) == 8 ? probs[p].ia64 : probs[p].ia;
  const int64_t wave = lb * (blockDim.x >> 6) + (threadIdx.x >> 6);
  k2y_body<T, D, LPS, false, 1>(a, wave, (const C*)k_ + probs[p].k_off,
                                (C*)y_ + probs[p].y_off, 0);
}

template <typename T, int D>
__global__ void __launch_bounds__(256)
y2k_batch_kernel(const BatchProblem* __restrict__ probs, const int2* __restrict__ items,
                 int64_t n_items, int wave_lds, int batch_entries,
                 const void* __restrict__ y_, void* __restrict__ k_, void* __restrict__ part_,
                 const int* __restrict__ done, const int* __restrict__ all_done) {
  typedef typename cplx<T>::t C;
  if (all_done && *all_done) return;
  extern __shared__ unsigned char smem_raw[];
  const int wave = threadIdx.x >> 6;
  const int64_t gi = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
  if (gi >= n_items) return;
  const int2 pi = items[gi];
  const int p = pi.x;
  if (done && done[p]) return;
  const InterpArgs a = sizeof(T) == 8 ? probs[p].ia64 : probs[p].ia;
  BinArgs b = probs[p].bins;
  b.batch = batch_entries;
  b.wave_lds_bytes = wave_lds;
  const int4 it = reinterpret_cast<const int4*>(b.items)[pi.y];
  y2k_body<T, D, false, 1>(a, b, it, smem_raw + (size_t)wave * wave_lds,
                           (const C*)y_ + probs[p].y_off, (C*)k_ + probs[p].k_off,
                           (C*)part_ + probs[p].part_off * b.tile_cells, 0);
}

template <typename T, int D>
__global__ void __launch_bounds__(256)
merge_batch_kernel(const BatchProblem* __restrict__ probs, const int2* __restrict__ merges,
                   int64_t n_merge, const void* __restrict__ part_, void* __restrict__ k_,
                   const int* __restrict__ done, const int* __restrict__ all_done) {
  typedef typename cplx<T>::t C;
  if (all_done && *all_done) return;
  for (int64_t g = blockIdx.y; g < n_merge; g += gridDim.y) {
    const int2 pm = merges[g];
    const int p = pm.x;
    if (done && done[p]) continue;
    const InterpArgs& a = probs[p].ia;
    const BinArgs& b = probs[p].bins;
    const int3 me = reinterpret_cast<const int3*>(b.merge)[pm.y];
    const int cells = b.tile_cells;
    int tid = me.x;
    int T0[PNH_MAXD], E[PNH_MAXD];
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      const int tg = b.tgrid[d];
      const int td = tid % tg;
      tid /= tg;
      T0[d] = td * b.tile[d];
      E[d] = min(b.tile[d], a.K[d] - T0[d]);
    }
    const int E1 = D > 1 ? E[1] : 1;
    const int E2 = D > 2 ? E[2] : 1;
    const int ncell = E[0] * E1 * E2;
    const C* part = (const C*)part_ + probs[p].part_off * cells;
    C* grid = (C*)k_ + probs[p].k_off;
    for (int cell = blockIdx.x * blockDim.x + threadIdx.x; cell < ncell;
         cell += gridDim.x * blockDim.x) {
      const int l2 = D > 2 ? cell % E2 : 0;
      const int t = D > 2 ? cell / E2 : cell;
      const int l1 = D > 1 ? t % E1 : 0;
      const int l0 = D > 1 ? t / E1 : t;
      int64_t gcell = (int64_t)(T0[0] + l0) * a.Kstride[0];
      if (D > 1) gcell += (int64_t)(T0[1] + l1) * a.Kstride[1];
      if (D > 2) gcell += (int64_t)(T0[2] + l2);
      C s = part[(int64_t)me.y * cells + cell];
      for (int k = 1; k < me.z; ++k) s = cadd(s, part[(int64_t)(me.y + k) * cells + cell]);
      grid[gcell] = s;
    }
  }
}

template <typename T, int D>
int batch_k2y_d(const BatchProblem* probs, const BatchTable& t, int J_last, const void* k,
                void* y, const int* done, const int* all_done, hipStream_t s) {
  dim3 g((unsigned)t.n_blocks), b(256);
  switch (pnh_lanes_per_sample(J_last)) {
    case 1: hipLaunchKernelGGL((k2y_batch_kernel<T, D, 1>), g, b, 0, s, probs, t, k, y, done, all_done); break;
    case 2: hipLaunchKernelGGL((k2y_batch_kernel<T, D, 2>), g, b, 0, s, probs, t, k, y, done, all_done); break;
    case 3: hipLaunchKernelGGL((k2y_batch_kernel<T, D, 3>), g, b, 0, s, probs, t, k, y, done, all_done); break;
    case 4: hipLaunchKernelGGL((k2y_batch_kernel<T, D, 4>), g, b, 0, s, probs, t, k, y, done, all_done); break;
    case 6: hipLaunchKernelGGL((k2y_batch_kernel<T, D, 6>), g, b, 0, s, probs, t, k, y, done, all_done); break;
    case 8: hipLaunchKernelGGL((k2y_batch_kernel<T, D, 8>), g, b, 0, s, probs, t, k, y, done, all_done); break;
    case 16: hipLaunchKernelGGL((k2y_batch_kernel<T, D, 16>), g, b, 0, s, probs, t, k, y, done, all_done); break;
    default: pnh_set_error("interpolator size above 16"); return -2;
  }
  PNH_LAUNCH_CHECK();
  return 0;
}

}  // namespace

int pnh_batch_k2y_launch(const BatchProblem* probs, const BatchTable& t, int D, int J_last,
                         int fp64, const void* k, void* y, const int* done,
                         const int* all_done, hipStream_t s) {
  if (t.n_blocks == 0) return 0;
  if (fp64) {
    switch (D) {
      case 1: return batch_k2y_d<double, 1>(probs, t, J_last, k, y, done, all_done, s);
      case 2: return batch_k2y_d<double, 2>(probs, t, J_last, k, y, done, all_done, s);
      case 3: return batch_k2y_d<double, 3>(probs, t, J_last, k, y, done, all_done, s);
    }
  } else {
    switch (D) {
      case 1: return batch_k2y_d<float, 1>(probs, t, J_last, k, y, done, all_done, s);
      case 2: return batch_k2y_d<float, 2>(probs, t, J_last, k, y, done, all_done, s);
      case 3: return batch_k2y_d<float, 3>(probs, t, J_last, k, y, done, all_done, s);
    }
  }
  pnh_set_error("unsupported ndim %d", D);
  return -2;
}

int pnh_batch_y2k_launch(const BatchProblem* probs, const int2* items, int64_t n_items,
                         const int2* merges, int64_t n_merge, int D, int fp64,
                         int tile_cells, int wave_lds, int batch_entries, int waves,
                         const void* y, void* k, void* part, const int* done,
                         const int* all_done, hipStream_t s) {
  const int64_t nblk = (n_items + waves - 1) / waves;
  const size_t lds = (size_t)waves * wave_lds;
  dim3 g((unsigned)nblk), b(64 * waves);
#define YB(T, DD) hipLaunchKernelGGL((y2k_batch_kernel<T, DD>), g, b, lds, s, probs, items, \
                                     n_items, wave_lds, batch_entries, y, k, part, done, all_done)
#define MB(T, DD) hipLaunchKernelGGL((merge_batch_kernel<T, DD>), mg, dim3(256), 0, s, probs, \
                                     merges, n_merge, part, k, done, all_done)
  if (nblk > 0) {
    if (fp64) { if (D == 1) YB(double, 1); else if (D == 2) YB(double, 2); else YB(double, 3); }
    else { if (D == 1) YB(float, 1); else if (D == 2) YB(float, 2); else YB(float, 3); }
    PNH_LAUNCH_CHECK();
  }
  if (n_merge > 0) {
    dim3 mg((unsigned)((tile_cells + 255) / 256), (unsigned)(n_merge < 65535 ? n_merge : 65535));
    if (fp64) { if (D == 1) MB(double, 1); else if (D == 2) MB(double, 2); else MB(double, 3); }
    else { if (D == 1) MB(float, 1); else if (D == 2) MB(float, 2); else MB(float, 3); }
    PNH_LAUNCH_CHECK();
  }
#undef YB
#undef MB
  return 0;
}
