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

template <int JC>
struct Jdims {
  __device__ __forceinline__ static int get(const int* J, int d) {
    return JC > 0 ? JC : J[d];
  }
};

__device__ __forceinline__ int wrap_add(int b, int j, int K) {
  int c = b + j;
  return c >= K ? c - K : c;
}

// ---------------------------------------------------------------- k2y ----
// y[perm[r]*nb + c0 + c] = sum_j coef(r, j) * grid[cell(r, j)*nb + c0 + c]
template <typename T, int D, int JC, bool CSRV, int NBC>
__global__ void __launch_bounds__(256)
k2y_kernel(InterpArgs a, const void* __restrict__ grid_, void* __restrict__ y_,
           int c0, const int* __restrict__ skip) {
  typedef typename cplx<T>::t C;
  if (skip && *skip) return;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.M) return;
  const C* __restrict__ grid = (const C*)grid_;
  C* __restrict__ y = (C*)y_;
  const C* __restrict__ u = (const C*)a.u;
  const int nb = a.nb;
  const int J0 = Jdims<JC>::get(a.J, 0);
  const int J1 = D > 1 ? Jdims<JC>::get(a.J, 1) : 1;
  const int J2 = D > 2 ? Jdims<JC>::get(a.J, 2) : 1;

  C acc[NBC];
#pragma unroll
  for (int c = 0; c < NBC; ++c) acc[c] = mk<T>((T)0, (T)0);

  if (CSRV) {
    const int P = J0 * J1 * J2;
    const int64_t slice = r >> 6, lane = r & 63;
    const C* __restrict__ val = (const C*)a.csr_val + (slice * P) * 64 + lane;
    const int* __restrict__ col = a.csr_col + (slice * P) * 64 + lane;
    for (int j = 0; j < P; ++j) {
      const C w = val[(int64_t)j * 64];
      const int64_t cell = col[(int64_t)j * 64];
#pragma unroll
      for (int c = 0; c < NBC; ++c)
        acc[c] = cadd(acc[c], cmul(w, grid[cell * nb + c0 + c]));
    }
  } else {
    const int b0 = a.base[r];
    const int b1 = D > 1 ? a.base[a.M + r] : 0;
    const int b2 = D > 2 ? a.base[2 * a.M + r] : 0;
    const C* u0 = u + a.u_off[0] + r * J0;
    const C* u1 = u + (D > 1 ? a.u_off[1] + r * J1 : 0);
    const C* u2 = u + (D > 2 ? a.u_off[2] + r * J2 : 0);
    // With a static J the last-dimension factors and wrapped cell offsets
    // are kept in registers; the product order stays ((u0*u1)*u2).
    constexpr int JR = JC > 0 ? JC : 1;
    C ul[JR];
    int64_t cl[JR];
    if (JC > 0 && D > 1) {
      const C* ulast = D == 2 ? u1 : u2;
      const int bl = D == 2 ? b1 : b2;
#pragma unroll
      for (int j = 0; j < JR; ++j) {
        ul[j] = ulast[j];
        cl[j] = wrap_add(bl, j, a.K[D - 1]);
      }
    }
    for (int j0 = 0; j0 < J0; ++j0) {
      const int64_t i0 = (int64_t)wrap_add(b0, j0, a.K[0]) * a.Kstride[0];
      const C w0 = u0[j0];
      if (D == 1) {
#pragma unroll
        for (int c = 0; c < NBC; ++c)
          acc[c] = cadd(acc[c], cmul(w0, grid[i0 * nb + c0 + c]));
        continue;
      }
      if (D == 2) {
        if (JC > 0) {
#pragma unroll
          for (int j1 = 0; j1 < JR; ++j1) {
            const C w = cmul(w0, ul[j1]);
            const int64_t cell = i0 + cl[j1];
#pragma unroll
            for (int c = 0; c < NBC; ++c)
              acc[c] = cadd(acc[c], cmul(w, grid[cell * nb + c0 + c]));
          }
        } else {
          for (int j1 = 0; j1 < J1; ++j1) {
            const C w = cmul(w0, u1[j1]);
            const int64_t cell = i0 + wrap_add(b1, j1, a.K[1]);
#pragma unroll
            for (int c = 0; c < NBC; ++c)
              acc[c] = cadd(acc[c], cmul(w, grid[cell * nb + c0 + c]));
          }
        }
        continue;
      }
      for (int j1 = 0; j1 < J1; ++j1) {
        const int64_t i1 = i0 + (int64_t)wrap_add(b1, j1, a.K[1]) * a.Kstride[1];
        const C w01 = cmul(w0, u1[j1]);
        if (JC > 0) {
#pragma unroll
          for (int j2 = 0; j2 < JR; ++j2) {
            const C w = cmul(w01, ul[j2]);
            const int64_t cell = i1 + cl[j2];
#pragma unroll
            for (int c = 0; c < NBC; ++c)
              acc[c] = cadd(acc[c], cmul(w, grid[cell * nb + c0 + c]));
          }
        } else {
          for (int j2 = 0; j2 < J2; ++j2) {
            const C w = cmul(w01, u2[j2]);
            const int64_t cell = i1 + wrap_add(b2, j2, a.K[2]);
#pragma unroll
            for (int c = 0; c < NBC; ++c)
              acc[c] = cadd(acc[c], cmul(w, grid[cell * nb + c0 + c]));
          }
        }
      }
    }
  }
  const int64_t m = a.perm[r];
#pragma unroll
  for (int c = 0; c < NBC; ++c) y[m * nb + c0 + c] = acc[c];
}

// ---------------------------------------------------------------- y2k ----
// One wavefront per work item (tile id, entry start, entry count, slot).
template <typename T, int D, bool CSRV, int NBC>
__global__ void __launch_bounds__(256)
y2k_binned_kernel(InterpArgs a, BinArgs b, const void* __restrict__ y_,
                  void* __restrict__ grid_, void* __restrict__ part_, int c0,
                  const int* __restrict__ skip) {
  typedef typename cplx<T>::t C;
  if (skip && *skip) return;
  extern __shared__ unsigned char smem_raw[];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t item = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
  if (item >= b.n_items) return;  // whole wave leaves together
  const int cells = b.tile_cells;
  T* acc = reinterpret_cast<T*>(smem_raw) + (size_t)wave * cells * 2 * NBC;
  const C* __restrict__ y = (const C*)y_;
  const C* __restrict__ u = (const C*)a.u;
  const int nb = a.nb;

  const int4 it = reinterpret_cast<const int4*>(b.items)[item];
  // tile origin and extent
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

  for (int i = lane; i < cells * 2 * NBC; i += 64) acc[i] = (T)0;

  const int J0 = a.J[0];
  const int J1 = D > 1 ? a.J[1] : 1;
  const int J2 = D > 2 ? a.J[2] : 1;
  const int P = J0 * J1 * J2;
  const int J12 = J1 * J2;

  const int* __restrict__ rows = b.bin_rows + it.y;
  const int cnt = it.z;
  for (int e = 0; e < cnt; ++e) {
    const int64_t r = rows[e];
    const int64_t m = a.perm[r];
    int bb[PNH_MAXD];
#pragma unroll
    for (int d = 0; d < D; ++d) bb[d] = a.base[d * a.M + r];
    C yv[NBC];
#pragma unroll
    for (int c = 0; c < NBC; ++c) yv[c] = y[m * nb + c0 + c];
    for (int q = lane; q < P; q += 64) {
      const int j0 = q / J12;
      const int rem = q - j0 * J12;
      const int j1 = D > 1 ? rem / J2 : 0;
      const int j2 = D > 2 ? rem - j1 * J2 : 0;
      // tile-local coordinates (unsigned compare rejects cells outside)
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
        const int64_t slice = r >> 6, ln = r & 63;
        w = ((const C*)a.csr_val)[(slice * P + q) * 64 + ln];
      } else {
        w = u[a.u_off[0] + r * J0 + j0];
        if (D > 1) w = cmul(w, u[a.u_off[1] + r * J1 + j1]);
        if (D > 2) w = cmul(w, u[a.u_off[2] + r * J2 + j2]);
      }
      const int cell = (l0 * E1 + l1) * E2 + l2;
#pragma unroll
      for (int c = 0; c < NBC; ++c) {
        const C v = cmulc(w, yv[c]);
        atomicAdd(&acc[(c * 2 + 0) * cells + cell], v.x);
        atomicAdd(&acc[(c * 2 + 1) * cells + cell], v.y);
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
#pragma unroll
    for (int c = 0; c < NBC; ++c) {
      const C v = mk<T>(acc[(c * 2 + 0) * cells + cell], acc[(c * 2 + 1) * cells + cell]);
      if (it.w < 0) grid[g * nb + c0 + c] = v;
      else part[((int64_t)it.w * cells + cell) * NBC + c] = v;
    }
  }
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

template <typename T, int D, int JC, bool CSRV, int NBC>
int launch_k2y_t(const InterpArgs& a, const void* grid, void* y, int c0,
                 const int* skip, hipStream_t s) {
  const int64_t nblk = (a.M + 255) / 256;
  hipLaunchKernelGGL((k2y_kernel<T, D, JC, CSRV, NBC>), dim3((unsigned)nblk),
                     dim3(256), 0, s, a, grid, y, c0, skip);
  PNH_LAUNCH_CHECK();
  return 0;
}

template <typename T, int D, bool CSRV, int NBC>
int launch_k2y_j(const InterpArgs& a, const void* grid, void* y, int c0,
                 const int* skip, hipStream_t s) {
  bool allj6 = true;
  for (int d = 0; d < D; ++d) allj6 = allj6 && a.J[d] == 6;
  if (!CSRV && allj6 && D > 1)
    return launch_k2y_t<T, D, 6, CSRV, NBC>(a, grid, y, c0, skip, s);
  return launch_k2y_t<T, D, 0, CSRV, NBC>(a, grid, y, c0, skip, s);
}

template <typename T, int D, bool CSRV>
int launch_k2y_nb(const InterpArgs& a, const void* grid, void* y,
                  const int* skip, hipStream_t s) {
  int c0 = 0;
  while (c0 < a.nb) {
    const int left = a.nb - c0;
    int rc;
    if (left >= 8) { rc = launch_k2y_j<T, D, CSRV, 8>(a, grid, y, c0, skip, s); c0 += 8; }
    else if (left >= 4) { rc = launch_k2y_j<T, D, CSRV, 4>(a, grid, y, c0, skip, s); c0 += 4; }
    else if (left >= 2) { rc = launch_k2y_j<T, D, CSRV, 2>(a, grid, y, c0, skip, s); c0 += 2; }
    else { rc = launch_k2y_j<T, D, CSRV, 1>(a, grid, y, c0, skip, s); c0 += 1; }
    if (rc) return rc;
  }
  return 0;
}

template <typename T, int D, bool CSRV>
int launch_y2k_nb(const InterpArgs& a, const BinArgs& b, const void* y,
                  void* grid, void* part, const int* skip, hipStream_t s) {
  const int waves = b.waves_per_block;
  const int64_t nblk = (b.n_items + waves - 1) / waves;
  int c0 = 0;
  while (c0 < a.nb) {
    // LDS per wave: tile_cells * 2 * NBC * sizeof(T); keep it within budget
    const int left = a.nb - c0;
    int nbc = left >= 4 ? 4 : (left >= 2 ? 2 : 1);
    while (nbc > 1 && (size_t)waves * b.tile_cells * 2 * nbc * sizeof(T) > b.lds_budget)
      nbc >>= 1;
    const size_t lds = (size_t)waves * b.tile_cells * 2 * nbc * sizeof(T);
    if (lds > b.lds_budget) {
      pnh_set_error("y2k tile needs %zu bytes of LDS (budget %zu)", lds, b.lds_budget);
      return -2;
    }
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

}  // namespace

int pnh_k2y_launch(const InterpArgs& a, int fp64, const void* grid, void* y,
                   const int* skip, hipStream_t s) {
  const bool csrv = a.csr_val != nullptr;
#define K2Y_CASE(T, D)                                                       \
  return csrv ? launch_k2y_nb<T, D, true>(a, grid, y, skip, s)               \
              : launch_k2y_nb<T, D, false>(a, grid, y, skip, s);
  if (fp64) {
    switch (a.D) { case 1: K2Y_CASE(double, 1) case 2: K2Y_CASE(double, 2) case 3: K2Y_CASE(double, 3) }
  } else {
    switch (a.D) { case 1: K2Y_CASE(float, 1) case 2: K2Y_CASE(float, 2) case 3: K2Y_CASE(float, 3) }
  }
#undef K2Y_CASE
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
  if (fp64) {
    switch (a.D) { case 1: Y2K_CASE(double, 1) case 2: Y2K_CASE(double, 2) case 3: Y2K_CASE(double, 3) }
  } else {
    switch (a.D) { case 1: Y2K_CASE(float, 1) case 2: Y2K_CASE(float, 2) case 3: Y2K_CASE(float, 3) }
  }
#undef Y2K_CASE
  pnh_set_error("unsupported ndim %d", a.D);
  return -2;
}
