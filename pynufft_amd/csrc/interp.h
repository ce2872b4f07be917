// Argument blocks for the interpolation / gridding kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "common.h"

struct InterpArgs {
  int D;                       // ndims (1..3)
  int nb;                      // coils sharing the trajectory
  int64_t M;                   // samples
  int K[PNH_MAXD];             // grid size per dim
  int64_t Kstride[PNH_MAXD];   // cell stride per dim (C order, in cells)
  int J[PNH_MAXD];             // support per dim
  const int* perm;             // [M] sorted row -> sample
  const int* base;             // [D][M] first support cell per dim
  const void* u;               // per-dim factors (complex, dtype of the op)
  int64_t u_off[PNH_MAXD];     // element offset of dim d's [M][J_d] block
  const void* du;              // d u_d / d om_d, same layout as u (or NULL)
  const void* csr_val;         // complex64 slice-interleaved table or NULL
  const int* csr_col;          // int32 slice-interleaved table or NULL
};

struct BinArgs {
  int tile[PNH_MAXD];          // tile extent per dim
  int tgrid[PNH_MAXD];         // tiles per dim
  int pad[PNH_MAXD];           // accumulator extent per dim: tile + J - 1, or K
                               // when one tile spans the axis (wraps inside)
  int64_t pad_cells;           // prod(pad): cells of one partial slot
  int64_t n_tiles;             // prod(tgrid)
  double u_bound;              // bound of |coefficient| over all rows and offsets
  int threads;                 // workgroup size of the gridding kernel
  int k2y_threads;             // workgroup size of the interpolation (k2y) kernel
  int max_tile_slots;          // largest slot count of any tile
  int64_t n_items;
  const int* items;            // [n_items][4] tile id, first row, rows, slot
  const int* tile_slots;       // [n_tiles][2] first slot, slot count
  // [n_items] or NULL: per work item, a bound of every cell's sum of
  // |coefficient| over the item's rows (complex64 factors).  With it the
  // complex64 gridding packs each cell's real and imaginary fixed-point
  // parts into one 64-bit accumulator (see y2k_item).
  const double* item_bound;
  // >= max over rows of the sum of |coefficient| over the row's support, or
  // 0.  With it the FP64 P^H P grids each work item in one pass over its
  // rows (the fixed-point exponent from the row bound and the input box).
  double row_bound;
};

// Workgroups per tile in the gridding gather pass: one, or one per 256
// cells when some tile has several slots (so that the slots of the heavy
// tiles of a dense trajectory region are not summed by one workgroup).
__host__ __device__ inline int pnh_gather_parts(const BinArgs& b) {
  if (b.max_tile_slots <= 1) return 1;
  int cells = 1;
  for (int d = 0; d < PNH_MAXD; ++d) cells *= b.tile[d] > 0 ? b.tile[d] : 1;
  return (cells + 255) / 256;
}

// Padded-box extent every split axis has with 16-cell tiles and a support
// of 6.  The gridding kernel and the one-pass FP64 P^H P have variants with
// it as a compile-time constant (box and accumulator offsets become
// instruction immediates, fewer registers).
constexpr int PNH_FIXED_PAD = 21;
// Whether the work list's every axis is split with padded extent
// PNH_FIXED_PAD (the box never wraps).
__host__ __device__ inline bool pnh_fixed_pad(const BinArgs& b, int D) {
  if (D < 2) return false;
  for (int d = 0; d < D; ++d)
    if (!(b.tgrid[d] > 1 && b.pad[d] == PNH_FIXED_PAD)) return false;
  return true;
}

// Lanes per sample used by k2y (and by the CSR table layout): one lane per
// support offset on the last axis, rounded to a divisor-friendly width.
__host__ __device__ inline int pnh_lanes_per_sample(int J_last) {
  if (J_last <= 4) return J_last;
  if (J_last == 6) return 6;
  if (J_last <= 8) return 8;
  if (J_last <= 16) return 16;
  return 0;
}

// y[perm[r]*nb + c] = sum_j coef(r, j) * grid[cell(r, j)*nb + c], by the
// gridding work list (tile ownership): each tile's padded box staged in LDS.
int pnh_k2y_launch(const InterpArgs& a, const BinArgs& b, int fp64, const void* grid, void* y,
                   const int* skip, hipStream_t s);
// out[m*D + d] (+)= sum_c Re(conj(g[m, c]) * sum_j d_d coef(m, j) * grid[cell(m, j), c])
int pnh_k2y_dom_launch(const InterpArgs& a, int fp64, const void* grid,
                       const void* g, void* out, hipStream_t s);
// Bytes at the head of the gridding scratch for the work items' fixed-point
// exponents (int32 each, rounded up to 256 B).
inline size_t pnh_item_exponent_bytes(int64_t n_items) {
  return ((size_t)n_items * 4 + 255) & ~(size_t)255;
}
// Work items from which the gridding computes every item's fixed-point
// exponent in a pass ahead, which also stores the samples in row order.
constexpr int64_t PNH_EXPONENT_PASS_ITEMS = 2048;
// Bytes after the exponents for those row-ordered samples (the coils of one
// gridding pass, at most 4, of a complex type of esz bytes), or 0 when the
// work list is too short for the exponent pass.
inline size_t pnh_row_samples_bytes(const BinArgs& b, int64_t M, int nb, size_t esz) {
  if (b.n_items < PNH_EXPONENT_PASS_ITEMS) return 0;
  const int nbc = nb >= 4 ? 4 : (nb >= 2 ? 2 : 1);
  return ((size_t)M * nbc * esz + 255) & ~(size_t)255;
}
// grid[cell*nb + c] = sum over rows r of conj(coef(r, j)) * y[perm[r]*nb + c]
// for every support cell of r; `part` holds pnh_item_exponent_bytes(n_items),
// then pnh_row_samples_bytes(...), then n_slots * pad_cells * min(nb, 4)
// elements of the operator's complex type.
int pnh_y2k_launch(const InterpArgs& a, const BinArgs& b, int fp64,
                   const void* y, void* grid, void* part, const int* skip,
                   hipStream_t s);
// grid = P^H P k (y2k of k2y): y is scratch of M * nb elements that receives
// P k (in row order when fused, else in sample order); one fused gridding
// pass per work item for one coil and per-axis factors, else k2y then y2k.
// `part` as for pnh_y2k_launch.  grid may alias k.
int pnh_k2y2k_launch(const InterpArgs& a, const BinArgs& b, int fp64, const void* k, void* y,
                     void* grid, void* part, const int* skip, hipStream_t s);
// bound[i] = max over the cells of item i's padded box of the sum of
// |u_0| |u_1| ... |u_last| (complex64 factors) over the item's rows, rounded
// up; one workgroup per item.  a: complex64 factors.
int pnh_item_bounds_launch(const InterpArgs& a, const BinArgs& b, double* bound, hipStream_t s);
// LDS bytes available to one workgroup on the current device.
size_t pnh_lds_limit();
// Coils gridded per pass for this tile and LDS limit (0 if none fits);
// packed: one accumulator word per cell and coil instead of two.
int pnh_y2k_coils_per_pass(const BinArgs& b, int nb, bool packed);
