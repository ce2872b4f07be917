// extern "C" entry points of libpynufft_hip.so (see include/pynufft_hip.h).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>

#include "../../include/pynufft_hip.h"
#include "common.h"
#include "fft.h"
#include "interp.h"
#include "solvers.h"
#include "vec.h"

#define PNH_VERSION "pynufft_hip 0.1.0 (gfx950)"

static thread_local char g_err[512];

void pnh_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

#define REQUIRE(cond, ...)            \
  do {                                \
    if (!(cond)) {                    \
      pnh_set_error(__VA_ARGS__);     \
      return -1;                      \
    }                                 \
  } while (0)

static const size_t kLdsBudget = 65536;      // FFT line buffers
static const int64_t kGridLdsMax = 163840;   // largest gridding accumulator + scratch

struct pnh_plan {
  pnh_plan_desc d;
  int D, nb;
  int64_t M, Ntot, Ktot;
  int64_t N[PNH_MAXD], K[PNH_MAXD];
  int J[PNH_MAXD];
  BinArgs bins;
  int slot_nbc;            // coils per partial slot element (min(nb, 4))
};

namespace {

inline size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

InterpArgs interp_args(const pnh_plan* p, int fp64) {
  InterpArgs a;
  memset(&a, 0, sizeof(a));
  a.D = p->D;
  a.nb = p->nb;
  a.M = p->M;
  int64_t stride = 1;
  for (int d = p->D - 1; d >= 0; --d) {
    a.K[d] = (int)p->K[d];
    a.Kstride[d] = stride;
    a.J[d] = p->J[d];
    stride *= p->K[d];
  }
  a.perm = p->d.perm;
  a.base = p->d.base;
  a.u = fp64 ? p->d.u_c128 : p->d.u_c64;
  a.du = fp64 ? nullptr : p->d.du_c64;
  for (int d = 0; d < PNH_MAXD; ++d) a.u_off[d] = p->d.u_off[d];
  // The stored coefficient table is complex64; FP64 operators always build
  // coefficients from the FP64 factors.
  a.csr_val = fp64 ? nullptr : p->d.csr_c64;
  a.csr_col = fp64 ? nullptr : p->d.csr_col;
  return a;
}

// gridding scratch: item exponents, the row-ordered samples of the exponent
// pass (long work lists), then the partial slots
size_t partial_bytes(const pnh_plan* p, int fp64) {
  const size_t esz = fp64 ? 16 : 8;
  return pnh_item_exponent_bytes(p->bins.n_items) +
         pnh_row_samples_bytes(p->bins, p->M, p->nb, esz) +
         (size_t)p->d.n_slots * p->bins.pad_cells * p->slot_nbc * esz;
}

// Build the line-FFT descriptor for axis `a`.
//  forward: reads N_a valid inputs (rest zero), writes K_a outputs;
//  inverse: reads K_a, writes the first N_a.
//  src_image / dst_image: that side is the Nd x nb image instead of the grid.
//  sn_side: 0 none, 1 multiply on load, 2 multiply on store.
int build_pass(const pnh_plan* p, int a, bool inverse, bool src_image,
               bool dst_image, int sn_side, int fp64, double scale, FftDesc* f) {
  memset(f, 0, sizeof(*f));
  const int D = p->D, nb = p->nb;
  int64_t kst[PNH_MAXD], nst[PNH_MAXD], sst[PNH_MAXD];
  int64_t sk = nb, sn_ = nb, ss = 1;
  for (int d = D - 1; d >= 0; --d) {
    kst[d] = sk; sk *= p->K[d];
    nst[d] = sn_; sn_ *= p->N[d];
    sst[d] = ss; ss *= p->N[d];
  }
  f->n = (int)p->K[a];
  f->n_in = inverse ? (int)p->K[a] : (int)p->N[a];
  f->n_out = inverse ? (int)p->N[a] : (int)p->K[a];
  f->sign = inverse ? 1 : -1;
  f->scale = scale;
  // outer dims: d < a with image extents
  int64_t n_outer = 1;
  for (int d = 0; d < a; ++d) n_outer *= p->N[d];
  f->n_outer = n_outer;
  f->o_split = (a >= 2) ? p->N[1] : 1;
  int64_t n_inner = nb;
  for (int d = a + 1; d < D; ++d) n_inner *= p->K[d];
  f->n_inner = n_inner;
  const int64_t* s_st = src_image ? nst : kst;
  const int64_t* d_st = dst_image ? nst : kst;
  f->src.t_stride = s_st[a];
  f->dst.t_stride = d_st[a];
  f->src.b_stride = 1;
  f->dst.b_stride = 1;
  f->src.o_stride[0] = a >= 1 ? s_st[0] : 0;
  f->src.o_stride[1] = a >= 2 ? s_st[1] : 0;
  f->dst.o_stride[0] = a >= 1 ? d_st[0] : 0;
  f->dst.o_stride[1] = a >= 2 ? d_st[1] : 0;
  f->sn_side = sn_side;
  f->sn = fp64 ? (const void*)p->d.sn_f64 : (const void*)p->d.sn_f32;
  if (sn_side && !f->sn) {
    pnh_set_error("apodization array missing for this precision");
    return -1;
  }
  f->sn_o_stride[0] = a >= 1 ? sst[0] : 0;
  f->sn_o_stride[1] = a >= 2 ? sst[1] : 0;
  f->sn_t_stride = sst[a];
  if (pnh_fft_factor(f->n, f->radix, f->ns, &f->npass)) return -2;
  const size_t esz = fp64 ? 16 : 8;
  const int n = f->n;
  // LDS per line: line buffers (none for one pass, one for two, two for
  // more) plus 24 B of per-line offsets
  const int nbuf = f->npass >= 3 ? 2 : (f->npass == 2 ? 1 : 0);
  const size_t line_bytes = (size_t)nbuf * pnh_fft_lds_line(n) * esz + 24;
  if (line_bytes > kLdsBudget) {
    pnh_set_error("FFT length %d exceeds the LDS-resident limit", n);
    return -2;
  }
  const int64_t n_lines = n_outer * n_inner;
  int LB;
  if (n_inner >= 8) {
    f->inner_major = 1;
    LB = 16;
    while (LB > 1 && (size_t)LB * line_bytes > kLdsBudget) LB >>= 1;
  } else {
    f->inner_major = 0;
    LB = (int)(kLdsBudget / line_bytes);
    if (LB > 16) LB = 16;
    if (LB < 1) LB = 1;
    // keep at least ~1024 workgroups when the lines allow it
    while (LB > 1 && (n_lines + LB - 1) / LB < 1024) LB >>= 1;
  }
  f->lines_per_block = LB;
  int rmin = f->radix[0];
  for (int q = 1; q < f->npass; ++q) rmin = f->radix[q] < rmin ? f->radix[q] : rmin;
  const int64_t work = (int64_t)LB * n / rmin;     // butterflies of the widest pass
  f->threads = work >= 256 ? 256 : (work >= 128 ? 128 : 64);
  return 0;
}

int run_passes_forward(const pnh_plan* p, const void* x_img, void* grid, bool apodize,
                       int fp64, const int* skip, hipStream_t s) {
  const int D = p->D;
  for (int a = D - 1; a >= 0; --a) {
    FftDesc f;
    const bool first = (a == D - 1);
    int rc = build_pass(p, a, false, first, false, (first && apodize) ? 1 : 0, fp64,
                        1.0, &f);
    if (rc) return rc;
    const void* tw = fp64 ? p->d.tw_c128[a] : p->d.tw_c64[a];
    if (!tw) { pnh_set_error("twiddle table missing for axis %d", a); return -1; }
    rc = pnh_fft_launch(f, fp64, first ? x_img : grid, grid, tw, skip, s);
    if (rc) return rc;
  }
  return 0;
}

int run_passes_inverse(const pnh_plan* p, void* grid, void* x_img, bool apodize,
                       int fp64, const int* skip, hipStream_t s) {
  const int D = p->D;
  for (int a = 0; a < D; ++a) {
    FftDesc f;
    const bool last = (a == D - 1);
    const double scale = last ? 1.0 / (double)p->Ktot : 1.0;
    int rc = build_pass(p, a, true, false, last, (last && apodize) ? 2 : 0, fp64,
                        scale, &f);
    if (rc) return rc;
    const void* tw = fp64 ? p->d.tw_c128[a] : p->d.tw_c64[a];
    if (!tw) { pnh_set_error("twiddle table missing for axis %d", a); return -1; }
    rc = pnh_fft_launch(f, fp64, grid, last ? x_img : grid, tw, skip, s);
    if (rc) return rc;
  }
  return 0;
}

int check_dtype_ready(const pnh_plan* p, int dtype) {
  if (dtype != PNH_C64 && dtype != PNH_C128) {
    pnh_set_error("dtype must be PNH_C64 or PNH_C128");
    return -1;
  }
  if (dtype == PNH_C128 && !p->d.u_c128) {
    pnh_set_error("plan has no complex128 factors (u_c128 is NULL)");
    return -1;
  }
  if (dtype == PNH_C64 && !p->d.u_c64 && !p->d.csr_c64) {
    pnh_set_error("plan has no complex64 interpolator");
    return -1;
  }
  return 0;
}

int do_y2k(const pnh_plan* p, int fp64, const void* y, void* k, void* part,
           const int* skip, hipStream_t s) {
  InterpArgs ia = interp_args(p, fp64);
  if (p->bins.n_items > 0 && !part) {
    pnh_set_error("workspace required for gridding");
    return -1;
  }
  return pnh_y2k_launch(ia, p->bins, fp64, y, k, part, skip, s);
}

int do_k2y(const pnh_plan* p, int fp64, const void* k, void* y, const int* skip,
           hipStream_t s) {
  InterpArgs ia = interp_args(p, fp64);
  return pnh_k2y_launch(ia, p->bins, fp64, k, y, skip, s);
}

// k_out = P^H P k, y (samples) receives P k; k_out may alias k
int do_k2y2k(const pnh_plan* p, int fp64, const void* k, void* y, void* k_out, void* part,
             const int* skip, hipStream_t s) {
  InterpArgs ia = interp_args(p, fp64);
  if (p->bins.n_items > 0 && !part) {
    pnh_set_error("workspace required for gridding");
    return -1;
  }
  return pnh_k2y2k_launch(ia, p->bins, fp64, k, y, k_out, part, skip, s);
}


// Validate a descriptor and fill a host-side plan (no device work).
int plan_init(const pnh_plan_desc& d, pnh_plan* p) {
  REQUIRE(d.ndim >= 1 && d.ndim <= PNH_MAXD, "ndim must be 1..%d (got %d)", PNH_MAXD, d.ndim);
  REQUIRE(d.batch >= 1, "batch must be >= 1");
  REQUIRE(d.M >= 0 && d.M < (1LL << 31), "M out of range");
  REQUIRE(d.perm && d.base, "perm/base pointers are required");
  REQUIRE(d.u_c64 || d.u_c128 || d.csr_c64, "no interpolator data");
  REQUIRE(!d.csr_c64 || d.csr_col, "csr_col required with csr_c64");
  REQUIRE(d.sn_f32, "sn_f32 is required");
  memset(p, 0, sizeof(*p));
  p->d = d;
  p->D = d.ndim;
  p->nb = d.batch;
  p->M = d.M;
  p->Ntot = 1;
  p->Ktot = 1;
  int64_t pad_cells = 1, n_tiles = 1;
  for (int a = 0; a < p->D; ++a) {
    p->N[a] = d.Nd[a];
    p->K[a] = d.Kd[a];
    p->J[a] = d.Jd[a];
    REQUIRE(d.Nd[a] >= 1 && d.Kd[a] >= d.Nd[a] && d.Kd[a] < (1LL << 30) && d.Jd[a] >= 1 &&
                d.Jd[a] <= d.Kd[a] && d.tile[a] >= 1,
            "bad geometry on axis %d: Nd=%lld Kd=%lld Jd=%d tile=%d", a,
            (long long)d.Nd[a], (long long)d.Kd[a], d.Jd[a], d.tile[a]);
    p->Ntot *= d.Nd[a];
    p->Ktot *= d.Kd[a];
    const int64_t tile = d.tile[a] < d.Kd[a] ? d.tile[a] : d.Kd[a];
    const int64_t tg = (d.Kd[a] + tile - 1) / tile;
    p->bins.tile[a] = (int)tile;
    p->bins.tgrid[a] = (int)tg;
    if (tg == 1) {
      p->bins.pad[a] = (int)d.Kd[a];
    } else {
      REQUIRE(tile >= d.Jd[a] - 1 && d.Kd[a] - (tg - 1) * tile >= d.Jd[a] - 1,
              "gridding tile %lld on axis %d leaves a tile narrower than Jd - 1 = %d "
              "(Kd = %lld)", (long long)tile, a, d.Jd[a] - 1, (long long)d.Kd[a]);
      p->bins.pad[a] = (int)(tile + d.Jd[a] - 1);
    }
    pad_cells *= p->bins.pad[a];
    n_tiles *= tg;
    int radix[PNH_FFT_MAX_PASS], ns[PNH_FFT_MAX_PASS], np;
    if (pnh_fft_factor((int)d.Kd[a], radix, ns, &np)) return -2;
    REQUIRE(d.tw_c64[a], "tw_c64[%d] missing", a);
  }
  REQUIRE(p->Ktot < (1LL << 31), "grids of 2^31 or more cells are not supported");
  REQUIRE(pnh_lanes_per_sample(d.Jd[d.ndim - 1]) > 0, "Jd above 16 on the last axis");
  for (int a = 0; a + 1 < p->D; ++a)
    REQUIRE(d.Jd[a] <= 16, "Jd above 16 on axis %d", a);
  if (pad_cells * 16 + 256 > kGridLdsMax) {
    pnh_set_error("gridding tile needs %lld cells of LDS accumulator (limit %lld)",
                  (long long)pad_cells, (long long)((kGridLdsMax - 256) / 16));
    return -2;
  }
  p->bins.pad_cells = pad_cells;
  p->bins.n_tiles = n_tiles;
  REQUIRE(d.n_tiles == n_tiles, "n_tiles is %lld, the tile grid has %lld tiles",
          (long long)d.n_tiles, (long long)n_tiles);
  REQUIRE(d.tile_slots, "tile_slots is required");
  REQUIRE(d.n_items >= 0 && (d.n_items == 0 || d.items), "gridding work items are required");
  REQUIRE(d.n_slots >= d.n_items, "n_slots below n_items");
  REQUIRE(d.M == 0 || (d.u_bound > 0.0 && d.u_bound < 1e300), "u_bound must be positive");
  REQUIRE(d.max_tile_slots >= 0, "max_tile_slots must be >= 0");
  p->bins.max_tile_slots = d.max_tile_slots;
  p->bins.u_bound = d.u_bound;
  p->bins.n_items = d.n_items;
  p->bins.items = d.items;
  p->bins.tile_slots = d.tile_slots;
  p->bins.item_bound = d.item_bound;
  REQUIRE(d.row_bound >= 0.0 && d.row_bound < 1e300, "row_bound must be >= 0");
  p->bins.row_bound = d.row_bound;
  if (d.threads == 0) {
    // gridding: by the complex64 accumulator, one 64-bit word per cell when
    // packed, else two.  Two packed boxes of 40-80 KiB share a CU's LDS
    // with 768 threads each: the 3-D 16^3 packed kernel's 80 registers
    // allow six waves per SIMD, which 2 x 768 threads fill (0.332 ms per
    // y2k against 0.345 at 512 and 0.382 at 1024, same bits).
    // Interpolation: one workgroup per CU on large boxes (3-D 16^3 k2y
    // measured 0.20 ms at 1024 threads, 0.26 ms at 512)
    const int64_t bytes = pad_cells * (d.item_bound ? 8 : 16);
    const int mid = d.item_bound ? 768 : 512;
    p->bins.threads = bytes > 81920 ? 1024 : (bytes > 40960 ? mid : 256);
    const int64_t kbytes = pad_cells * 16;
    p->bins.k2y_threads = kbytes > 81920 ? 1024 : (kbytes > 40960 ? 512 : 256);
  } else {
    REQUIRE(d.threads >= 64 && d.threads <= 1024 && d.threads % 64 == 0,
            "threads must be a multiple of 64 in 64..1024 (got %d)", d.threads);
    p->bins.threads = d.threads;
    p->bins.k2y_threads = d.threads;
  }
  p->slot_nbc = p->nb >= 4 ? 4 : (p->nb >= 2 ? 2 : 1);
  REQUIRE(d.fp64_ops == 0 || d.fp64_ops == 1, "fp64_ops must be 0 or 1");
  if (d.fp64_ops) {
    REQUIRE(d.u_c128 && d.sn_f64, "fp64_ops needs u_c128 and sn_f64");
    for (int a = 0; a < p->D; ++a) REQUIRE(d.tw_c128[a], "fp64_ops needs tw_c128[%d]", a);
  }
  return 0;
}

// Scratch of the complex64 operators.  FP32 plans: forward the grid; adjoint
// the grid and the gridding partials; selfadjoint the grid, the samples and
// the partials.  fp64_ops plans: complex128 image, grid and samples plus the
// FP64 partials (forward needs no partials).
size_t forward_ws_bytes(const pnh_plan* p) {
  const size_t g = align_up((size_t)p->Ktot * p->nb * (p->d.fp64_ops ? 16 : 8));
  if (!p->d.fp64_ops) return g;
  return g + align_up((size_t)p->Ntot * p->nb * 16) + align_up((size_t)p->M * p->nb * 16);
}

size_t adjoint_ws_bytes(const pnh_plan* p) {
  if (!p->d.fp64_ops)
    return align_up((size_t)p->Ktot * p->nb * 8) + align_up(partial_bytes(p, 0));
  return align_up((size_t)p->M * p->nb * 16) + align_up((size_t)p->Ktot * p->nb * 16) +
         align_up((size_t)p->Ntot * p->nb * 16) + align_up(partial_bytes(p, 1));
}

size_t selfadjoint_ws_bytes(const pnh_plan* p) {
  if (!p->d.fp64_ops)
    return align_up((size_t)p->Ktot * p->nb * 8) + align_up((size_t)p->M * p->nb * 8) +
           align_up(partial_bytes(p, 0));
  return align_up((size_t)p->Ntot * p->nb * 16) + align_up((size_t)p->Ktot * p->nb * 16) +
         align_up((size_t)p->M * p->nb * 16) + align_up(partial_bytes(p, 1));
}

// x (complex64) = xx2x(k2xx(y2k(v))) for complex128 samples v, with the
// reference's precision: FP64 gridding and inverse FFT, rounded to complex64
// before the apodization.
int adjoint_c128(const pnh_plan* p, const void* v, void* x, void* g64, void* img64,
                 void* part, hipStream_t s) {
  int rc;
  if ((rc = do_y2k(p, 1, v, g64, part, nullptr, s))) return rc;
  if ((rc = run_passes_inverse(p, g64, img64, false, 1, nullptr, s))) return rc;
  return pnh_cast_mul_sn(img64, x, p->d.sn_f32, p->Ntot, p->nb, s);
}

// The complex64 operators of an fp64_ops plan: the image is upcast, x * sn,
// the FFTs and the interpolation run in FP64, the samples are rounded to
// complex64, and the image is rounded to complex64 before the final
// apodization (as the reference's k2xx / xx2x do).
int forward_fp64(const pnh_plan* p, const void* x, void* y, void* ws, hipStream_t s) {
  char* w = (char*)ws;
  void* g64 = w; w += align_up((size_t)p->Ktot * p->nb * 16);
  void* x64 = w; w += align_up((size_t)p->Ntot * p->nb * 16);
  void* y64 = w;
  int rc;
  if ((rc = pnh_c64_to_c128(x, x64, p->Ntot * p->nb, s))) return rc;
  if ((rc = run_passes_forward(p, x64, g64, true, 1, nullptr, s))) return rc;
  if (p->M == 0) return 0;
  if ((rc = do_k2y(p, 1, g64, y64, nullptr, s))) return rc;
  return pnh_c128_to_c64(y64, y, p->M * p->nb, s);
}

int adjoint_fp64(const pnh_plan* p, const void* y, void* x, void* ws, hipStream_t s) {
  char* w = (char*)ws;
  void* y64 = w; w += align_up((size_t)p->M * p->nb * 16);
  void* g64 = w; w += align_up((size_t)p->Ktot * p->nb * 16);
  void* img64 = w; w += align_up((size_t)p->Ntot * p->nb * 16);
  void* part = w;
  int rc;
  if (p->M > 0 && (rc = pnh_c64_to_c128(y, y64, p->M * p->nb, s))) return rc;
  return adjoint_c128(p, y64, x, g64, img64, part, s);
}

int selfadjoint_fp64(const pnh_plan* p, const void* x, void* x_out, void* ws, hipStream_t s) {
  char* w = (char*)ws;
  void* img64 = w; w += align_up((size_t)p->Ntot * p->nb * 16);
  void* g64 = w; w += align_up((size_t)p->Ktot * p->nb * 16);
  void* y64 = w; w += align_up((size_t)p->M * p->nb * 16);
  void* part = w;
  int rc;
  if ((rc = pnh_c64_to_c128(x, img64, p->Ntot * p->nb, s))) return rc;
  if ((rc = run_passes_forward(p, img64, g64, true, 1, nullptr, s))) return rc;
  if ((rc = do_k2y2k(p, 1, g64, y64, g64, part, nullptr, s))) return rc;
  if ((rc = run_passes_inverse(p, g64, img64, false, 1, nullptr, s))) return rc;
  return pnh_cast_mul_sn(img64, x_out, p->d.sn_f32, p->Ntot, p->nb, s);
}

}  // namespace

extern "C" {

const char* pnh_version(void) { return PNH_VERSION; }

const char* pnh_last_error(void) { return g_err; }

int pnh_device_count(void) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    pnh_set_error("hipGetDeviceCount: %s", hipGetErrorString(e));
    return -3;
  }
  return n;
}

int pnh_plan_create(const pnh_plan_desc* desc, pnh_plan** out) {
  REQUIRE(desc && out, "null argument");
  *out = nullptr;
  pnh_plan* p = new (std::nothrow) pnh_plan;
  if (!p) { pnh_set_error("out of host memory"); return -1; }
  const int rc = plan_init(*desc, p);
  if (rc) { delete p; return rc; }
  *out = p;
  return 0;
}

int pnh_plan_destroy(pnh_plan* plan) {
  delete plan;
  return 0;
}

// Scratch of density(): ones [M nb], image [N nb], then the adjoint's
// scratch, which afterwards holds the complex64 grid of the forward FFT.
static size_t density_ws_bytes(const pnh_plan* p) {
  const size_t adj = adjoint_ws_bytes(p);
  const size_t grid = align_up((size_t)p->Ktot * p->nb * 8);
  return align_up((size_t)p->M * p->nb * 8) + align_up((size_t)p->Ntot * p->nb * 8) +
         (adj > grid ? adj : grid);
}

// Scratch of the L1TVOLS solve: uker and |density| (float [K]), images AHy,
// AHyk, rhs, q and dd/bb per axis, then the self-adjoint scratch, which also
// serves density(), the adjoint and the k-space solve's complex64 grid.
static size_t l1tv_ws_bytes(const pnh_plan* p) {
  const size_t img = align_up((size_t)p->Ntot * 8);
  size_t sa = selfadjoint_ws_bytes(p);
  const size_t dens = density_ws_bytes(p);
  const size_t adj = adjoint_ws_bytes(p);
  sa = sa > dens ? sa : dens;
  sa = sa > adj ? sa : adj;
  return align_up((size_t)p->Ktot * 8) + align_up((size_t)p->Ktot * 4) +
         (4 + 2 * (size_t)p->D) * img + sa;
}

// Scratch of the 'dc' solve: W, E and the weighted samples (complex128
// [M nb] each), a complex128 and a complex64 grid, the complex128 image, the
// FP64 gridding partials.
static size_t dc_ws_bytes(const pnh_plan* p) {
  const int64_t my = p->M * p->nb;
  return 3 * align_up((size_t)my * 16) + align_up((size_t)p->Ktot * p->nb * 16) +
         align_up((size_t)p->Ktot * p->nb * 8) + align_up((size_t)p->Ntot * p->nb * 16) +
         align_up(partial_bytes(p, 1));
}

// Scratch of the LSQR solve: u and A v (complex128 [M]), v, w, x and A^H u
// (complex128 [K]), FP64 gridding partials, the complex128 image, block
// partials and the scalar state.
static size_t lsqr_ws_bytes(const pnh_plan* p) {
  return 2 * align_up((size_t)p->M * 16) + 4 * align_up((size_t)p->Ktot * 16) +
         align_up(partial_bytes(p, 1)) + align_up((size_t)p->Ntot * 16) +
         align_up((size_t)PNH_LSQR_BLOCKS * 8) + align_up(sizeof(LsqrState));
}

int pnh_workspace_bytes(const pnh_plan* p, int op, int dtype, size_t* bytes) {
  REQUIRE(p && bytes, "null argument");
  const int fp64 = dtype == PNH_C128;
  const size_t esz = fp64 ? 16 : 8;
  const size_t grid = align_up((size_t)p->Ktot * p->nb * esz);
  const size_t ys = align_up((size_t)p->M * p->nb * esz);
  const size_t part = align_up(partial_bytes(p, fp64));
  switch (op) {
    case PNH_OP_FORWARD: *bytes = forward_ws_bytes(p); break;
    case PNH_OP_ADJOINT: *bytes = adjoint_ws_bytes(p); break;
    case PNH_OP_SELFADJOINT: *bytes = selfadjoint_ws_bytes(p); break;
    case PNH_OP_Y2K: *bytes = part; break;
    case PNH_OP_K2Y2K: *bytes = ys + part; break;
    case PNH_OP_FORWARD_ONE2MANY:
      *bytes = align_up((size_t)p->Ntot * p->nb * 8) + forward_ws_bytes(p);
      break;
    case PNH_OP_ADJOINT_MANY2ONE:
      *bytes = align_up((size_t)p->Ntot * p->nb * 8) + adjoint_ws_bytes(p);
      break;
    case PNH_OP_PRECOMPUTE_SP:
      *bytes = density_ws_bytes(p);
      break;
    case PNH_OP_SELFADJOINT2: *bytes = align_up((size_t)p->Ktot * p->nb * 8); break;
    case PNH_OP_SOLVE_DC: *bytes = dc_ws_bytes(p); break;
    case PNH_OP_SOLVE_L1TVOLS:
      *bytes = l1tv_ws_bytes(p);
      break;
    case PNH_OP_SOLVE_LSQR:
      *bytes = lsqr_ws_bytes(p);
      break;
    case PNH_OP_SOLVE_BICG:
      *bytes = align_up((size_t)p->M * 16) + 7 * align_up((size_t)p->Ktot * 16) +
               align_up(partial_bytes(p, 1)) + align_up((size_t)p->Ntot * 16) +
               align_up((size_t)2 * PNH_LSQR_BLOCKS * 16) + align_up(sizeof(KrylovState));
      break;
    case PNH_OP_SOLVE_LSMR:
      *bytes = lsqr_ws_bytes(p) + align_up((size_t)p->Ktot * 16) + align_up(sizeof(LsmrState));
      break;
    case PNH_OP_SOLVE_CG: {
      const size_t g2 = align_up((size_t)p->Ktot * p->nb * 16);
      const size_t y2 = align_up((size_t)p->M * p->nb * 16);
      const size_t part2 = align_up(partial_bytes(p, 1));
      const size_t red = align_up((size_t)PNH_RED_BLOCKS * p->nb * 16);
      *bytes = 4 * g2 + y2 + part2 + 2 * red + align_up(sizeof(CgScalars));
      break;
    }
    default: pnh_set_error("unknown op %d", op); return -1;
  }
  return 0;
}

int pnh_item_bounds(const pnh_plan* p, double* bound, void* stream) {
  REQUIRE(p && (bound || p->bins.n_items == 0), "null argument");
  REQUIRE(p->d.u_c64, "item bounds need the complex64 factors (u_c64)");
  return pnh_item_bounds_launch(interp_args(p, 0), p->bins, bound, (hipStream_t)stream);
}

int pnh_x2xx(const pnh_plan* p, const void* x, void* xx, void* stream) {
  REQUIRE(p && x && xx, "null argument");
  return pnh_scale_sn(0, x, xx, p->d.sn_f32, p->Ntot, p->nb, (hipStream_t)stream);
}

int pnh_xx2x(const pnh_plan* p, const void* xx, void* x, void* stream) {
  REQUIRE(p && x && xx, "null argument");
  return pnh_scale_sn(0, xx, x, p->d.sn_f32, p->Ntot, p->nb, (hipStream_t)stream);
}

int pnh_xx2k(const pnh_plan* p, const void* xx, void* k, void* stream) {
  REQUIRE(p && xx && k, "null argument");
  return run_passes_forward(p, xx, k, false, 0, nullptr, (hipStream_t)stream);
}

int pnh_k2xx(const pnh_plan* p, void* k, void* xx, void* stream) {
  REQUIRE(p && xx && k, "null argument");
  return run_passes_inverse(p, k, xx, false, 0, nullptr, (hipStream_t)stream);
}

int pnh_k2y(const pnh_plan* p, int dtype, const void* k, void* y, void* stream) {
  REQUIRE(p && k && (y || p->M == 0), "null argument");
  int rc = check_dtype_ready(p, dtype);
  if (rc) return rc;
  if (p->M == 0) return 0;
  return do_k2y(p, dtype == PNH_C128, k, y, nullptr, (hipStream_t)stream);
}

int pnh_y2k(const pnh_plan* p, int dtype, const void* y, void* k, void* ws, void* stream) {
  REQUIRE(p && k && (y || p->M == 0), "null argument");
  int rc = check_dtype_ready(p, dtype);
  if (rc) return rc;
  return do_y2k(p, dtype == PNH_C128, y, k, ws, nullptr, (hipStream_t)stream);
}

int pnh_k2y2k(const pnh_plan* p, int dtype, const void* k, void* k_out, void* ws,
              void* stream) {
  REQUIRE(p && k && k_out && ws, "null argument");
  int rc = check_dtype_ready(p, dtype);
  if (rc) return rc;
  const int fp64 = dtype == PNH_C128;
  const size_t esz = fp64 ? 16 : 8;
  char* w = (char*)ws;
  void* ytmp = w;
  void* part = w + align_up((size_t)p->M * p->nb * esz);
  hipStream_t s = (hipStream_t)stream;
  return do_k2y2k(p, fp64, k, ytmp, k_out, part, nullptr, s);
}

int pnh_forward(const pnh_plan* p, const void* x, void* y, void* ws, void* stream) {
  REQUIRE(p && x && ws && (y || p->M == 0), "null argument");
  int rc = check_dtype_ready(p, PNH_C64);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (p->d.fp64_ops) return forward_fp64(p, x, y, ws, s);
  if ((rc = run_passes_forward(p, x, ws, true, 0, nullptr, s))) return rc;
  if (p->M == 0) return 0;
  return do_k2y(p, 0, ws, y, nullptr, s);
}

int pnh_adjoint(const pnh_plan* p, const void* y, void* x, void* ws, void* stream) {
  REQUIRE(p && x && ws && (y || p->M == 0), "null argument");
  int rc = check_dtype_ready(p, PNH_C64);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (p->d.fp64_ops) return adjoint_fp64(p, y, x, ws, s);
  char* w = (char*)ws;
  void* grid = w;
  void* part = w + align_up((size_t)p->Ktot * p->nb * 8);
  if ((rc = do_y2k(p, 0, y, grid, part, nullptr, s))) return rc;
  return run_passes_inverse(p, grid, x, true, 0, nullptr, s);
}

int pnh_selfadjoint(const pnh_plan* p, const void* x, void* x_out, void* ws, void* stream) {
  REQUIRE(p && x && x_out && ws, "null argument");
  int rc = check_dtype_ready(p, PNH_C64);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (p->d.fp64_ops) return selfadjoint_fp64(p, x, x_out, ws, s);
  char* w = (char*)ws;
  void* grid = w;
  void* ytmp = w + align_up((size_t)p->Ktot * p->nb * 8);
  void* part = (char*)ytmp + align_up((size_t)p->M * p->nb * 8);
  if ((rc = run_passes_forward(p, x, grid, true, 0, nullptr, s))) return rc;
  if ((rc = do_k2y2k(p, 0, grid, ytmp, grid, part, nullptr, s))) return rc;
  return run_passes_inverse(p, grid, x_out, true, 0, nullptr, s);
}

int pnh_forward_one2many(const pnh_plan* p, const void* x, const void* sense, void* y,
                         void* ws, void* stream) {
  REQUIRE(p && x && sense && ws && (y || p->M == 0), "null argument");
  char* w = (char*)ws;
  void* stack = w;
  void* fws = w + align_up((size_t)p->Ntot * p->nb * 8);
  int rc = pnh_coil_expand(x, sense, stack, p->Ntot, p->nb, (hipStream_t)stream);
  if (rc) return rc;
  return pnh_forward(p, stack, y, fws, stream);
}

int pnh_adjoint_many2one(const pnh_plan* p, const void* y, const void* sense, void* x,
                         void* ws, void* stream) {
  REQUIRE(p && x && sense && ws && (y || p->M == 0), "null argument");
  char* w = (char*)ws;
  void* stack = w;
  void* aws = w + align_up((size_t)p->Ntot * p->nb * 8);
  int rc = pnh_adjoint(p, y, stack, aws, stream);
  if (rc) return rc;
  return pnh_coil_combine(stack, sense, x, p->Ntot, p->nb, (hipStream_t)stream);
}

// W[i] = |xx2k(adjoint(1))[i]| of the first coil (float [Ktot]).
static int density(const pnh_plan* p, float* W, char* w, hipStream_t s) {
  void* ones = w;
  void* img = w + align_up((size_t)p->M * p->nb * 8);
  char* aws = (char*)img + align_up((size_t)p->Ntot * p->nb * 8);   // adjoint scratch, then grid
  int rc;
  if ((rc = pnh_fill_c64(ones, p->M * p->nb, 1.f, 0.f, s))) return rc;
  if ((rc = pnh_adjoint(p, ones, img, aws, s))) return rc;
  if ((rc = run_passes_forward(p, img, aws, false, 0, nullptr, s))) return rc;
  return pnh_abs_first_coil(aws, W, p->Ktot, p->nb, s);
}

int pnh_precompute_sp(const pnh_plan* p, float* W, void* ws, void* stream) {
  REQUIRE(p && W && ws, "null argument");
  int rc = check_dtype_ready(p, PNH_C64);
  if (rc) return rc;
  return density(p, W, (char*)ws, (hipStream_t)stream);
}

int pnh_selfadjoint2(const pnh_plan* p, const void* x, const float* W, void* x_out,
                     void* ws, void* stream) {
  REQUIRE(p && x && W && x_out && ws, "null argument");
  hipStream_t s = (hipStream_t)stream;
  int rc;
  if ((rc = run_passes_forward(p, x, ws, false, 0, nullptr, s))) return rc;
  if ((rc = pnh_scale_sn(0, ws, ws, W, p->Ktot, p->nb, s))) return rc;
  return run_passes_inverse(p, ws, x_out, false, 0, nullptr, s);
}

int pnh_solve_dc(const pnh_plan* p, const void* y, void* x, int maxiter, void* ws,
                 void* stream) {
  REQUIRE(p && x && ws && (y || p->M == 0), "null argument");
  REQUIRE(maxiter >= 0, "maxiter must be >= 0");
  int rc = check_dtype_ready(p, PNH_C64);
  if (rc) return rc;
  if ((rc = check_dtype_ready(p, PNH_C128))) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int64_t my = p->M * p->nb;
  char* w = (char*)ws;
  void* W = w; w += align_up((size_t)my * 16);
  void* E = w; w += align_up((size_t)my * 16);
  void* Wy = w; w += align_up((size_t)my * 16);
  void* g64 = w; w += align_up((size_t)p->Ktot * p->nb * 16);
  void* g32 = w; w += align_up((size_t)p->Ktot * p->nb * 8);
  void* img64 = w; w += align_up((size_t)p->Ntot * p->nb * 16);
  void* part = w;
  // W = 1; W <- W / forward(adjoint(W)): the forward's FFT in complex64 and
  // its interpolation in FP64 (complex128 result), as the reference runs them
  if ((rc = pnh_fill_c128(W, my, 1.0, 0.0, s))) return rc;
  for (int it = 0; it < maxiter; ++it) {
    if ((rc = adjoint_c128(p, W, x, g64, img64, part, s))) return rc;
    if ((rc = run_passes_forward(p, x, g32, true, 0, nullptr, s))) return rc;
    if ((rc = pnh_c64_to_c128(g32, g64, p->Ktot * p->nb, s))) return rc;
    if (p->M > 0 && (rc = do_k2y(p, 1, g64, E, nullptr, s))) return rc;
    if ((rc = pnh_cdiv_c128(W, E, my, s))) return rc;
  }
  if ((rc = pnh_cmul_c128_c64(W, y, Wy, my, s))) return rc;
  return adjoint_c128(p, Wy, x, g64, img64, part, s);
}

int pnh_solve_l1tvols(const pnh_plan* p, const void* y, void* x, int maxiter, double rho,
                      const void* lap, void* ws, void* stream) {
  REQUIRE(p && x && ws && (y || p->M == 0), "null argument");
  REQUIRE(maxiter >= 0, "maxiter must be >= 0");
  REQUIRE(p->nb == 1, "L1TVOLS solves one coil (plan without batch)");
  REQUIRE(rho > 0.0, "rho must be > 0");
  int rc = check_dtype_ready(p, PNH_C64);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int64_t N = p->Ntot, K = p->Ktot;
  const size_t img = align_up((size_t)N * 8);
  char* w = (char*)ws;
  void* uker = w; w += align_up((size_t)K * 8);
  float* wabs = (float*)w; w += align_up((size_t)K * 4);
  void* AHy = w; w += img;
  void* AHyk = w; w += img;
  void* rhs = w; w += img;
  void* q = w; w += img;
  void* dd = w; w += p->D * img;
  void* bb = w; w += p->D * img;
  char* sws = w;          // self-adjoint scratch: grid first
  const float mu = 1.0f, lam = (float)(rho * 1.0);
  ShapeArgs gk, gn;
  gk.D = gn.D = p->D;
  for (int a = 0; a < 3; ++a) {
    gk.dim[a] = a < p->D ? p->K[a] : 1;
    gn.dim[a] = a < p->D ? p->N[a] : 1;
  }
  gk.n = K;
  gn.n = N;
  if ((rc = density(p, wabs, sws, s))) return rc;
  if ((rc = pnh_l1tv_uker(wabs, lap, uker, gk, mu, lam, s))) return rc;
  if ((rc = pnh_adjoint(p, y, AHy, sws, s))) return rc;
  if ((rc = pnh_fill_c64(AHyk, N, 0.f, 0.f, s))) return rc;
  if ((rc = pnh_fill_c64(dd, p->D * N, 0.f, 0.f, s))) return rc;
  if ((rc = pnh_fill_c64(bb, p->D * N, 0.f, 0.f, s))) return rc;
  if ((rc = pnh_fill_c64(x, N, 0.f, 0.f, s))) return rc;
  const float thr = (float)(1.0 / (double)lam);
  for (int it = 0; it < maxiter; ++it) {
    if ((rc = pnh_l1tv_rhs(AHyk, dd, bb, rhs, gn, mu, lam, s))) return rc;
    if ((rc = run_passes_forward(p, rhs, sws, false, 0, nullptr, s))) return rc;
    if ((rc = pnh_l1tv_kdiv(sws, uker, K, s))) return rc;
    if ((rc = run_passes_inverse(p, sws, x, false, 0, nullptr, s))) return rc;
    if ((rc = pnh_selfadjoint(p, x, q, sws, s))) return rc;
    if ((rc = pnh_l1tv_shrink(x, q, AHy, AHyk, dd, bb, gn, thr, s))) return rc;
  }
  return 0;
}

int pnh_solve_lsqr(const pnh_plan* p, const void* y, void* x, int iter_lim, double damp,
                   double atol, double btol, double conlim, void* ws, int32_t* info_out,
                   void* stream) {
  REQUIRE(p && x && ws && (y || p->M == 0), "null argument");
  REQUIRE(p->nb == 1, "lsqr solves one coil (plan without batch)");
  REQUIRE(damp >= 0.0, "damp must be >= 0");
  int rc = check_dtype_ready(p, PNH_C128);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int64_t M = p->M, K = p->Ktot;
  if (iter_lim <= 0) iter_lim = (int)(2 * K < (1LL << 30) ? 2 * K : (1LL << 30));
  char* w = (char*)ws;
  void* u = w; w += align_up((size_t)M * 16);
  void* tM = w; w += align_up((size_t)M * 16);
  void* v = w; w += align_up((size_t)K * 16);
  void* wv = w; w += align_up((size_t)K * 16);
  void* xk = w; w += align_up((size_t)K * 16);
  void* tK = w; w += align_up((size_t)K * 16);
  void* part = w; w += align_up(partial_bytes(p, 1));
  void* img64 = w; w += align_up((size_t)p->Ntot * 16);
  double* red = (double*)w; w += align_up((size_t)PNH_LSQR_BLOCKS * 8);
  LsqrState* st = (LsqrState*)w;
  const int* done = &st->done;
  const int* skip_v = &st->skip_v;
  // beta u = b, alfa v = A^H u, w = v, x = 0
  if ((rc = pnh_lsqr_init(st, atol, btol, conlim, damp, iter_lim, s))) return rc;
  if ((rc = pnh_fill_c128(xk, K, 0.0, 0.0, s))) return rc;
  if (M > 0 && (rc = pnh_c64_to_c128(y, u, M, s))) return rc;
  if ((rc = pnh_lsqr_norm2(u, M, red, nullptr, s))) return rc;
  if ((rc = pnh_lsqr_scalar(st, red, 0, s))) return rc;
  if ((rc = pnh_lsqr_scale(u, &st->beta, M, nullptr, done, s))) return rc;
  if ((rc = do_y2k(p, 1, u, v, part, done, s))) return rc;
  if ((rc = pnh_lsqr_norm2(v, K, red, done, s))) return rc;
  if ((rc = pnh_lsqr_scalar(st, red, 1, s))) return rc;
  if ((rc = pnh_lsqr_scale(v, &st->alfa, K, wv, done, s))) return rc;
  for (int it = 0; it < iter_lim; ++it) {
    // beta u = A v - alfa u; alfa v = A^H u - beta v
    if (M > 0 && (rc = do_k2y(p, 1, v, tM, done, s))) return rc;
    if ((rc = pnh_lsqr_axpby(u, tM, &st->alfa, M, red, done, s))) return rc;
    if ((rc = pnh_lsqr_scalar(st, red, 2, s))) return rc;
    if ((rc = pnh_lsqr_scale(u, &st->beta, M, nullptr, skip_v, s))) return rc;
    if ((rc = do_y2k(p, 1, u, tK, part, skip_v, s))) return rc;
    if ((rc = pnh_lsqr_axpby(v, tK, &st->beta, K, red, skip_v, s))) return rc;
    if ((rc = pnh_lsqr_scalar(st, red, 3, s))) return rc;
    if ((rc = pnh_lsqr_scale(v, &st->alfa, K, nullptr, skip_v, s))) return rc;
    if ((rc = pnh_lsqr_xw(xk, wv, v, st, K, red, s))) return rc;
    if ((rc = pnh_lsqr_scalar(st, red, 4, s))) return rc;
    if ((it & 63) == 63 && it + 1 < iter_lim) {
      // long runs: stop launching once the device has converged
      int flag = 0;
      PNH_CHECK_HIP(hipMemcpyAsync(&flag, done, sizeof(int), hipMemcpyDeviceToHost, s));
      PNH_CHECK_HIP(hipStreamSynchronize(s));
      if (flag) break;
    }
  }
  if ((rc = run_passes_inverse(p, xk, img64, false, 1, nullptr, s))) return rc;
  if ((rc = pnh_cast_div_sn(img64, x, p->d.sn_f32, p->Ntot, 1, s))) return rc;
  if (info_out) {
    PNH_CHECK_HIP(hipMemcpyAsync(info_out, &st->istop, sizeof(int), hipMemcpyDeviceToDevice, s));
    PNH_CHECK_HIP(hipMemcpyAsync(info_out + 1, &st->itn, sizeof(int), hipMemcpyDeviceToDevice, s));
  }
  return 0;
}

int pnh_solve_lsmr(const pnh_plan* p, const void* y, void* x, int maxiter, double damp,
                   double atol, double btol, double conlim, void* ws, int32_t* info_out,
                   void* stream) {
  REQUIRE(p && x && ws && (y || p->M == 0), "null argument");
  REQUIRE(p->nb == 1, "lsmr solves one coil (plan without batch)");
  REQUIRE(damp >= 0.0, "damp must be >= 0");
  int rc = check_dtype_ready(p, PNH_C128);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int64_t M = p->M, K = p->Ktot;
  if (maxiter <= 0) maxiter = (int)(M < K ? (M > 0 ? M : 1) : K);
  char* w = (char*)ws;
  void* u = w; w += align_up((size_t)M * 16);
  void* tM = w; w += align_up((size_t)M * 16);
  void* v = w; w += align_up((size_t)K * 16);
  void* h = w; w += align_up((size_t)K * 16);
  void* hbar = w; w += align_up((size_t)K * 16);
  void* xk = w; w += align_up((size_t)K * 16);
  void* tK = w; w += align_up((size_t)K * 16);
  void* part = w; w += align_up(partial_bytes(p, 1));
  void* img64 = w; w += align_up((size_t)p->Ntot * 16);
  double* red = (double*)w; w += align_up((size_t)PNH_LSQR_BLOCKS * 8);
  LsmrState* st = (LsmrState*)w;
  const int* done = &st->done;
  const int* skip_v = &st->skip_v;
  if ((rc = pnh_lsmr_init(st, atol, btol, conlim, damp, maxiter, s))) return rc;
  if ((rc = pnh_fill_c128(xk, K, 0.0, 0.0, s))) return rc;
  if ((rc = pnh_fill_c128(hbar, K, 0.0, 0.0, s))) return rc;
  if (M > 0 && (rc = pnh_c64_to_c128(y, u, M, s))) return rc;
  if ((rc = pnh_lsqr_norm2(u, M, red, nullptr, s))) return rc;
  if ((rc = pnh_lsmr_scalar(st, red, 0, s))) return rc;
  if ((rc = pnh_lsqr_scale(u, &st->beta, M, nullptr, done, s))) return rc;
  if ((rc = do_y2k(p, 1, u, v, part, done, s))) return rc;
  if ((rc = pnh_lsqr_norm2(v, K, red, done, s))) return rc;
  if ((rc = pnh_lsmr_scalar(st, red, 1, s))) return rc;
  if ((rc = pnh_lsqr_scale(v, &st->alpha, K, h, done, s))) return rc;
  for (int it = 0; it < maxiter; ++it) {
    if (M > 0 && (rc = do_k2y(p, 1, v, tM, done, s))) return rc;
    if ((rc = pnh_lsqr_axpby(u, tM, &st->alpha, M, red, done, s))) return rc;
    if ((rc = pnh_lsmr_scalar(st, red, 2, s))) return rc;
    if ((rc = pnh_lsqr_scale(u, &st->beta, M, nullptr, skip_v, s))) return rc;
    if ((rc = do_y2k(p, 1, u, tK, part, skip_v, s))) return rc;
    if ((rc = pnh_lsqr_axpby(v, tK, &st->beta, K, red, skip_v, s))) return rc;
    if ((rc = pnh_lsmr_scalar(st, red, 3, s))) return rc;
    if ((rc = pnh_lsqr_scale(v, &st->alpha, K, nullptr, skip_v, s))) return rc;
    if ((rc = pnh_lsmr_hx(xk, h, hbar, v, st, K, red, s))) return rc;
    if ((rc = pnh_lsmr_scalar(st, red, 4, s))) return rc;
    if ((it & 63) == 63 && it + 1 < maxiter) {
      int flag = 0;
      PNH_CHECK_HIP(hipMemcpyAsync(&flag, done, sizeof(int), hipMemcpyDeviceToHost, s));
      PNH_CHECK_HIP(hipStreamSynchronize(s));
      if (flag) break;
    }
  }
  if ((rc = run_passes_inverse(p, xk, img64, false, 1, nullptr, s))) return rc;
  if ((rc = pnh_cast_div_sn(img64, x, p->d.sn_f32, p->Ntot, 1, s))) return rc;
  if (info_out) {
    PNH_CHECK_HIP(hipMemcpyAsync(info_out, &st->istop, sizeof(int), hipMemcpyDeviceToDevice, s));
    PNH_CHECK_HIP(hipMemcpyAsync(info_out + 1, &st->itn, sizeof(int), hipMemcpyDeviceToDevice, s));
  }
  return 0;
}

int pnh_solve_bicg(const pnh_plan* p, int kind, const void* y, void* x, int maxiter,
                   double rtol, double atol, void* ws, int32_t* info_out, void* stream) {
  REQUIRE(p && x && ws && (y || p->M == 0), "null argument");
  REQUIRE(kind == PNH_BICGSTAB || kind == PNH_BICG, "unknown Krylov solver %d", kind);
  REQUIRE(p->nb == 1, "bicg/bicgstab solve one coil (plan without batch)");
  REQUIRE(atol >= 0.0, "atol must be >= 0");
  int rc = check_dtype_ready(p, PNH_C128);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int64_t M = p->M, K = p->Ktot;
  if (maxiter <= 0) maxiter = (int)(10 * K < (1LL << 30) ? 10 * K : (1LL << 30));
  const bool stab = kind == PNH_BICGSTAB;
  char* w = (char*)ws;
  void* ym = w; w += align_up((size_t)M * 16);
  void* xk = w; w += align_up((size_t)K * 16);
  void* r = w; w += align_up((size_t)K * 16);
  void* rt = w; w += align_up((size_t)K * 16);
  void* pv = w; w += align_up((size_t)K * 16);
  void* v = w; w += align_up((size_t)K * 16);   // bicgstab v, bicg q
  void* t = w; w += align_up((size_t)K * 16);   // bicgstab t, bicg qt
  void* pt = w; w += align_up((size_t)K * 16);  // bicg ptilde
  void* part = w; w += align_up(partial_bytes(p, 1));
  void* img64 = w; w += align_up((size_t)p->Ntot * 16);
  void* red = w; w += align_up((size_t)2 * PNH_LSQR_BLOCKS * 16);
  KrylovState* st = (KrylovState*)w;
  const int* done = &st->done;
  const int* skip_t = &st->skip_t;
  auto normal_op = [&](const void* in, void* out, const int* skip) -> int {
    return do_k2y2k(p, 1, in, ym, out, part, skip, s);
  };
  // b = P^H y; r = b, rt = r, x = 0
  if ((rc = pnh_krylov_init(st, stab ? 1 : 0, rtol, atol, maxiter, s))) return rc;
  if (M > 0 && (rc = pnh_c64_to_c128(y, ym, M, s))) return rc;
  if ((rc = do_y2k(p, 1, ym, r, part, nullptr, s))) return rc;
  if ((rc = pnh_fill_c128(xk, K, 0.0, 0.0, s))) return rc;
  PNH_CHECK_HIP(hipMemcpyAsync(rt, r, (size_t)K * 16, hipMemcpyDeviceToDevice, s));
  if ((rc = pnh_krylov_vdot2(r, r, nullptr, nullptr, K, red, done, s))) return rc;
  if ((rc = pnh_krylov_scalar(st, red, 0, s))) return rc;
  for (int it = 0; it < maxiter; ++it) {
    if ((rc = pnh_krylov_vdot2(r, r, rt, r, K, red, done, s))) return rc;
    if ((rc = pnh_krylov_scalar(st, red, 1, s))) return rc;
    if ((rc = pnh_krylov_p(pv, r, v, stab ? nullptr : pt, rt, st, K, s))) return rc;
    if ((rc = normal_op(pv, v, done))) return rc;
    if (stab) {
      if ((rc = pnh_krylov_vdot2(rt, v, nullptr, nullptr, K, red, done, s))) return rc;
      if ((rc = pnh_krylov_scalar(st, red, 2, s))) return rc;
      if ((rc = pnh_krylov_r(r, v, st, K, red, s))) return rc;
      if ((rc = pnh_krylov_scalar(st, red, 3, s))) return rc;
      if ((rc = normal_op(r, t, skip_t))) return rc;
      if ((rc = pnh_krylov_vdot2(t, r, t, t, K, red, skip_t, s))) return rc;
      if ((rc = pnh_krylov_scalar(st, red, 4, s))) return rc;
      if ((rc = pnh_krylov_x(xk, pv, r, t, nullptr, nullptr, st, K, s))) return rc;
    } else {
      if ((rc = normal_op(pt, t, done))) return rc;
      if ((rc = pnh_krylov_vdot2(pt, v, nullptr, nullptr, K, red, done, s))) return rc;
      if ((rc = pnh_krylov_scalar(st, red, 2, s))) return rc;
      if ((rc = pnh_krylov_x(xk, pv, r, v, rt, t, st, K, s))) return rc;
    }
    if ((rc = pnh_krylov_scalar(st, red, 5, s))) return rc;
    if ((it & 63) == 63 && it + 1 < maxiter) {
      int flag = 0;
      PNH_CHECK_HIP(hipMemcpyAsync(&flag, done, sizeof(int), hipMemcpyDeviceToHost, s));
      PNH_CHECK_HIP(hipStreamSynchronize(s));
      if (flag) break;
    }
  }
  if ((rc = run_passes_inverse(p, xk, img64, false, 1, nullptr, s))) return rc;
  if ((rc = pnh_cast_div_sn(img64, x, p->d.sn_f32, p->Ntot, 1, s))) return rc;
  if (info_out) {
    PNH_CHECK_HIP(hipMemcpyAsync(info_out, &st->info, sizeof(int), hipMemcpyDeviceToDevice, s));
    PNH_CHECK_HIP(hipMemcpyAsync(info_out + 1, &st->it, sizeof(int), hipMemcpyDeviceToDevice, s));
  }
  return 0;
}

static int gmres_restart(const pnh_plan* p, int restart) {
  if (restart <= 0) restart = 20;
  if ((int64_t)restart > p->Ktot) restart = (int)p->Ktot;
  return restart;
}

int pnh_gmres_workspace_bytes(const pnh_plan* p, int restart, size_t* bytes) {
  REQUIRE(p && bytes, "null argument");
  restart = gmres_restart(p, restart);
  REQUIRE(restart >= 1 && restart <= PNH_GMRES_MAXR, "restart must be 1..%d", PNH_GMRES_MAXR);
  const size_t kv = align_up((size_t)p->Ktot * 16);
  *bytes = align_up((size_t)p->M * 16) + (size_t)(restart + 1) * kv + 4 * kv +
           align_up(partial_bytes(p, 1)) + align_up((size_t)p->Ntot * 16) +
           align_up((size_t)2 * PNH_LSQR_BLOCKS * 16) + align_up(sizeof(GmresState));
  return 0;
}

int pnh_solve_gmres(const pnh_plan* p, const void* y, void* x, int restart, int maxiter,
                    double rtol, double atol, void* ws, int32_t* info_out, void* stream) {
  REQUIRE(p && x && ws && (y || p->M == 0), "null argument");
  REQUIRE(p->nb == 1, "gmres solves one coil (plan without batch)");
  REQUIRE(atol >= 0.0, "atol must be >= 0");
  restart = gmres_restart(p, restart);
  REQUIRE(restart >= 1 && restart <= PNH_GMRES_MAXR, "restart must be 1..%d", PNH_GMRES_MAXR);
  int rc = check_dtype_ready(p, PNH_C128);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int64_t M = p->M, K = p->Ktot;
  if (maxiter <= 0) maxiter = (int)(10 * K < (1LL << 30) ? 10 * K : (1LL << 30));
  const size_t kv = align_up((size_t)K * 16);
  char* w = (char*)ws;
  void* ym = w; w += align_up((size_t)M * 16);
  char* V = w; w += (size_t)(restart + 1) * kv;
  void* wv = w; w += kv;
  void* xk = w; w += kv;
  void* r = w; w += kv;
  void* bv = w; w += kv;
  void* part = w; w += align_up(partial_bytes(p, 1));
  void* img64 = w; w += align_up((size_t)p->Ntot * 16);
  void* red = w; w += align_up((size_t)2 * PNH_LSQR_BLOCKS * 16);
  GmresState* st = (GmresState*)w;
  const int* done = &st->done;
  const int* skip = &st->skip;
  const int64_t ldv = (int64_t)(kv / 16);
  auto vec = [&](int k) -> void* { return V + (size_t)k * kv; };
  auto normal_op = [&](const void* in, void* out, const int* sk) -> int {
    return do_k2y2k(p, 1, in, ym, out, part, sk, s);
  };
  if ((rc = pnh_gmres_init(st, restart, rtol, atol, maxiter, 0, s))) return rc;
  if (M > 0 && (rc = pnh_c64_to_c128(y, ym, M, s))) return rc;
  if ((rc = do_y2k(p, 1, ym, bv, part, nullptr, s))) return rc;
  PNH_CHECK_HIP(hipMemcpyAsync(r, bv, (size_t)K * 16, hipMemcpyDeviceToDevice, s));
  if ((rc = pnh_fill_c128(xk, K, 0.0, 0.0, s))) return rc;
  if ((rc = pnh_krylov_vdot2(bv, bv, nullptr, nullptr, K, red, done, s))) return rc;
  if ((rc = pnh_gmres_scalar(st, red, 0, 0, 0, s))) return rc;
  for (int outer = 0; outer < maxiter; ++outer) {
    if ((rc = pnh_krylov_vdot2(r, r, nullptr, nullptr, K, red, done, s))) return rc;
    if ((rc = pnh_gmres_scalar(st, red, 1, 0, 0, s))) return rc;
    if ((rc = pnh_gmres_normalize(vec(0), r, st, K, -1, s))) return rc;
    for (int col = 0; col < restart; ++col) {
      if ((rc = normal_op(vec(col), wv, skip))) return rc;
      if ((rc = pnh_krylov_vdot2(wv, wv, nullptr, nullptr, K, red, skip, s))) return rc;
      if ((rc = pnh_gmres_scalar(st, red, 2, col, 0, s))) return rc;
      for (int k = 0; k <= col; ++k) {
        if ((rc = pnh_krylov_vdot2(vec(k), wv, nullptr, nullptr, K, red, skip, s))) return rc;
        if ((rc = pnh_gmres_scalar(st, red, 3, col, k, s))) return rc;
        if ((rc = pnh_gmres_mgs(wv, vec(k), st, col, k, K, s))) return rc;
      }
      if ((rc = pnh_krylov_vdot2(wv, wv, nullptr, nullptr, K, red, skip, s))) return rc;
      if ((rc = pnh_gmres_scalar(st, red, 4, col, 0, s))) return rc;
      if ((rc = pnh_gmres_normalize(vec(col + 1), wv, st, K, col, s))) return rc;
    }
    if ((rc = pnh_gmres_solve(st, s))) return rc;
    if ((rc = pnh_gmres_x(xk, V, ldv, st, K, s))) return rc;
    if ((rc = normal_op(xk, wv, done))) return rc;
    if ((rc = pnh_gmres_residual(r, bv, wv, st, K, red, s))) return rc;
    if ((rc = pnh_gmres_scalar(st, red, 6, 0, 0, s))) return rc;
    if (outer + 1 < maxiter) {
      int flag = 0;
      PNH_CHECK_HIP(hipMemcpyAsync(&flag, done, sizeof(int), hipMemcpyDeviceToHost, s));
      PNH_CHECK_HIP(hipStreamSynchronize(s));
      if (flag) break;
    }
  }
  if ((rc = run_passes_inverse(p, xk, img64, false, 1, nullptr, s))) return rc;
  if ((rc = pnh_cast_div_sn(img64, x, p->d.sn_f32, p->Ntot, 1, s))) return rc;
  if (info_out) {
    PNH_CHECK_HIP(hipMemcpyAsync(info_out, &st->info, sizeof(int), hipMemcpyDeviceToDevice, s));
    PNH_CHECK_HIP(hipMemcpyAsync(info_out + 1, &st->inner_iter, sizeof(int),
                                 hipMemcpyDeviceToDevice, s));
  }
  return 0;
}

int pnh_lgmres_workspace_bytes(const pnh_plan* p, int inner_m, int outer_k, size_t* bytes) {
  REQUIRE(p && bytes, "null argument");
  REQUIRE(inner_m >= 1 && outer_k >= 0 && outer_k <= 8 && inner_m + outer_k <= PNH_GMRES_MAXR,
          "lgmres needs inner_m >= 1, 0 <= outer_k <= 8, inner_m + outer_k <= %d",
          PNH_GMRES_MAXR);
  const size_t kv = align_up((size_t)p->Ktot * 16);
  *bytes = align_up((size_t)p->M * 16) + (size_t)(inner_m + outer_k + 1) * kv +
           2 * (size_t)outer_k * kv + 6 * kv + align_up(partial_bytes(p, 1)) +
           align_up((size_t)p->Ntot * 16) + align_up((size_t)2 * PNH_LSQR_BLOCKS * 16) +
           align_up(sizeof(GmresState));
  return 0;
}

int pnh_solve_lgmres(const pnh_plan* p, const void* y, void* x, int maxiter, int inner_m,
                     int outer_k, double rtol, double atol, void* ws, int32_t* info_out,
                     void* stream) {
  REQUIRE(p && x && ws && (y || p->M == 0), "null argument");
  REQUIRE(p->nb == 1, "lgmres solves one coil (plan without batch)");
  REQUIRE(atol >= 0.0, "atol must be >= 0");
  REQUIRE(maxiter >= 1, "maxiter must be >= 1");
  REQUIRE(inner_m >= 1 && outer_k >= 0 && outer_k <= 8 && inner_m + outer_k <= PNH_GMRES_MAXR,
          "lgmres needs inner_m >= 1, 0 <= outer_k <= 8, inner_m + outer_k <= %d",
          PNH_GMRES_MAXR);
  int rc = check_dtype_ready(p, PNH_C128);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int64_t M = p->M, K = p->Ktot;
  const int mmax = inner_m + outer_k;
  const size_t kv = align_up((size_t)K * 16);
  char* w = (char*)ws;
  void* ym = w; w += align_up((size_t)M * 16);
  char* V = w; w += (size_t)(mmax + 1) * kv;
  char* OZ = w; w += (size_t)outer_k * kv;
  char* OAZ = w; w += (size_t)outer_k * kv;
  void* wv = w; w += kv;
  void* xk = w; w += kv;
  void* r = w; w += kv;
  void* bv = w; w += kv;
  void* dx = w; w += kv;
  void* tK = w; w += kv;
  void* part = w; w += align_up(partial_bytes(p, 1));
  void* img64 = w; w += align_up((size_t)p->Ntot * 16);
  void* red = w; w += align_up((size_t)2 * PNH_LSQR_BLOCKS * 16);
  GmresState* st = (GmresState*)w;
  const int* done = &st->done;
  const int* skip = &st->skip;
  const int64_t ldv = (int64_t)(kv / 16);
  auto vec = [&](int k) -> void* { return V + (size_t)k * kv; };
  auto normal_op = [&](const void* in, void* out, const int* sk) -> int {
    return do_k2y2k(p, 1, in, ym, out, part, sk, s);
  };
  // b = P^H y; x = 0, so the first residual b - A x is b itself
  if ((rc = pnh_gmres_init(st, mmax, rtol, atol, maxiter, 1, s))) return rc;
  if (M > 0 && (rc = pnh_c64_to_c128(y, ym, M, s))) return rc;
  if ((rc = do_y2k(p, 1, ym, bv, part, nullptr, s))) return rc;
  if ((rc = pnh_fill_c128(xk, K, 0.0, 0.0, s))) return rc;
  PNH_CHECK_HIP(hipMemcpyAsync(r, bv, (size_t)K * 16, hipMemcpyDeviceToDevice, s));
  if ((rc = pnh_krylov_vdot2(bv, bv, nullptr, nullptr, K, red, done, s))) return rc;
  if ((rc = pnh_gmres_scalar(st, red, 0, 0, 0, s))) return rc;
  for (int outer = 0; outer < maxiter; ++outer) {
    // augmentation vectors, oldest first (one stored per outer iteration)
    const int n_aug = outer < outer_k ? outer : outer_k;
    const int first_it = outer - n_aug;
    LgmresZ zo;
    memset(&zo, 0, sizeof(zo));
    for (int t = 0; t < n_aug; ++t) zo.z[t] = OZ + (size_t)((first_it + t) % outer_k) * kv;
    const int m = inner_m + n_aug;
    if (outer > 0) {
      if ((rc = normal_op(xk, tK, done))) return rc;
      if ((rc = pnh_gmres_residual(r, bv, tK, st, K, red, s))) return rc;
    } else {
      if ((rc = pnh_krylov_vdot2(r, r, nullptr, nullptr, K, red, done, s))) return rc;
    }
    if ((rc = pnh_gmres_scalar(st, red, 7, 0, 0, s))) return rc;
    if ((rc = pnh_gmres_normalize(vec(0), r, st, K, -1, s))) return rc;
    for (int col = 0; col < m; ++col) {
      if (col < inner_m) {
        if ((rc = normal_op(vec(col), wv, skip))) return rc;
      } else {
        const int t = col - inner_m;
        const void* az = OAZ + (size_t)((first_it + t) % outer_k) * kv;
        PNH_CHECK_HIP(hipMemcpyAsync(wv, az, (size_t)K * 16, hipMemcpyDeviceToDevice, s));
      }
      if ((rc = pnh_krylov_vdot2(wv, wv, nullptr, nullptr, K, red, skip, s))) return rc;
      if ((rc = pnh_gmres_scalar(st, red, 2, col, 0, s))) return rc;
      for (int k = 0; k <= col; ++k) {
        if ((rc = pnh_krylov_vdot2(vec(k), wv, nullptr, nullptr, K, red, skip, s))) return rc;
        if ((rc = pnh_gmres_scalar(st, red, 3, col, k, s))) return rc;
        if ((rc = pnh_gmres_mgs(wv, vec(k), st, col, k, K, s))) return rc;
      }
      if ((rc = pnh_krylov_vdot2(wv, wv, nullptr, nullptr, K, red, skip, s))) return rc;
      if ((rc = pnh_gmres_scalar(st, red, 4, col, 0, s))) return rc;
      if ((rc = pnh_gmres_normalize(vec(col + 1), wv, st, K, col, s))) return rc;
    }
    if ((rc = pnh_gmres_solve(st, s))) return rc;
    if ((rc = pnh_lgmres_dx(dx, V, ldv, inner_m, zo, st, K, red, s))) return rc;
    if ((rc = pnh_gmres_scalar(st, red, 8, 0, 0, s))) return rc;
    if (outer_k > 0) {
      const size_t slot = (size_t)(outer % outer_k) * kv;
      if ((rc = pnh_lgmres_store(xk, dx, V, ldv, OZ + slot, OAZ + slot, st, K, s))) return rc;
    } else if ((rc = pnh_lgmres_store(xk, dx, V, ldv, nullptr, nullptr, st, K, s))) {
      return rc;
    }
    if ((rc = pnh_gmres_scalar(st, red, 9, 0, 0, s))) return rc;
    if (outer + 1 < maxiter) {
      int flag = 0;
      PNH_CHECK_HIP(hipMemcpyAsync(&flag, done, sizeof(int), hipMemcpyDeviceToHost, s));
      PNH_CHECK_HIP(hipStreamSynchronize(s));
      if (flag) break;
    }
  }
  if ((rc = run_passes_inverse(p, xk, img64, false, 1, nullptr, s))) return rc;
  if ((rc = pnh_cast_div_sn(img64, x, p->d.sn_f32, p->Ntot, 1, s))) return rc;
  if (info_out) {
    PNH_CHECK_HIP(hipMemcpyAsync(info_out, &st->info, sizeof(int), hipMemcpyDeviceToDevice, s));
    PNH_CHECK_HIP(hipMemcpyAsync(info_out + 1, &st->inner_iter, sizeof(int),
                                 hipMemcpyDeviceToDevice, s));
  }
  return 0;
}

int pnh_forward_grad_om(const pnh_plan* p, const void* k, const void* g, float* om_grad,
                        void* stream) {
  REQUIRE(p && k && (g || p->M == 0) && (om_grad || p->M == 0), "null argument");
  REQUIRE(p->d.du_c64 && p->d.u_c64, "plan has no factor derivatives (du_c64)");
  InterpArgs ia = interp_args(p, 0);
  return pnh_k2y_dom_launch(ia, 0, k, g, om_grad, (hipStream_t)stream);
}

// The CG solve; host_checks: read the convergence flag back every 64
// iterations and stop launching (off while a graph is being captured).
static int cg_solve(const pnh_plan* p, const void* y, void* x, int maxiter, double rtol,
                    double atol, void* ws, int32_t* iters_out, double* resid_out,
                    void* stream, bool host_checks) {
  REQUIRE(p && x && ws && (y || p->M == 0), "null argument");
  REQUIRE(maxiter >= 0, "maxiter must be >= 0");
  REQUIRE(p->nb <= PNH_CG_MAXNB, "solve supports at most %d coils", PNH_CG_MAXNB);
  REQUIRE(p->d.u_c128 && p->d.tw_c128[0] && p->d.sn_f32,
          "FP64 solve needs u_c128 and tw_c128 in the plan");
  for (int a = 0; a < p->D; ++a)
    REQUIRE(p->d.tw_c128[a], "tw_c128[%d] missing", a);
  hipStream_t s = (hipStream_t)stream;
  const int nb = p->nb;
  const size_t g2 = align_up((size_t)p->Ktot * nb * 16);
  const size_t y2 = align_up((size_t)p->M * nb * 16);
  const size_t part2 = align_up(partial_bytes(p, 1));
  const size_t red = align_up((size_t)PNH_RED_BLOCKS * nb * 16);
  char* w = (char*)ws;
  void* vx = w; w += g2;
  void* vr = w; w += g2;
  void* vp = w; w += g2;
  void* vq = w; w += g2;
  void* yt = w; w += y2;
  void* part = w; w += part2;
  void* redp = w; w += red;
  void* redpq = w; w += red;
  CgScalars* sc = (CgScalars*)w;
  int rc;
  const int64_t K = p->Ktot;
  // b = P^H y in FP64 (y is complex64; upcast exactly as scipy does)
  if (p->M > 0 && (rc = pnh_c64_to_c128(y, yt, p->M * nb, s))) return rc;
  if ((rc = do_y2k(p, 1, yt, vr, part, nullptr, s))) return rc;
  if ((rc = pnh_zero(1, vx, K * nb, s))) return rc;
  if ((rc = pnh_cg_dot(1, vr, vr, K, nb, redp, nullptr, s))) return rc;
  if ((rc = pnh_cg_init(redp, nb, sc, maxiter, rtol, atol, s))) return rc;
  for (int it = 0; it < maxiter; ++it) {
    // the p update finishes the previous iteration's |r|^2 and stopping test
    const int* all_done = &sc->all_done[it & 1];
    if ((rc = pnh_cg_p_update(vp, vr, K, nb, sc, redp, it, s))) return rc;
    if ((rc = do_k2y2k(p, 1, vp, yt, vq, part, all_done, s))) return rc;
    if ((rc = pnh_cg_dot(0, vp, vq, K, nb, redpq, all_done, s))) return rc;
    if ((rc = pnh_cg_xr_update(vx, vr, vp, vq, K, nb, sc, it, redpq, redp, s))) return rc;
    if (host_checks && (it & 63) == 63 && it + 1 < maxiter) {
      // long runs: stop launching once every coil has converged (as seen by
      // this iteration's p update)
      int flag = 0;
      PNH_CHECK_HIP(hipMemcpyAsync(&flag, all_done, sizeof(int), hipMemcpyDeviceToHost, s));
      PNH_CHECK_HIP(hipStreamSynchronize(s));
      if (flag) break;
    }
  }
  // x = complex64(crop(ifftn(k))) / sn, inverse FFT in FP64 (as numpy does
  // for the complex128 CG result)
  if ((rc = run_passes_inverse(p, vx, vq, false, 1, nullptr, s))) return rc;
  if ((rc = pnh_cast_div_sn(vq, x, p->d.sn_f32, p->Ntot, nb, s))) return rc;
  if (iters_out)
    PNH_CHECK_HIP(hipMemcpyAsync(iters_out, sc->iters, sizeof(int) * nb,
                                 hipMemcpyDeviceToDevice, s));
  if (resid_out && (rc = pnh_cg_resid(redp, nb, sc, resid_out, s))) return rc;
  return 0;
}

int pnh_solve_cg(const pnh_plan* p, const void* y, void* x, int maxiter, double rtol,
                 double atol, void* ws, int32_t* iters_out, double* resid_out,
                 void* stream) {
  return cg_solve(p, y, x, maxiter, rtol, atol, ws, iters_out, resid_out, stream, true);
}

}  // extern "C"

// ------------------------------------------------ CG with static scratch --
// Workspace of pnh_solve_cg_noalloc: a file-scope device array, one copy per
// device, held until the process ends.  Host code never uses the array's
// name as an address (on the host it is only a stand-in for the symbol); it
// uses the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[PNH_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices];

// Device address of the current device's scratch array, looked up once per
// device (with that device current) and kept.
int scratch_address(void** out) {
  int dev = -1;
  PNH_CHECK_HIP(hipGetDevice(&dev));
  REQUIRE(dev >= 0 && dev < kScratchDevices, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      pnh_set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev,
                    hipGetErrorString(e));
      return -3;
    }
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" {

int pnh_solve_cg_noalloc(const pnh_plan* p, const void* y, void* x, int maxiter, double rtol,
                         double atol, int32_t* iters_out, double* resid_out, void* stream) {
  REQUIRE(p && x && (y || p->M == 0), "null argument");
  size_t need = 0;
  int rc = pnh_workspace_bytes(p, PNH_OP_SOLVE_CG, PNH_C128, &need);
  if (rc) return rc;
  REQUIRE(need <= (size_t)PNH_NOALLOC_SCRATCH_BYTES,
          "this solve needs %zu bytes of scratch, above the %zu bytes of "
          "pnh_solve_cg_noalloc's static scratch; call pnh_solve_cg with a workspace",
          need, (size_t)PNH_NOALLOC_SCRATCH_BYTES);
  void* ws = nullptr;
  if ((rc = scratch_address(&ws))) return rc;
  // Every CG buffer in the scratch is written before it is read (partial
  // slots and item exponents by the gridding, x by a zero kernel, p by the
  // first p update, the scalars by the init kernel), so nothing is cleared.
  return cg_solve(p, y, x, maxiter, rtol, atol, ws, iters_out, resid_out, stream, true);
}

}  // extern "C"

// ================================================================ batches ==
#include <algorithm>
#include <vector>

#include "batch.h"

struct pnh_batch {
  int n, D, J_last, J_max, grid_threads;
  int packed;                          // every problem has per-item bounds
  int fixed_pad;                       // every problem's boxes have extent PNH_FIXED_PAD
                                       // and every problem a row bound
  int J[PNH_MAXD];
  int64_t N_total, M_total, K_total, part_elems, pad_cells_max;
  int fft_threads;
  BatchProblem* probs;                 // device (inside the caller's table)
  const int2* items;
  int64_t n_items;
  BatchTable gather_t;                 // one block per tile
  BatchTable fft_t[3][PNH_MAXD];       // [kind][axis]
  size_t fft_lds[3][PNH_MAXD];
  int fft_pow2[3][PNH_MAXD];
  BatchTable vec_t[3];                 // [0] grid cells, [1] samples, [2] image
  int64_t vec_chunk;
  size_t table_bytes;
};

struct pnh_graph {
  hipGraphExec_t exec;
};

namespace {

const int64_t kVecChunk = 4096;

struct Blob {
  std::vector<unsigned char> bytes;
  size_t add(const void* src, size_t n) {
    const size_t off = (bytes.size() + 255) & ~(size_t)255;
    bytes.resize(off + n);
    if (n) memcpy(bytes.data() + off, src, n);
    return off;
  }
};

void add_table(Blob& blob, const std::vector<int64_t>& blocks_per_problem,
               size_t* off_tab, size_t* off_first, int64_t* n_blocks) {
  std::vector<int> tab, first;
  int64_t acc = 0;
  for (size_t p = 0; p < blocks_per_problem.size(); ++p) {
    first.push_back((int)acc);
    for (int64_t b = 0; b < blocks_per_problem[p]; ++b) tab.push_back((int)p);
    acc += blocks_per_problem[p];
  }
  first.push_back((int)acc);
  if (tab.empty()) tab.push_back(0);
  *off_tab = blob.add(tab.data(), tab.size() * sizeof(int));
  *off_first = blob.add(first.data(), first.size() * sizeof(int));
  *n_blocks = acc;
}

// Build every host-side table of a batch.  With `table` NULL only sizes are
// computed; otherwise the blob is copied to `table` and device pointers set.
int build_batch(const pnh_plan_desc* descs, int n, void* table, hipStream_t stream,
                pnh_batch* b) {
  REQUIRE(descs && n >= 1, "need at least one problem");
  memset(b, 0, sizeof(*b));
  std::vector<pnh_plan> plans(n);
  for (int i = 0; i < n; ++i) {
    int rc = plan_init(descs[i], &plans[i]);
    if (rc) return rc;
    const pnh_plan& p = plans[i];
    REQUIRE(p.nb == 1, "batched problems are single-coil (problem %d has batch %d)", i, p.nb);
    REQUIRE(p.d.u_c64 && p.d.u_c128 && p.d.tw_c128[0], "problem %d lacks c64/c128 factors", i);
    REQUIRE(!p.d.csr_c64, "batched problems use the per-axis factors");
    REQUIRE(!p.d.fp64_ops, "problem %d needs FP64 operators (fp64_ops); batches run the "
            "complex64 operators in FP32", i);
    if (i > 0) {
      REQUIRE(p.D == plans[0].D, "all problems must have the same ndim");
      for (int a = 0; a < p.D; ++a) {
        REQUIRE(p.J[a] == plans[0].J[a], "all problems must have the same Jd");
      }
    }
  }
  const int D = plans[0].D;
  b->n = n;
  b->D = D;
  b->J_last = plans[0].J[D - 1];
  b->J_max = 1;
  for (int a = 0; a < PNH_MAXD; ++a) b->J[a] = a < D ? plans[0].J[a] : 1;
  for (int a = 0; a < D; ++a) b->J_max = std::max(b->J_max, plans[0].J[a]);
  b->pad_cells_max = 1;
  b->grid_threads = 64;
  b->packed = 1;
  b->fixed_pad = 1;
  for (int i = 0; i < n; ++i) {
    b->pad_cells_max = std::max(b->pad_cells_max, plans[i].bins.pad_cells);
    b->grid_threads = std::max(b->grid_threads, plans[i].bins.threads);
    if (!plans[i].bins.item_bound && plans[i].bins.n_items > 0) b->packed = 0;
    if (!pnh_fixed_pad(plans[i].bins, D) || !(plans[i].bins.row_bound > 0.0)) b->fixed_pad = 0;
  }
  b->fft_threads = 256;
  b->vec_chunk = kVecChunk;
  std::vector<BatchProblem> probs(n);
  int64_t xo = 0, yo = 0, ko = 0, so = 0;
  for (int i = 0; i < n; ++i) {
    const pnh_plan& p = plans[i];
    BatchProblem& q = probs[i];
    memset(&q, 0, sizeof(q));
    q.ia = interp_args(&p, 0);
    q.ia64 = interp_args(&p, 1);
    q.bins = p.bins;
    q.x_off = xo; q.y_off = yo; q.k_off = ko; q.part_off = so;
    q.n_img = p.Ntot;
    q.n_grid = p.Ktot;
    q.sn32 = p.d.sn_f32;
    for (int a = 0; a < D; ++a) {
      int rc = build_pass(&p, a, false, a == D - 1, false, a == D - 1 ? 1 : 0, 0, 1.0, &q.fwd[a]);
      const double sc = a == D - 1 ? 1.0 / (double)p.Ktot : 1.0;
      if (!rc) rc = build_pass(&p, a, true, false, a == D - 1, a == D - 1 ? 2 : 0, 0, sc, &q.inv[a]);
      if (!rc) rc = build_pass(&p, a, true, false, a == D - 1, 0, 1, sc, &q.inv64[a]);
      if (rc) return rc;
      q.tw32[a] = p.d.tw_c64[a];
      q.tw64[a] = p.d.tw_c128[a];
    }
    xo += p.Ntot; yo += p.M; ko += p.Ktot; so += p.d.n_slots * p.bins.pad_cells;
  }
  b->N_total = xo; b->M_total = yo; b->K_total = ko; b->part_elems = so;

  Blob blob;
  const size_t off_probs = blob.add(probs.data(), probs.size() * sizeof(BatchProblem));
  // gridding gather: pnh_gather_parts blocks per tile
  std::vector<int64_t> nb_tiles(n);
  for (int i = 0; i < n; ++i)
    nb_tiles[i] = plans[i].bins.n_tiles * pnh_gather_parts(plans[i].bins);
  size_t og_tab, og_first;
  add_table(blob, nb_tiles, &og_tab, &og_first, &b->gather_t.n_blocks);
  // FFT passes per kind and axis
  size_t of_tab[3][PNH_MAXD], of_first[3][PNH_MAXD];
  for (int kind = 0; kind < 3; ++kind)
    for (int a = 0; a < D; ++a) {
      std::vector<int64_t> nbk(n);
      size_t lds = 0;
      int pow2 = 1;
      for (int i = 0; i < n; ++i) {
        FftDesc& f = kind == BFFT_FWD ? probs[i].fwd[a]
                                      : (kind == BFFT_INV ? probs[i].inv[a] : probs[i].inv64[a]);
        const int64_t lines = f.n_outer * f.n_inner;
        nbk[i] = (lines + f.lines_per_block - 1) / f.lines_per_block;
        lds = std::max(lds, pnh_fft_lds_bytes(f, kind == BFFT_INV64));
        for (int q = 0; q < f.npass; ++q)
          pow2 &= (f.radix[q] == 2 || f.radix[q] == 4 || f.radix[q] == 8 || f.radix[q] == 16);
      }
      add_table(blob, nbk, &of_tab[kind][a], &of_first[kind][a], &b->fft_t[kind][a].n_blocks);
      b->fft_lds[kind][a] = lds;
      b->fft_pow2[kind][a] = pow2;
    }
  // vector tables: grid cells, samples, image
  size_t ov_tab[3], ov_first[3];
  for (int e = 0; e < 3; ++e) {
    std::vector<int64_t> nbv(n);
    for (int i = 0; i < n; ++i) {
      const int64_t len = e == 0 ? plans[i].Ktot : (e == 1 ? plans[i].M : plans[i].Ntot);
      nbv[i] = (len + kVecChunk - 1) / kVecChunk;
    }
    add_table(blob, nbv, &ov_tab[e], &ov_first[e], &b->vec_t[e].n_blocks);
  }
  // gridding items of all problems, heaviest first (needs the counts)
  std::vector<std::pair<int, int2>> all;  // (count, (problem, item))
  for (int i = 0; i < n; ++i) {
    std::vector<int4> it(plans[i].d.n_items);
    if (table && !it.empty()) {
      // ordered after the caller's uploads of the work lists on `stream`
      PNH_CHECK_HIP(hipMemcpyAsync(it.data(), plans[i].d.items, it.size() * sizeof(int4),
                                   hipMemcpyDeviceToHost, stream));
      PNH_CHECK_HIP(hipStreamSynchronize(stream));
    }
    for (int64_t k = 0; k < plans[i].d.n_items; ++k)
      all.push_back(std::make_pair(table ? it[k].z : 0, make_int2(i, (int)k)));
  }
  std::stable_sort(all.begin(), all.end(), [](const std::pair<int, int2>& x,
                                              const std::pair<int, int2>& y) {
    return x.first > y.first;
  });
  std::vector<int2> items(all.size());
  for (size_t k = 0; k < all.size(); ++k) items[k] = all[k].second;
  if (items.empty()) items.push_back(make_int2(0, 0));
  const size_t oi = blob.add(items.data(), items.size() * sizeof(int2));
  b->n_items = (int64_t)all.size();
  b->table_bytes = blob.bytes.size();
  if (!table) return 0;

  unsigned char* base = (unsigned char*)table;
  PNH_CHECK_HIP(hipMemcpyAsync(base, blob.bytes.data(), blob.bytes.size(),
                               hipMemcpyHostToDevice, stream));
  PNH_CHECK_HIP(hipStreamSynchronize(stream));
  b->probs = (BatchProblem*)(base + off_probs);
  b->gather_t.tab = (const int*)(base + og_tab);
  b->gather_t.first = (const int*)(base + og_first);
  for (int kind = 0; kind < 3; ++kind)
    for (int a = 0; a < D; ++a) {
      b->fft_t[kind][a].tab = (const int*)(base + of_tab[kind][a]);
      b->fft_t[kind][a].first = (const int*)(base + of_first[kind][a]);
    }
  for (int e = 0; e < 3; ++e) {
    b->vec_t[e].tab = (const int*)(base + ov_tab[e]);
    b->vec_t[e].first = (const int*)(base + ov_first[e]);
  }
  b->items = (const int2*)(base + oi);
  return 0;
}

// Forward kind: image -> grid, last axis first (image read by the first
// pass).  Inverse kinds: grid -> image, first axis first (image written by
// the last pass).  Intermediate passes work in place on the grid.
int batch_passes(const pnh_batch* b, int kind, int fp64, const void* image_in,
                 void* image_out, void* grid, const int* done, const int* all_done,
                 hipStream_t s) {
  for (int step = 0; step < b->D; ++step) {
    const int a = kind == BFFT_FWD ? b->D - 1 - step : step;
    const bool img_in = kind == BFFT_FWD && a == b->D - 1;
    const bool img_out = kind != BFFT_FWD && a == b->D - 1;
    int rc = pnh_batch_fft_launch(b->probs, b->fft_t[kind][a], a, kind, fp64,
                                  b->fft_pow2[kind][a], b->fft_lds[kind][a], b->fft_threads,
                                  img_in ? image_in : grid, img_in, img_out ? image_out : grid,
                                  img_out, done, all_done, s);
    if (rc) return rc;
  }
  return 0;
}

int batch_y2k(const pnh_batch* b, int fp64, const void* y, void* k, void* part,
              const int* done, const int* all_done, hipStream_t s) {
  return pnh_batch_y2k_launch(b->probs, b->items, b->n_items, b->gather_t,
                              b->D, b->J, fp64, b->packed, b->pad_cells_max, b->grid_threads,
                              y, k, part, done, all_done, s);
}

struct BatchWs {
  void *x, *r, *p, *q, *yt, *part;
  BatchCg cg;
};

// Carve the CG workspace; with ws == NULL only the size is computed.
size_t batch_cg_layout(const pnh_batch* b, void* ws, BatchWs* w) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    void* ptr = ws ? (void*)((char*)ws + off) : nullptr;
    off += align_up(bytes);
    return ptr;
  };
  const size_t g2 = (size_t)b->K_total * 16;
  const size_t n = b->n;
  w->x = take(g2);
  w->r = take(g2);
  w->p = take(g2);
  w->q = take(g2);
  w->yt = take((size_t)b->M_total * 16);
  w->part = take((size_t)b->part_elems * 16);
  w->cg.tol = (double*)take(n * 8);
  w->cg.rr = (double*)take(2 * n * 8);
  w->cg.done = (int*)take(2 * n * 4);
  w->cg.iters = (int*)take(n * 4);
  w->cg.stopped = (int*)take(n * 4);
  w->cg.res = (double*)take(n * 8);
  w->cg.part_rr = (double2*)take((size_t)b->vec_t[0].n_blocks * 16);
  w->cg.part_pq = (double2*)take((size_t)b->vec_t[0].n_blocks * 16);
  w->cg.n = b->n;
  return off;
}

int batch_solve(const pnh_batch* b, const void* y, void* x, int maxiter, double rtol,
                double atol, void* ws, int32_t* iters_out, double* resid_out, hipStream_t s) {
  BatchWs w;
  batch_cg_layout(b, ws, &w);
  BatchCg cg = w.cg;
  cg.maxiter = maxiter;
  cg.rtol = rtol;
  cg.atol = atol;
  const int64_t ch = b->vec_chunk;
  const BatchTable& gt = b->vec_t[0];
  int rc;
  // b = P^H y in FP64 (complex64 samples upcast exactly) into r, x = 0, |b|^2
  if ((rc = pnh_batch_vec_launch(BVEC_CAST_Y, b->probs, b->vec_t[1], ch, (void*)y, w.yt,
                                 nullptr, nullptr, cg, 0, s))) return rc;
  if ((rc = batch_y2k(b, 1, w.yt, w.r, w.part, nullptr, nullptr, s))) return rc;
  if ((rc = pnh_batch_vec_launch(BVEC_ZERO_X, b->probs, gt, ch, w.x, nullptr, nullptr,
                                 nullptr, cg, 0, s))) return rc;
  if ((rc = pnh_batch_vec_launch(BVEC_DOT_RR, b->probs, gt, ch, w.r, nullptr, nullptr,
                                 nullptr, cg, 0, s))) return rc;
  for (int it = 0; it < maxiter; ++it) {
    // the p update finishes the previous iteration's |r|^2 and stopping test
    const int* done = cg.done + (it & 1) * b->n;
    if ((rc = pnh_batch_vec_launch(BVEC_P_UPDATE, b->probs, gt, ch, w.p, w.r, nullptr,
                                   nullptr, cg, it, s))) return rc;
    if ((rc = pnh_batch_k2y2k_launch(b->probs, b->items, b->n_items, b->gather_t, b->D, b->J,
                                     1, b->pad_cells_max, b->grid_threads, b->fixed_pad, w.p,
                                     w.yt, w.q, w.part, done, nullptr, s))) return rc;
    if ((rc = pnh_batch_vec_launch(BVEC_DOT_PQ, b->probs, gt, ch, w.p, w.q, nullptr,
                                   nullptr, cg, it, s))) return rc;
    if ((rc = pnh_batch_vec_launch(BVEC_XR_UPDATE, b->probs, gt, ch, w.x, w.r, w.p,
                                   w.q, cg, it, s))) return rc;
  }
  // image = crop(ifftn(k)) in FP64 into q (packed image layout), then
  // x = complex64(image) / sn
  if ((rc = batch_passes(b, BFFT_INV64, 1, nullptr, w.q, w.x, nullptr, nullptr, s))) return rc;
  if ((rc = pnh_batch_vec_launch(BVEC_CAST_DIV, b->probs, b->vec_t[2], ch, w.q, x, nullptr,
                                 nullptr, cg, 0, s))) return rc;
  return pnh_batch_cg_out_launch(gt, cg, iters_out, resid_out, s);
}

}  // namespace

extern "C" {

int pnh_batch_table_bytes(const pnh_plan_desc* descs, int n, size_t* bytes) {
  REQUIRE(bytes, "null argument");
  pnh_batch b;
  int rc = build_batch(descs, n, nullptr, nullptr, &b);
  if (rc) return rc;
  *bytes = b.table_bytes;
  return 0;
}

int pnh_batch_create(const pnh_plan_desc* descs, int n, void* table, void* stream,
                     pnh_batch** out) {
  REQUIRE(out && table, "null argument");
  *out = nullptr;
  pnh_batch* b = new (std::nothrow) pnh_batch;
  if (!b) { pnh_set_error("out of host memory"); return -1; }
  int rc = build_batch(descs, n, table, (hipStream_t)stream, b);
  if (rc) { delete b; return rc; }
  *out = b;
  return 0;
}

int pnh_batch_destroy(pnh_batch* b) {
  delete b;
  return 0;
}

int pnh_batch_sizes(const pnh_batch* b, int64_t* n_image, int64_t* n_samples, int64_t* n_grid) {
  REQUIRE(b && n_image && n_samples && n_grid, "null argument");
  *n_image = b->N_total;
  *n_samples = b->M_total;
  *n_grid = b->K_total;
  return 0;
}

int pnh_batch_workspace_bytes(const pnh_batch* b, int op, int dtype, size_t* bytes) {
  REQUIRE(b && bytes, "null argument");
  (void)dtype;
  const size_t grid = align_up((size_t)b->K_total * 8);
  const size_t part = align_up((size_t)b->part_elems * 8);
  switch (op) {
    case PNH_OP_FORWARD: *bytes = grid; break;
    case PNH_OP_ADJOINT: *bytes = grid + part; break;
    case PNH_OP_SOLVE_CG: { BatchWs w; *bytes = batch_cg_layout(b, nullptr, &w); break; }
    default: pnh_set_error("op %d is not provided for batches", op); return -1;
  }
  return 0;
}

int pnh_batch_forward(const pnh_batch* b, const void* x, void* y, void* ws, void* stream) {
  REQUIRE(b && x && ws && (y || b->M_total == 0), "null argument");
  hipStream_t s = (hipStream_t)stream;
  int rc = batch_passes(b, BFFT_FWD, 0, x, nullptr, ws, nullptr, nullptr, s);
  if (rc) return rc;
  return pnh_batch_k2y_launch(b->probs, b->items, b->n_items, b->D, b->J, 0,
                              b->pad_cells_max, b->grid_threads, ws, y, nullptr, nullptr, s);
}

int pnh_batch_adjoint(const pnh_batch* b, const void* y, void* x, void* ws, void* stream) {
  REQUIRE(b && x && ws && (y || b->M_total == 0), "null argument");
  hipStream_t s = (hipStream_t)stream;
  char* w = (char*)ws;
  void* grid = w;
  void* part = w + align_up((size_t)b->K_total * 8);
  int rc = batch_y2k(b, 0, y, grid, part, nullptr, nullptr, s);
  if (rc) return rc;
  return batch_passes(b, BFFT_INV, 0, nullptr, x, grid, nullptr, nullptr, s);
}

int pnh_batch_solve_cg(const pnh_batch* b, const void* y, void* x, int maxiter, double rtol,
                       double atol, void* ws, int32_t* iters_out, double* resid_out,
                       void* stream) {
  REQUIRE(b && x && ws && (y || b->M_total == 0), "null argument");
  REQUIRE(maxiter >= 0, "maxiter must be >= 0");
  return batch_solve(b, y, x, maxiter, rtol, atol, ws, iters_out, resid_out,
                     (hipStream_t)stream);
}

int pnh_batch_cg_graph_create(const pnh_batch* b, const void* y, void* x, int maxiter,
                              double rtol, double atol, void* ws, int32_t* iters_out,
                              double* resid_out, pnh_graph** out) {
  REQUIRE(b && x && ws && out, "null argument");
  *out = nullptr;
  hipStream_t cs;
  PNH_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(cs);
    pnh_set_error("hipStreamBeginCapture: %s", hipGetErrorString(e));
    return -3;
  }
  const int rc = batch_solve(b, y, x, maxiter, rtol, atol, ws, iters_out, resid_out, cs);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(cs, &g);
  (void)hipStreamDestroy(cs);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) {
    pnh_set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
    return -3;
  }
  pnh_graph* gr = new (std::nothrow) pnh_graph;
  if (!gr) { (void)hipGraphDestroy(g); pnh_set_error("out of host memory"); return -1; }
  e = hipGraphInstantiate(&gr->exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    delete gr;
    pnh_set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
    return -3;
  }
  *out = gr;
  return 0;
}

int pnh_cg_graph_create(const pnh_plan* p, const void* y, void* x, int maxiter, double rtol,
                        double atol, void* ws, int32_t* iters_out, double* resid_out,
                        pnh_graph** out) {
  REQUIRE(p && x && ws && out, "null argument");
  *out = nullptr;
  hipStream_t cs;
  PNH_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(cs);
    pnh_set_error("hipStreamBeginCapture: %s", hipGetErrorString(e));
    return -3;
  }
  const int rc = cg_solve(p, y, x, maxiter, rtol, atol, ws, iters_out, resid_out, cs, false);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(cs, &g);
  (void)hipStreamDestroy(cs);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) {
    pnh_set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
    return -3;
  }
  pnh_graph* gr = new (std::nothrow) pnh_graph;
  if (!gr) { (void)hipGraphDestroy(g); pnh_set_error("out of host memory"); return -1; }
  e = hipGraphInstantiate(&gr->exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    delete gr;
    pnh_set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
    return -3;
  }
  *out = gr;
  return 0;
}

int pnh_graph_launch(pnh_graph* g, void* stream) {
  REQUIRE(g, "null argument");
  PNH_CHECK_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return 0;
}

int pnh_graph_destroy(pnh_graph* g) {
  if (g) {
    (void)hipGraphExecDestroy(g->exec);
    delete g;
  }
  return 0;
}

}  // extern "C"
