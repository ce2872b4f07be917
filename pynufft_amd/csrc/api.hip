// extern "C" entry points of libpynufft_hip.so (see include/pynufft_hip.h).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <new>

#include "../../include/pynufft_hip.h"
#include "common.h"
#include "fft.h"
#include "interp.h"
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

static const size_t kLdsBudget = 65536;

struct pnh_plan {
  pnh_plan_desc d;
  int D, nb;
  int64_t M, Ntot, Ktot;
  int64_t N[PNH_MAXD], K[PNH_MAXD];
  int J[PNH_MAXD];
  BinArgs bins;
  int slot_nbc;            // coils per y2k launch chunk used to size partial slots
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

size_t partial_bytes(const pnh_plan* p, int fp64) {
  const size_t esz = fp64 ? 16 : 8;
  return (size_t)p->d.n_slots * p->bins.tile_cells * p->slot_nbc * esz;
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
  f->lds_pad = 1;
  const int64_t cap = (int64_t)(kLdsBudget / (2 * esz));  // complex elements per buffer
  if ((int64_t)(n + f->lds_pad) > cap) f->lds_pad = 0;
  if ((int64_t)(n + f->lds_pad) > cap) {
    pnh_set_error("FFT length %d exceeds the LDS-resident limit (%lld)", n, (long long)cap);
    return -2;
  }
  const int64_t n_lines = n_outer * n_inner;
  int LB;
  if (n_inner >= 8) {
    f->inner_major = 1;
    LB = 16;
    while (LB > 8 && (int64_t)LB * (n + f->lds_pad) > cap) LB >>= 1;
    while (LB > 1 && (int64_t)LB * (n + f->lds_pad) > cap) LB >>= 1;
  } else {
    f->inner_major = 0;
    LB = (int)(cap / (n + f->lds_pad));
    if (LB > 16) LB = 16;
    if (LB < 1) LB = 1;
    // keep at least ~1024 workgroups when the lines allow it
    while (LB > 1 && (n_lines + LB - 1) / LB < 1024) LB >>= 1;
  }
  f->lines_per_block = LB;
  int64_t work = (int64_t)LB * n / 2;
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
  if (p->d.n_slots > 0 && !part) {
    pnh_set_error("workspace required for split gridding tiles");
    return -1;
  }
  return pnh_y2k_launch(ia, p->bins, fp64, y, k, part, skip, s);
}

int do_k2y(const pnh_plan* p, int fp64, const void* k, void* y, const int* skip,
           hipStream_t s) {
  InterpArgs ia = interp_args(p, fp64);
  return pnh_k2y_launch(ia, fp64, k, y, skip, s);
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
  const pnh_plan_desc& d = *desc;
  REQUIRE(d.ndim >= 1 && d.ndim <= PNH_MAXD, "ndim must be 1..%d (got %d)", PNH_MAXD, d.ndim);
  REQUIRE(d.batch >= 1, "batch must be >= 1");
  REQUIRE(d.M >= 0 && d.M < (1LL << 31), "M out of range");
  REQUIRE(d.perm && d.base, "perm/base pointers are required");
  REQUIRE(d.u_c64 || d.u_c128 || d.csr_c64, "no interpolator data");
  REQUIRE(!d.csr_c64 || d.csr_col, "csr_col required with csr_c64");
  REQUIRE(d.sn_f32, "sn_f32 is required");
  pnh_plan* p = new (std::nothrow) pnh_plan;
  if (!p) { pnh_set_error("out of host memory"); return -1; }
  memset(p, 0, sizeof(*p));
  p->d = d;
  p->D = d.ndim;
  p->nb = d.batch;
  p->M = d.M;
  p->Ntot = 1;
  p->Ktot = 1;
  int64_t cells = 1;
  for (int a = 0; a < p->D; ++a) {
    p->N[a] = d.Nd[a];
    p->K[a] = d.Kd[a];
    p->J[a] = d.Jd[a];
    if (d.Nd[a] < 1 || d.Kd[a] < d.Nd[a] || d.Kd[a] >= (1LL << 30) || d.Jd[a] < 1 ||
        d.Jd[a] > d.Kd[a] || d.tile[a] < 1) {
      pnh_set_error("bad geometry on axis %d: Nd=%lld Kd=%lld Jd=%d tile=%d", a,
                    (long long)d.Nd[a], (long long)d.Kd[a], d.Jd[a], d.tile[a]);
      delete p;
      return -1;
    }
    p->Ntot *= d.Nd[a];
    p->Ktot *= d.Kd[a];
    p->bins.tile[a] = d.tile[a];
    p->bins.tgrid[a] = (int)((d.Kd[a] + d.tile[a] - 1) / d.tile[a]);
    cells *= d.tile[a];
    int radix[PNH_FFT_MAX_PASS], ns[PNH_FFT_MAX_PASS], np;
    if (pnh_fft_factor((int)d.Kd[a], radix, ns, &np)) { delete p; return -2; }
    if (!d.tw_c64[a]) { pnh_set_error("tw_c64[%d] missing", a); delete p; return -1; }
  }
  p->bins.tile_cells = (int)cells;
  p->bins.waves_per_block = 4;
  p->bins.lds_budget = kLdsBudget;
  while (p->bins.waves_per_block > 1 &&
         (size_t)p->bins.waves_per_block * cells * 2 * sizeof(double) > kLdsBudget)
    p->bins.waves_per_block >>= 1;
  if ((size_t)cells * 2 * sizeof(double) > kLdsBudget) {
    pnh_set_error("gridding tile of %lld cells exceeds the LDS budget", (long long)cells);
    delete p;
    return -2;
  }
  p->bins.n_items = d.n_items;
  p->bins.items = d.items;
  p->bins.bin_rows = d.bin_rows;
  p->bins.n_merge = d.n_merge;
  p->bins.merge = d.merge;
  p->slot_nbc = p->nb >= 4 ? 4 : (p->nb >= 2 ? 2 : 1);
  REQUIRE(d.n_items > 0 && d.items && (d.bin_rows || d.M == 0),
          "gridding work items are required");
  *out = p;
  return 0;
}

int pnh_plan_destroy(pnh_plan* plan) {
  delete plan;
  return 0;
}

int pnh_workspace_bytes(const pnh_plan* p, int op, int dtype, size_t* bytes) {
  REQUIRE(p && bytes, "null argument");
  const int fp64 = dtype == PNH_C128;
  const size_t esz = fp64 ? 16 : 8;
  const size_t grid = align_up((size_t)p->Ktot * p->nb * esz);
  const size_t ys = align_up((size_t)p->M * p->nb * esz);
  const size_t part = align_up(partial_bytes(p, fp64));
  switch (op) {
    case PNH_OP_FORWARD: *bytes = grid; break;
    case PNH_OP_ADJOINT: *bytes = grid + part; break;
    case PNH_OP_SELFADJOINT: *bytes = grid + ys + part; break;
    case PNH_OP_Y2K: *bytes = part; break;
    case PNH_OP_K2Y2K: *bytes = ys + part; break;
    case PNH_OP_SOLVE_CG: {
      const size_t g2 = align_up((size_t)p->Ktot * p->nb * 16);
      const size_t y2 = align_up((size_t)p->M * p->nb * 16);
      const size_t part2 = align_up(partial_bytes(p, 1));
      const size_t red = align_up((size_t)PNH_RED_BLOCKS * p->nb * 16);
      *bytes = 4 * g2 + y2 + part2 + red + align_up(sizeof(CgScalars));
      break;
    }
    default: pnh_set_error("unknown op %d", op); return -1;
  }
  return 0;
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
  if (p->M > 0 && (rc = do_k2y(p, fp64, k, ytmp, nullptr, s))) return rc;
  return do_y2k(p, fp64, ytmp, k_out, part, nullptr, s);
}

int pnh_forward(const pnh_plan* p, const void* x, void* y, void* ws, void* stream) {
  REQUIRE(p && x && ws && (y || p->M == 0), "null argument");
  int rc = check_dtype_ready(p, PNH_C64);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = run_passes_forward(p, x, ws, true, 0, nullptr, s))) return rc;
  if (p->M == 0) return 0;
  return do_k2y(p, 0, ws, y, nullptr, s);
}

int pnh_adjoint(const pnh_plan* p, const void* y, void* x, void* ws, void* stream) {
  REQUIRE(p && x && ws && (y || p->M == 0), "null argument");
  int rc = check_dtype_ready(p, PNH_C64);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
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
  char* w = (char*)ws;
  void* grid = w;
  void* ytmp = w + align_up((size_t)p->Ktot * p->nb * 8);
  void* part = (char*)ytmp + align_up((size_t)p->M * p->nb * 8);
  if ((rc = run_passes_forward(p, x, grid, true, 0, nullptr, s))) return rc;
  if (p->M > 0 && (rc = do_k2y(p, 0, grid, ytmp, nullptr, s))) return rc;
  if ((rc = do_y2k(p, 0, ytmp, grid, part, nullptr, s))) return rc;
  return run_passes_inverse(p, grid, x_out, true, 0, nullptr, s);
}

int pnh_forward_grad_om(const pnh_plan* p, const void* k, const void* g, float* om_grad,
                        void* stream) {
  REQUIRE(p && k && (g || p->M == 0) && (om_grad || p->M == 0), "null argument");
  REQUIRE(p->d.du_c64 && p->d.u_c64, "plan has no factor derivatives (du_c64)");
  InterpArgs ia = interp_args(p, 0);
  return pnh_k2y_dom_launch(ia, 0, k, g, om_grad, (hipStream_t)stream);
}

int pnh_solve_cg(const pnh_plan* p, const void* y, void* x, int maxiter, double rtol,
                 double atol, void* ws, int32_t* iters_out, void* stream) {
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
  CgScalars* sc = (CgScalars*)w;
  const int* all_done = &sc->all_done;
  int rc;
  const int64_t K = p->Ktot;
  // b = P^H y in FP64 (y is complex64; upcast exactly as scipy does)
  if (p->M > 0 && (rc = pnh_c64_to_c128(y, yt, p->M * nb, s))) return rc;
  if ((rc = do_y2k(p, 1, yt, vr, part, nullptr, s))) return rc;
  if ((rc = pnh_zero(1, vx, K * nb, s))) return rc;
  if ((rc = pnh_cg_dot(1, vr, vr, K, nb, redp, nullptr, s))) return rc;
  if ((rc = pnh_cg_finalize(redp, nb, sc, CG_PHASE_INIT, 0, maxiter, rtol, atol, s))) return rc;
  for (int it = 0; it < maxiter; ++it) {
    if ((rc = pnh_cg_p_update(vp, vr, K, nb, sc, it == 0, s))) return rc;
    if (p->M > 0 && (rc = do_k2y(p, 1, vp, yt, all_done, s))) return rc;
    if ((rc = do_y2k(p, 1, yt, vq, part, all_done, s))) return rc;
    if ((rc = pnh_cg_dot(0, vp, vq, K, nb, redp, all_done, s))) return rc;
    if ((rc = pnh_cg_finalize(redp, nb, sc, CG_PHASE_PQ, it, maxiter, rtol, atol, s))) return rc;
    if ((rc = pnh_cg_xr_update(vx, vr, vp, vq, K, nb, sc, redp, s))) return rc;
    if ((rc = pnh_cg_finalize(redp, nb, sc, CG_PHASE_RR, it, maxiter, rtol, atol, s))) return rc;
  }
  // x = complex64(crop(ifftn(k))) / sn, inverse FFT in FP64 (as numpy does
  // for the complex128 CG result)
  if ((rc = run_passes_inverse(p, vx, vq, false, 1, nullptr, s))) return rc;
  if ((rc = pnh_cast_div_sn(vq, x, p->d.sn_f32, p->Ntot, nb, s))) return rc;
  if (iters_out)
    PNH_CHECK_HIP(hipMemcpyAsync(iters_out, sc->iters, sizeof(int) * nb,
                                 hipMemcpyDeviceToDevice, s));
  return 0;
}

}  // extern "C"
