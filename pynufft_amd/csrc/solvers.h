// Element-wise kernels of the Toeplitz self-adjoint and of the 'dc' and
// 'L1TVOLS' solvers (complex64; host launch wrappers).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// A C-order array of D <= 3 dimensions (n = prod(dim)).
struct ShapeArgs {
  int D;
  int64_t dim[3];
  int64_t n;
};

// p[i] = re + i im
int pnh_fill_c64(void* p, int64_t n, float re, float im, hipStream_t s);
// w[i] /= e[i] (complex, Smith's scaling)
int pnh_cdiv_c64(void* w, const void* e, int64_t n, hipStream_t s);
// complex128 variants: fill, w /= e, out = a (c128) * b (c64)
int pnh_fill_c128(void* p, int64_t n, double re, double im, hipStream_t s);
int pnh_cdiv_c128(void* w, const void* e, int64_t n, hipStream_t s);
int pnh_cmul_c128_c64(const void* a, const void* b, void* out, int64_t n, hipStream_t s);
// out = complex64(in) * sn[i / nb] (complex128 image -> apodized complex64)
int pnh_cast_mul_sn(const void* in, void* out, const float* sn, int64_t n_img, int nb,
                    hipStream_t s);
// out[i] = a[i] * b[i]
int pnh_cmul_c64(const void* a, const void* b, void* out, int64_t n, hipStream_t s);
// w[i] = |k[i * nb]| (first coil)
int pnh_abs_first_coil(const void* k, float* w, int64_t n, int nb, hipStream_t s);
// uker (complex64) = mu w - lam * Lap on the grid sh; Lap = lap (complex64)
// or, with lap null, sum_d 2 cos(2 pi k_d / K_d) - 2
int pnh_l1tv_uker(const float* w, const void* lap, void* uker, const ShapeArgs& sh, float mu,
                  float lam, hipStream_t s);
// k = (k + 1e-7) / (uker + 1e-7), uker complex64
int pnh_l1tv_kdiv(void* k, const void* uker, int64_t n, hipStream_t s);
// rhs = mu AHyk + lam sum_p D+_p(dd_p - bb_p) (periodic forward differences)
int pnh_l1tv_rhs(const void* AHyk, const void* dd, const void* bb, void* rhs,
                 const ShapeArgs& sh, float mu, float lam, hipStream_t s);
// gradient D-_p x, shrinkage by thr, Bregman update of dd/bb, AHyk -= q - AHy
int pnh_l1tv_shrink(const void* x, const void* q, const void* AHy, void* AHyk, void* dd,
                    void* bb, const ShapeArgs& sh, float thr, hipStream_t s);

// ---- LSQR (scipy.sparse.linalg.lsqr semantics), FP64, scalars on the device
#define PNH_LSQR_BLOCKS 512

struct LsqrState {
  double alfa, beta, rhobar, phibar, anorm, acond, ddnorm, res2, xnorm, xxnorm, z, cs2, sn2;
  double bnorm, t1, t2, rho, theta, phi, psi, tau;
  double atol, btol, ctol, damp;
  int itn, iter_lim, istop, done, skip_v;
};

// part[b] = block partials of sum |a|^2 (PNH_LSQR_BLOCKS blocks)
int pnh_lsqr_norm2(const void* a, int64_t n, double* part, const int* skip, hipStream_t s);
// a = b - (*coef) a with partials of |a|^2
int pnh_lsqr_axpby(void* a, const void* b, const double* coef, int64_t n, double* part,
                   const int* skip, hipStream_t s);
// a /= *norm when *norm > 0; copy (optional) = a
int pnh_lsqr_scale(void* a, const double* norm, int64_t n, void* copy, const int* skip,
                   hipStream_t s);
// partials of |w / rho|^2, x += t1 w, w = v + t2 w
int pnh_lsqr_xw(void* x, void* w, const void* v, const LsqrState* st, int64_t n, double* part,
                hipStream_t s);
// state reset with the caller's tolerances (ctol = 1 / conlim, 0 if conlim <= 0)
int pnh_lsqr_init(LsqrState* st, double atol, double btol, double conlim, double damp,
                  int iter_lim, hipStream_t s);
// scalar recurrence step `phase` (see solvers.hip)
int pnh_lsqr_scalar(LsqrState* st, const double* part, int phase, hipStream_t s);

// ---- LSMR (scipy.sparse.linalg.lsmr semantics), FP64, scalars on the device
struct LsmrState {
  double alpha, beta, alphabar, rho, rhobar, cbar, sbar, zetabar, zeta;
  double betadd, betad, rhodold, tautildeold, thetatilde, d;
  double normA2, maxrbar, minrbar, normA, condA, normx, normb, normr;
  double c1, c2, c3;
  double atol, btol, ctol, damp;
  int itn, maxiter, istop, done, skip_v;
};

int pnh_lsmr_init(LsmrState* st, double atol, double btol, double conlim, double damp,
                  int maxiter, hipStream_t s);
// hbar = h + c1 hbar; x += c2 hbar; h = v + c3 h; partials of |x|^2
int pnh_lsmr_hx(void* x, void* h, void* hbar, const void* v, const LsmrState* st, int64_t n,
                double* part, hipStream_t s);
int pnh_lsmr_scalar(LsmrState* st, const double* part, int phase, hipStream_t s);

// ---- BiCGSTAB / BiCG (scipy semantics) on P^H P, FP64, complex scalars
struct KrylovState {
  double2 rho, rho_prev, alpha, omega, beta;
  double rtol, atol, tol, rhotol;
  int stab, it, maxiter, done, info, half, skip_t;
};

int pnh_krylov_init(KrylovState* st, int stab, double rtol, double atol, int maxiter,
                    hipStream_t s);
// block partials of vdot(a0, b0) (and of vdot(a1, b1) when a1 is set), 2 x
// PNH_LSQR_BLOCKS complex128 entries
int pnh_krylov_vdot2(const void* a0, const void* b0, const void* a1, const void* b1, int64_t n,
                     void* part, const int* skip, hipStream_t s);
int pnh_krylov_scalar(KrylovState* st, const void* part, int phase, hipStream_t s);
int pnh_krylov_p(void* p, const void* r, const void* v, void* pt, const void* rt,
                 const KrylovState* st, int64_t n, hipStream_t s);
int pnh_krylov_r(void* r, const void* v, const KrylovState* st, int64_t n, void* part,
                 hipStream_t s);
int pnh_krylov_x(void* x, const void* p, void* r, const void* t, void* rt, const void* qt,
                 const KrylovState* st, int64_t n, hipStream_t s);

// ---- GMRES (scipy semantics: restarted, modified Gram-Schmidt) on P^H P
#define PNH_GMRES_MAXR 64

struct GmresState {
  double2 S[PNH_GMRES_MAXR + 1];
  double2 H[PNH_GMRES_MAXR][PNH_GMRES_MAXR + 1];   // H[col][row], rotated
  double2 H0[PNH_GMRES_MAXR][PNH_GMRES_MAXR + 1];  // as built (LGMRES: A dx)
  double gc[PNH_GMRES_MAXR];
  double2 gs[PNH_GMRES_MAXR];
  double2 y[PNH_GMRES_MAXR];
  double2 q[PNH_GMRES_MAXR + 1];                   // LGMRES: H0 y
  double rtol, atol, tol, bnrm2, h0, vnorm, presid, ptol, ptol_max_factor, rnorm, nx;
  int restart, maxiter, outer, inner_iter, col_end, breakdown, skip, done, info;
  int flex;        // LGMRES inner cycle (scipy _fgmres): no H zeroing, strict res < ptol
};

// flex = 1: LGMRES inner cycles (scipy's _fgmres rules)
int pnh_gmres_init(GmresState* st, int restart, double rtol, double atol, int maxiter,
                   int flex, hipStream_t s);
// scalar step `phase` for column col / basis vector k (see solvers.hip)
int pnh_gmres_scalar(GmresState* st, const void* part, int phase, int col, int k,
                     hipStream_t s);
// y = H^-1 S of the cycle (one thread)
int pnh_gmres_solve(GmresState* st, hipStream_t s);
// w -= H[col][k] v_k
int pnh_gmres_mgs(void* w, const void* vk, const GmresState* st, int col, int k, int64_t n,
                  hipStream_t s);
// out = in / vnorm (skipped once the cycle has ended, except, for LGMRES,
// the basis vector of the column col that ended it; col < 0: a cycle start)
int pnh_gmres_normalize(void* out, const void* in, const GmresState* st, int64_t n, int col,
                        hipStream_t s);
// x += sum_{k <= col_end} y_k V[k]
int pnh_gmres_x(void* x, const void* V, int64_t ldv, const GmresState* st, int64_t n,
                hipStream_t s);
// r = b - t with partials of |r|^2
int pnh_gmres_residual(void* r, const void* b, const void* t, const GmresState* st, int64_t n,
                       void* part, hipStream_t s);

// ---- LGMRES (scipy semantics, augmentation vectors stored with A z)
struct LgmresZ {
  const void* z[8];   // augmentation vectors of the cycle's last columns, oldest first
};
// dx = sum_{c <= col_end} y_c z_c (z_c = V[c] for c < inner_m, else zo.z[c - inner_m]);
// partials of |dx|^2
int pnh_lgmres_dx(void* dx, const void* V, int64_t ldv, int inner_m, const LgmresZ& zo,
                  const GmresState* st, int64_t n, void* part, hipStream_t s);
// ax = sum_{c <= col_end + 1} q_c V[c]; z_new = dx / nx; az_new = ax / nx; x += dx
int pnh_lgmres_store(void* x, const void* dx, const void* V, int64_t ldv, void* z_new,
                     void* az_new, const GmresState* st, int64_t n, hipStream_t s);
