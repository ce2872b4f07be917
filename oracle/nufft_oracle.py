"""numpy/scipy restatement of pynufft's CPU NUFFT path (oracle; test use only).

Every function follows one piece of the reference and cites it.  The dtype
behaviour of the reference is kept on purpose, because it decides the bits of
the result:

* the alpha fit runs ``lstsq`` in complex64 (reference ``helper.py:8,947-949``);
* the interpolation matrix is complex128, so ``k2y``/``y2k`` accumulate in FP64;
* ``fftn`` of a complex64 grid stays complex64 (numpy >= 2, pocketfft);
* the CG solve runs in complex128 through ``scipy.sparse.linalg.cg``.

Pinning (checks that do not restate the formula):

=====================  ======================================================
function               pinned by
=====================  ======================================================
kb_shape / kb_ft       golden alpha/beta/sn from the reference (tests/golden)
fit_alpha_beta         golden alpha (complex64 lstsq bits) + J=1 closed form
scale_factor_1d        golden sn_d; J=1 closed form sn == 1
interp_T / interp_r    golden CSR data (tiny 1D/2D/3D cases) + J=1 closed form
interp_u / interp_cols golden CSR data/indices; grid-aligned J=1 == NUDFT
plan (CSR)             golden CSR of the reference at several shapes
OracleNUFFT stages     golden forward/adjoint/selfadjoint/solve outputs;
                       adjoint identity; FFT vs brute-force DFT
cg_scipy               golden 30-iteration image; dense FP64 solve (tiny)
OracleNUDFT            brute-force triple loop on tiny inputs
laplacian_spectrum     closed form sum_d 2 cos(2 pi k_d / K_d) - 2
precompute_sp,         golden W and selfadjoint2 of the reference
  selfadjoint2         (tests/golden/solvers.npz)
solve 'dc', 'L1TVOLS', golden outputs of the reference's solve for each
  'lsqr', 'lsmr',      solver on four shapes (tests/golden/solvers.npz)
  'bicgstab', 'bicg',
  'gmres', 'lgmres'
=====================  ======================================================
"""
import numpy as np
import scipy.linalg
import scipy.sparse
import scipy.sparse.linalg
import scipy.special

_C64 = np.complex64

# Kaiser-Bessel shape table used when K/N == 2 (reference helper.py:973-988).
_KB_BEST = {2: 2.5, 3: 2.27, 4: 2.31, 5: 2.34, 6: 2.32, 7: 2.32, 8: 2.35,
            9: 2.34, 10: 2.34, 11: 2.35, 12: 2.34, 13: 2.35, 14: 2.35,
            15: 2.35, 16: 2.33}


def kb_shape(J, K_over_N):
    """KB shape parameter (alpha, m).  Reference helper.py:961-997.

    K/N != 2 -> 2.34*J; K/N == 2 -> J*table[J].  For J outside the table the
    reference evaluates ``dict_keys - ndarray`` and raises TypeError
    (helper.py:991-995); the same error is raised here.
    """
    if K_over_N != 2:
        return 2.34 * J, 0
    if J in _KB_BEST:
        return J * _KB_BEST[J], 0
    raise TypeError("unsupported operand type(s) for -: 'dict_keys' and "
                    "'float' (reference kaiser_bessel, J=%r not tabulated)" % J)


def kb_ft(u, J, a, m, d):
    """Fourier transform of the KB window.  Reference helper.py:1000-1012."""
    u = u * (1.0 + 0.0j)
    z = np.sqrt((2 * np.pi * (J / 2) * u) ** 2.0 - a ** 2.0)
    nu = d / 2 + m
    y = ((2 * np.pi) ** (d / 2)) * ((J / 2) ** d) * (a ** m) / \
        scipy.special.iv(m, a) * scipy.special.jv(nu, z) / (z ** nu)
    return np.real(y)


def fit_alpha_beta(N, J, K):
    """Cosine-series fit of 1/KB_ft.  Reference helper.py:910-958.

    Returns alpha as the reference does: a float32 array of shape (L+1, 1)
    (complex64 lstsq, real part), and beta = 1.
    """
    beta = 1
    Nmid = (N - 1.0) / 2.0
    L = 13 if N > 40 else int(np.ceil(N / 3))
    nlist = np.arange(0, N) * 1.0 - Nmid
    kb_a, kb_m = kb_shape(J, K / N)
    if J > 1:
        target = 1 / kb_ft(nlist / K, J, kb_a, kb_m, 1.0)
    else:
        target = np.ones((1, N), dtype=_C64)
    gam = 2 * np.pi / K
    X = np.cos(np.dot((beta * gam * nlist).reshape((N, 1)),
                      np.arange(0, L + 1).reshape((1, L + 1))))
    X = np.array(X, dtype=_C64)
    rhs = np.array(target.reshape((N, 1), order='F').conj(), dtype=_C64)
    coef = np.linalg.lstsq(np.nan_to_num(X), np.nan_to_num(rhs), rcond=-1)[0]
    if J > 1:
        coef[1:] = coef[1:] / 2.0
    else:
        coef[0] = 1.0
        coef[1:] = 0.0
    return np.real(coef), beta


def scale_factor_1d(N, K, alpha, beta):
    """1-D apodization s_d[n] = sum_l alpha_|l| e^{i gam (n-Nmid) beta l}.

    Reference helper.py:1015-1050 (nufft_scale1 via nufft_scale, dd == 1).
    Returns complex128 (N, 1).
    """
    Nmid = (N - 1) / 2.0
    alpha = np.real(alpha)
    L = len(alpha) - 1
    if L == 0:
        return np.dot(alpha, np.ones((N, 1)))
    sn = np.zeros((N, 1))
    n = np.arange(0, N).reshape((N, 1), order='F')
    ph = 1j * (2 * np.pi / K) * (n - Nmid) * beta
    for l1 in range(-L, L + 1):
        sn = sn + alpha[abs(l1)] * np.exp(ph * l1)
    return sn


def interp_T(N, J, K, alpha, beta):
    """pinv of the J x J overlap matrix.  Reference helper.py:1053-1083.

    alpha entries are float32 (1,)-arrays, so alpha*alpha is a float32
    product before it meets the float64 sinc matrix, as in the reference.
    """
    L = np.size(alpha) - 1
    cssc = np.zeros((J, J))
    j1, j2 = np.mgrid[1:J + 1, 1:J + 1]
    diff = j2 - j1
    for l1 in range(-L, L + 1):
        for l2 in range(-L, L + 1):
            w = alpha[abs(l1)] * alpha[abs(l2)]
            cssc = cssc + w * np.sinc(1.0 * (diff + beta * (l1 - l2)) /
                                      (1.0 * K / N))
    return scipy.linalg.pinv(cssc)


def offset(om, J, K):
    """koff = floor(om/gam - J/2).  Reference helper.py:900-907."""
    gam = 2.0 * np.pi / (K * 1.0)
    return np.floor(1.0 * om / gam - 1.0 * J / 2.0)


def interp_r(om, N, J, K, alpha, beta):
    """r[j,m] = sum_l alpha_|l| sinc((arg+l beta)/(K/N)).  helper.py:1086-1117."""
    gam = 2.0 * np.pi / (K * 1.0)
    dk = 1.0 * om / gam - offset(om, J, K)
    arg = np.add.outer(-np.arange(1, J + 1) * 1.0, dk)
    L = np.size(alpha) - 1
    rr = np.zeros((J, np.size(om)), dtype=np.float32)
    ratio = 1.0 * K / N
    for l1 in range(-L, L + 1):
        rr = rr + (alpha[abs(l1)] * 1.0) * np.sinc((arg + 1.0 * l1 * beta) / ratio)
    return rr, arg


def interp_u(om, N, J, K, alpha, beta):
    """Per-dimension min-max coefficients u_d (M, J), complex128.

    c = T r (helper.py:606-618); u = conj((e^{-i om N/2} e^{i gam (N-1)/2 arg}
    c)^T) (OMEGA_u, helper.py:148-162).
    """
    T = interp_T(N, J, K, alpha, beta)
    r, arg = interp_r(om, N, J, K, alpha, beta)
    c = T.dot(r)
    gam = 2.0 * np.pi / (K * 1.0)
    u = np.exp(1.0j * gam * (N * 1.0 - 1.0) / 2.0 * arg) * c
    u = np.exp(-1.0j * om * N / 2.0) * u
    return u.T.conj()


def interp_cols(om, J, K, stride):
    """Column offsets mod(koff + j, K) * stride, (M, J) float64.

    Reference OMEGA_k, helper.py:164-210 (ft axis branch).
    """
    k = np.mod(np.add.outer(np.arange(1, J + 1) * 1.0, offset(om, J, K)), K)
    return (k * stride).T


def plan(om, Nd, Kd, Jd):
    """FP64 plan with a CSR interpolator.  Reference helper.py:620-802.

    Returns a dict with alpha/beta per dimension, sn_d (complex128, (N_d,1)),
    sn (real FP64, Nd), u_d, k_d and the CSR matrix p (M x prod(Kd)).
    """
    if type(Nd) != tuple:
        raise TypeError('Nd must be tuple, e.g. (256, 256)')
    if type(Kd) != tuple:
        raise TypeError('Kd must be tuple, e.g. (512, 512)')
    if type(Jd) != tuple:
        raise TypeError('Jd must be tuple, e.g. (6, 6)')
    if len(Nd) != len(Kd) or len(Nd) != len(Jd):
        raise KeyError('Nd, Kd, Jd must be in the same length, e.g. '
                       'Nd=(256,256),Kd=(512,512),Jd=(6,6)')
    dd = len(Nd)
    M = om.shape[0]
    st = {'Nd': Nd, 'Kd': Kd, 'Jd': Jd, 'M': np.int32(M), 'om': om,
          'alpha': [], 'beta': []}
    for d in range(dd):
        a, b = fit_alpha_beta(Nd[d], Jd[d], Kd[d])
        st['alpha'].append(a)
        st['beta'].append(b)
    snd = [scale_factor_1d(Nd[d], Kd[d], st['alpha'][d], st['beta'][d])
           for d in range(dd)]
    ud = [interp_u(om[:, d], Nd[d], Jd[d], Kd[d], st['alpha'][d], st['beta'][d])
          for d in range(dd)]
    kd = [interp_cols(om[:, d], Jd[d], Kd[d], int(np.prod(Kd[d + 1:])))
          for d in range(dd)]
    # Khatri-Rao products, j in C order (helper.py:449-480)
    uu, kk = ud[0], kd[0]
    for d in range(1, dd):
        uu = np.einsum('mi,mj->mij', uu, ud[d]).reshape(M, -1)
        kk = (kk[:, :, None] + kd[d][:, None, :]).reshape(M, -1)
    P = int(np.prod(Jd))
    # create_csr (helper.py:285-311): fixed row length P, unsorted columns
    st['p'] = scipy.sparse.csr_matrix(
        (uu.ravel(), kk.ravel().astype(np.int32), np.arange(0, (M + 1) * P, P)),
        shape=(M, int(np.prod(Kd))))
    # kronecker_scale(snd).real (helper.py:558-584)
    sn = np.reshape(1.0, (1,) * dd)
    for d in range(dd):
        shp = [1] * dd
        shp[d] = Nd[d]
        sn = sn * np.reshape(snd[d], shp)
    st['sn'] = sn.real
    st['snd'] = snd
    st['ud'] = ud
    st['kd'] = kd
    return st


def pad_index_maps(Nd, Kd):
    """Front-corner copy maps between Nd and Kd.  Reference helper.py:830-881."""
    mn = tuple(min(a, b) for a, b in zip(Nd, Kd))
    idx = np.indices(mn).reshape(len(mn), -1)
    return (np.ravel_multi_index(idx, Nd).astype(np.int32),
            np.ravel_multi_index(idx, Kd).astype(np.int32))


def cg_scipy(matvec, b, maxiter, rtol=1e-5, atol=0.0):
    """scipy.sparse.linalg.cg on a complex128 LinearOperator.

    Reference linalg/solve_cpu.py:256-288 (LinearOperator + scipy cg).  The
    operator dtype is complex128 there because the CSR is complex128.
    """
    n = b.shape[0]
    A = scipy.sparse.linalg.LinearOperator((n, n), matvec=matvec,
                                           rmatvec=matvec, dtype=np.complex128)
    return scipy.sparse.linalg.cg(A, b, maxiter=maxiter, rtol=rtol, atol=atol)


def laplacian_spectrum(Kd):
    """fftn of the periodic discrete Laplacian stencil on the Kd grid: -2d at
    the origin, 1 at the two neighbours along each axis, in complex64.
    Reference src/_helper/helper.py:11-45 (create_laplacian_kernel)."""
    d = len(Kd)
    stencil = np.zeros(Kd, dtype=_C64)
    origin = (0,) * d
    stencil[origin] = -2.0 * d
    for p in range(d):
        for step in (1, -1):
            at = list(origin)
            at[p] = step
            stencil[tuple(at)] = 1
    return np.fft.fftn(stencil)


def diff_back(x, p):
    """x[i - e_p] - x[i], periodic (reference solve_cpu.py:11-23 cDiff with
    helper.py:46-70 indxmap_diff's d_indx)."""
    return np.roll(x, 1, axis=p) - x


def diff_fwd(x, p):
    """x[i + e_p] - x[i], periodic (cDiff with indxmap_diff's dt_indx)."""
    return np.roll(x, -1, axis=p) - x


def l1tvols(A, y, maxiter, rho):
    """L1 total-variation regularised least squares by split Bregman.
    Reference linalg/solve_cpu.py:42-163 (L1TVOLS) with mu = 1, lambda = rho:
    each outer iteration solves (mu W - lambda Lap) x = rhs in k-space, takes
    the periodic image gradient, shrinks its magnitude by 1/lambda, updates the
    Bregman variables and the data residual AHyk += AHy - AHA x.  complex64
    arithmetic as numpy runs it there."""
    mu = 1.0
    lam = rho * mu
    uker = mu * A.kspace_density() - lam * laplacian_spectrum(A.Kd)
    AHy = A.adjoint(y)
    x = np.zeros_like(AHy)
    AHyk = np.zeros_like(AHy)
    d = len(A.Nd)
    dd = [np.zeros(A.Nd, dtype=_C64) for _ in range(d)]
    bb = [np.zeros(A.Nd, dtype=_C64) for _ in range(d)]
    for _ in range(maxiter):
        rhs = mu * AHyk
        for p in range(d):
            rhs += lam * diff_fwd(dd[p] - bb[p], p)
        x = A.k2xx((A.xx2k(rhs) + 1e-7) / (uker + 1e-7))
        zz = [diff_back(x, p) for p in range(d)]
        zf = A.selfadjoint(x) - AHy
        mag = np.zeros_like(zz[0])
        t = []
        for p in range(d):
            t.append(zz[p] + bb[p])
            mag = np.hypot(np.hypot(t[p].real, t[p].imag), mag.real)
        mag += 1e-5
        thr = 1 / lam
        r = (mag > thr) * (mag - thr) / mag
        for p in range(d):
            dd[p] = t[p] * r
            bb[p] += zz[p] - dd[p]
        AHyk = AHyk - zf
    return x


class OracleNUFFT:
    """CPU NUFFT with the reference's stage semantics.

    Reference nufft/_nufft_class_methods_cpu.py:4-362 (``_init__cpu``,
    ``_plan_cpu`` and the ``_*_cpu`` stages).
    """

    def __init__(self):
        self.dtype = _C64
        self.Nd = ()
        self.Kd = ()
        self.Jd = ()
        self.ndims = 0
        self.batch = 1

    def plan(self, om, Nd, Kd, Jd):
        """Reference _nufft_class_methods_cpu.py:27-124."""
        self.ndims = len(Nd)
        self.ft_axes = tuple(range(self.ndims))
        self.st = plan(om, Nd, Kd, Jd)
        self.Nd, self.Kd, self.Jd = Nd, Kd, Jd
        self.sn = np.asarray(self.st['sn'].astype(self.dtype), order='C')
        self.M = (self.st['M'],)
        self.sp = self.st['p'].copy().tocsr()
        self.spH = self.st['p'].getH().copy().tocsr()
        self.Kdprod = int(np.prod(Kd))
        self.Jdprod = int(np.prod(Jd))
        self.NdCPUorder, self.KdCPUorder = pad_index_maps(Nd, Kd)
        return 0

    # stages (reference _nufft_class_methods_cpu.py:224-362)
    def x2xx(self, x):
        return x * self.sn

    def xx2x(self, xx):
        return self.x2xx(xx)

    def xx2k(self, xx):
        g = np.zeros(self.Kd, dtype=self.dtype)
        g.ravel()[self.KdCPUorder] = xx.ravel()[self.NdCPUorder]
        return np.fft.fftn(g, axes=self.ft_axes)

    def k2xx(self, k):
        k = np.fft.ifftn(k, axes=self.ft_axes)
        xx = np.zeros(self.Nd, dtype=self.dtype)
        xx.ravel()[self.NdCPUorder] = k.ravel()[self.KdCPUorder]
        return xx

    def k2y(self, k):
        return self.sp.dot(np.reshape(k, (self.Kdprod,)))

    def y2k(self, y):
        return np.reshape(self.spH.dot(y), self.Kd)

    def k2y2k(self, k):
        v = np.reshape(k, (self.Kdprod,))
        return np.reshape(self.spH.dot(self.sp.dot(v)), self.Kd)

    def forward(self, x):
        return self.k2y(self.xx2k(self.x2xx(x)))

    def adjoint(self, y):
        return self.xx2x(self.k2xx(self.y2k(y)))

    def selfadjoint(self, x):
        return self.xx2x(self.k2xx(self.k2y2k(self.xx2k(self.x2xx(x)))))

    def precompute_sp(self):
        """W = |xx2k(adjoint(1))| as (W W*)^0.5 in complex64.  Reference
        _nufft_class_methods_cpu.py:126-145 (_precompute_sp_cpu)."""
        W = self.xx2k(self.adjoint(np.ones(self.st['M'], dtype=_C64)))
        self.W = (W * W.conj()) ** 0.5

    def selfadjoint2(self, x):
        """k2xx(W * xx2k(x)), no apodization.  Reference
        _nufft_class_methods_cpu.py:216-222 (_selfadjoint2_cpu)."""
        return self.k2xx(self.W * self.xx2k(x))

    def kspace_density(self):
        """|xx2k(adjoint(1))|.  Reference solve_cpu.py:25-40."""
        return np.abs(self.xx2k(self.adjoint(np.ones(self.st['M'], dtype=_C64))))

    def pipe_density(self, maxiter):
        """W <- W / forward(adjoint(W)) from W = 1.  Reference
        solve_cpu.py:165-183 (_pipe_density)."""
        W = np.ones(self.st['M'], dtype=_C64)
        for _ in range(maxiter):
            W = W / self.forward(self.adjoint(W))
        return W

    def solve(self, y, solver=None, *args, **kwargs):
        """Reference linalg/solve_cpu.py:185-288: 'dc' (density-compensated
        adjoint), 'L1TVOLS', 'lsqr'/'lsmr' on P (k-space unknowns), and the
        scipy Krylov solvers on P^H P; results are crop(ifftn(k)) / sn."""
        if solver == 'dc':
            return self.adjoint(self.pipe_density(*args, **kwargs) * y)
        if solver == 'L1TVOLS':
            return l1tvols(self, y, *args, **kwargs)
        LO = scipy.sparse.linalg.LinearOperator
        if solver in ('lsqr', 'lsmr'):
            A = LO((self.st['M'], self.Kdprod), dtype=np.complex128,
                   matvec=lambda k: self.k2y(k.reshape(self.Kd)).ravel(),
                   rmatvec=lambda v: self.y2k(v.reshape(self.st['M'])).ravel())
            fn = {'lsqr': scipy.sparse.linalg.lsqr, 'lsmr': scipy.sparse.linalg.lsmr}[solver]
            k = fn(A, y.flatten(), *args, **kwargs)[0]
        else:
            fns = {'cg': scipy.sparse.linalg.cg, 'bicgstab': scipy.sparse.linalg.bicgstab,
                   'bicg': scipy.sparse.linalg.bicg, 'gmres': scipy.sparse.linalg.gmres,
                   'lgmres': scipy.sparse.linalg.lgmres}
            if solver not in fns:
                raise KeyError('unknown solver %r' % (solver,))

            def spHsp(v):
                return self.k2y2k(v.reshape(self.Kd)).ravel()
            A = LO((self.Kdprod, self.Kdprod), matvec=spHsp, rmatvec=spHsp,
                   dtype=np.complex128)
            k = fns[solver](A, self.y2k(y).ravel(), *args, **kwargs)[0]
        xx = self.k2xx(k.reshape(self.Kd))
        return xx / self.sn


class OracleNUDFT:
    """FP64 direct NUDFT.  Reference linalg/nudft_cpu.py:27-103.

    forward: y[m] = sum_n x[n] prod_d exp(-i om[m,d] (n_d - N_d/2));
    adjoint: F^H y / prod(Nd).
    """

    def plan(self, om, Nd):
        self.Nd = Nd
        self.F = []
        for d in range(len(Nd)):
            n = np.arange(Nd[d]) - Nd[d] / 2
            self.F.append(np.exp(-1.0j * np.multiply.outer(om[:, d], n)))
        self.scale = np.prod(Nd)
        return 0

    def forward(self, x):
        y = x.astype(np.complex128)
        # contract the last image axis first: y[m, i0..] -> per-sample sums
        t = np.einsum('...j,mj->m...', y, self.F[-1])
        for d in range(len(self.Nd) - 2, -1, -1):
            t = np.einsum('m...j,mj->m...', t, self.F[d])
        return t

    def adjoint(self, y):
        d = len(self.Nd)
        letters = 'abcdefgh'[:d]
        spec = 'm,' + ','.join('m' + c for c in letters) + '->' + letters
        x = np.einsum(spec, y, *[f.conj() for f in self.F], optimize=True)
        return x / self.scale
