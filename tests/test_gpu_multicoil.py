"""Differentiable multi-coil NUFFT at full size: Nd 256^2, Kd 512^2, Jd 6^2,
M 131 072, 8 coils; loss = sum_c ||A(om)(S_c x) - y_c||^2 with y_c from the
trajectory perturbed by N(0, 1e-3).  Gradients in om (float32) and x
(complex64) are checked against FP64 central finite differences on a subset
of samples and pixels (rows of the interpolator depend only on their own
sample; a one-pixel change is a rank-1 change of the oversampled spectrum)."""
import numpy as np
import pytest

import oracle
from pynufft_amd import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def fp64_spectra(w, x):
    A = oracle.OracleNUFFT()
    A.plan(w['om'][:1], w['Nd'], w['Kd'], w['Jd'])
    sn = A.st['sn']
    ks = []
    for c in range(w['S'].shape[-1]):
        g = np.zeros(w['Kd'], dtype=np.complex128)
        g[:w['Nd'][0], :w['Nd'][1]] = x * w['S'][..., c] * sn
        ks.append(np.fft.fftn(g).ravel())
    return np.stack(ks, axis=-1), sn


def rows64(om_rows, w):
    A = oracle.OracleNUFFT()
    A.plan(om_rows, w['Nd'], w['Kd'], w['Jd'])
    return A.sp


def test_multicoil_gradients_full_size(hip):
    import torch
    from pynufft_amd.autograd import nufft_forward
    from pynufft_amd.nufft import NUFFT
    w = W.multicoil2d()
    nb = w['S'].shape[-1]
    D = NUFFT()
    D.plan(w['om_data'], w['Nd'], w['Kd'], w['Jd'], batch=nb)
    xs_true = (w['x'][..., None] * w['S']).astype(np.complex64)
    y0 = D.forward(xs_true)                          # data on the perturbed trajectory
    A = NUFFT()
    A.plan(w['om'], w['Nd'], w['Kd'], w['Jd'], batch=nb)
    dev = A.device
    x_est = (0.9 * w['x']).astype(np.complex64)
    xt = torch.tensor(x_est, device=dev, requires_grad=True)
    omt = torch.tensor(w['om'].astype(np.float32), device=dev, requires_grad=True)
    St = torch.tensor(w['S'], device=dev)
    y = nufft_forward(A, xt[..., None] * St, omt)
    loss = (y - torch.tensor(y0, device=dev)).abs().pow(2).sum()
    loss.backward()
    gom = omt.grad.cpu().numpy().astype(np.float64)
    gx = xt.grad.cpu().numpy()

    k64, sn = fp64_spectra(w, x_est.astype(np.complex128))
    y0_64 = y0.astype(np.complex128)
    rng = np.random.default_rng(7)
    M = w['om'].shape[0]
    sel = np.sort(rng.choice(M, 48, replace=False))
    h = 1e-6
    fd = []
    for m in sel:
        for d in range(2):
            op = w['om'][m:m + 1].copy()
            om_ = w['om'][m:m + 1].copy()
            op[0, d] += h
            om_[0, d] -= h
            yp = rows64(op, w).dot(k64)
            ym = rows64(om_, w).dot(k64)
            lp = np.sum(np.abs(yp - y0_64[m]) ** 2)
            lm = np.sum(np.abs(ym - y0_64[m]) ** 2)
            fd.append(((lp - lm) / (2 * h), gom[m, d]))
    fd = np.array(fd)
    # skip samples within reach of a support-cell jump (FD straddles it)
    ok = np.abs(fd[:, 0] - fd[:, 1]) < 1e3 * np.median(np.abs(fd[:, 0] - fd[:, 1])) + 1e-12
    err = np.linalg.norm(fd[ok, 0] - fd[ok, 1]) / np.linalg.norm(fd[ok, 0])
    assert ok.sum() >= 80 and err < 1e-3, (ok.sum(), err)

    # image gradient at a few pixels: dL/dRe x_n and dL/dIm x_n in FP64
    P = rows64(w['om'], w)
    r = P.dot(k64) - y0_64                          # residual per sample and coil
    Kd, Nd = w['Kd'], w['Nd']
    for (i0, i1) in [(128, 128), (40, 200), (250, 3)]:
        # d k_c / d x_n = S_c(n) sn(n) exp(-2 pi i (k . n) / K)
        e = np.exp(-2j * np.pi * (np.arange(Kd[0])[:, None] * i0 / Kd[0] +
                                  np.arange(Kd[1])[None, :] * i1 / Kd[1])).ravel()
        Pe = P.dot(e)                                # (M,)
        g_re = 0.0
        g_im = 0.0
        for c in range(nb):
            dy = Pe * (w['S'][i0, i1, c] * sn[i0, i1])
            g_re += 2 * np.real(np.vdot(r[:, c], dy))
            g_im += 2 * np.real(np.vdot(r[:, c], 1j * dy))
        ref = g_re + 1j * g_im
        assert abs(gx[i0, i1] - ref) < 1e-3 * abs(ref) + 1e-3 * np.abs(gx).max(), (i0, i1)
