"""Gradients of the differentiable forward NUFFT against FP64 finite
differences.

The finite differences use an FP64 restatement of the forward operator: the
oracle's complex128 CSR rows (which depend only on each sample's own om) and
a complex128 FFT of the apodized, zero-padded image.  The loss is
L = sum_c ||A(om)(S_c x) - y_c||^2 (one coil when S is None).
"""
import numpy as np
import pytest

import oracle
from pynufft_amd import workloads as W

pytestmark = pytest.mark.gpu


def fp64_forward_rows(om_rows, Nd, Kd, Jd, xs):
    """y[m, c] for the given sample rows, FP64 end to end."""
    A = oracle.OracleNUFFT()
    A.plan(om_rows, Nd, Kd, Jd)
    sn = A.st['sn']
    out = []
    for c in range(xs.shape[-1]):
        g = np.zeros(Kd, dtype=np.complex128)
        g[tuple(slice(0, n) for n in Nd)] = xs[..., c] * sn
        out.append(A.sp.dot(np.fft.fftn(g).ravel()))
    return np.stack(out, axis=-1)


def loss64(om, Nd, Kd, Jd, xs, y0):
    y = fp64_forward_rows(om, Nd, Kd, Jd, xs)
    return float(np.sum(np.abs(y - y0) ** 2))


def far_from_jumps(om, Kd, Jd, h):
    ok = np.ones(om.shape[0], bool)
    for d in range(om.shape[1]):
        gam = 2 * np.pi / Kd[d]
        t = om[:, d] / gam - Jd[d] / 2.0
        frac = t - np.floor(t)
        ok &= (frac > 20 * h / gam) & (frac < 1 - 20 * h / gam)
    return ok


@pytest.mark.parametrize('coils', [None, 3])
def test_gradients_match_finite_differences(hip, coils):
    import torch
    from pynufft_amd.autograd import nufft_forward
    from pynufft_amd.nufft import NUFFT
    Nd, Kd, Jd = (12, 10), (24, 20), (6, 5)
    M = 60
    om = W.uniform_om(M, 2, 801)
    nb = coils or 1
    x = W.gaussian_image(Nd, 802).astype(np.complex128)
    rng = np.random.default_rng(803)
    S = (rng.standard_normal(Nd + (nb,)) + 1j * rng.standard_normal(Nd + (nb,))) * 0.5 + 1
    if not coils:
        S = np.ones(Nd + (1,), dtype=np.complex128)
    y0 = (rng.standard_normal((M, nb)) + 1j * rng.standard_normal((M, nb)))

    A = NUFFT()
    A.plan(om, Nd, Kd, Jd, batch=coils)
    dev = A.device
    xt = torch.tensor(x.astype(np.complex64), device=dev, requires_grad=True)
    omt = torch.tensor(om.astype(np.float32), device=dev, requires_grad=True)
    St = torch.tensor(S.astype(np.complex64), device=dev)
    y0t = torch.tensor(y0.astype(np.complex64), device=dev)
    xs = xt[..., None] * St if coils else xt
    y = nufft_forward(A, xs, omt)
    if not coils:
        y = y[:, None]
    loss = (y - y0t).abs().pow(2).sum()
    loss.backward()
    gom = omt.grad.cpu().numpy().astype(np.float64)
    gx = xt.grad.cpu().numpy().astype(np.complex128)

    xs64 = x[..., None] * S
    # trajectory gradient: per-sample central differences (rows independent)
    h = 1e-6
    ok = far_from_jumps(om, Kd, Jd, h)
    assert ok.sum() > M // 2
    fd = np.zeros_like(gom)
    for m in np.nonzero(ok)[0]:
        for d in range(2):
            op = om[m:m + 1].copy()
            om_ = om[m:m + 1].copy()
            op[0, d] += h
            om_[0, d] -= h
            fd[m, d] = (loss64(op, Nd, Kd, Jd, xs64, y0[m:m + 1]) -
                        loss64(om_, Nd, Kd, Jd, xs64, y0[m:m + 1])) / (2 * h)
    err = np.linalg.norm(gom[ok] - fd[ok]) / np.linalg.norm(fd[ok])
    assert err < 1e-4, err

    # image gradient: torch convention x.grad = dL/dRe + i dL/dIm
    hx = 1e-5
    for idx in [(0, 0), (5, 7), (11, 9)]:
        for part, unit in ((0, 1.0), (1, 1j)):
            xp = x.copy()
            xm = x.copy()
            xp[idx] += hx * unit
            xm[idx] -= hx * unit
            lp = loss64(om, Nd, Kd, Jd, xp[..., None] * S, y0)
            lm = loss64(om, Nd, Kd, Jd, xm[..., None] * S, y0)
            g_fd = (lp - lm) / (2 * hx)
            g_ad = gx[idx].real if part == 0 else gx[idx].imag
            assert abs(g_ad - g_fd) <= 1e-4 * (abs(g_fd) + np.abs(gx).max()), (idx, part)


def test_image_gradient_is_scaled_adjoint(hip):
    """dL/dx for L = Re<v, A x> is A^H v = prod(Kd) * adjoint(v)."""
    import torch
    from pynufft_amd.autograd import nufft_forward
    from pynufft_amd.nufft import NUFFT
    Nd, Kd, Jd = (16, 16), (32, 32), (6, 6)
    om = W.uniform_om(300, 2, 811)
    A = NUFFT()
    A.plan(om, Nd, Kd, Jd)
    xt = torch.tensor(W.gaussian_image(Nd, 812), device=A.device, requires_grad=True)
    v = torch.tensor(W.gaussian_image((300,), 813), device=A.device)
    y = nufft_forward(A, xt)
    (torch.vdot(v, y).real).backward()
    ref = A.adjoint(v) * float(np.prod(Kd))
    assert torch.allclose(xt.grad, ref, rtol=1e-5, atol=1e-5 * ref.abs().max().item())
