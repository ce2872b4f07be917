"""pnh_solve_cg_noalloc (static device scratch, no allocation) against
pnh_solve_cg with a caller workspace, through the C ABI.

Two solves in a row on the static scratch, the second on a smaller problem
(radial 2-D 256^2 / 512^2, M 205 824, then plain 2-D 64^2 / 128^2, M 4 096),
each compared bit for bit with pnh_solve_cg on the same inputs (same kernels,
same summation order): image, iterations and final residual.  Both plans are
also solved once more on the static scratch after the other, so a call that
read the previous call's leftovers would show.  A solve whose workspace
exceeds the static scratch is refused before anything is launched.
"""
import ctypes

import numpy as np
import pytest

from pynufft_amd import workloads as W

pytestmark = pytest.mark.gpu


def _plan(w):
    from pynufft_amd.nufft import NUFFT
    A = NUFFT()
    A.plan(w['om'], tuple(w['Nd']), tuple(w['Kd']), tuple(w['Jd']))
    A._ensure_fp64()
    return A


def _y(A, w, seed):
    rng = np.random.default_rng(seed)
    y = A.forward(w['x'])
    return (y + 0.01 * np.sqrt(np.mean(np.abs(y) ** 2)) *
            (rng.standard_normal(y.shape) + 1j * rng.standard_normal(y.shape))
            ).astype(np.complex64)


def _solve(A, y, maxiter, rtol, noalloc):
    import torch
    from pynufft_amd import _lib
    from pynufft_amd.nufft import _ptr
    lib = _lib.load()
    yt = torch.from_numpy(y).to(A.device)
    x = torch.full(tuple(A.Nd), complex(np.nan, np.nan), dtype=torch.complex64,
                   device=A.device)
    iters = torch.full((1,), -1, dtype=torch.int32, device=A.device)
    resid = torch.full((1,), -1.0, dtype=torch.float64, device=A.device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(A.device).cuda_stream)
    if noalloc:
        rc = lib.pnh_solve_cg_noalloc(A._plan, _ptr(yt), _ptr(x), int(maxiter), float(rtol),
                                      0.0, _ptr(iters), _ptr(resid), stream)
    else:
        ws = A._workspace(_lib.OP_SOLVE_CG, _lib.C128)
        rc = lib.pnh_solve_cg(A._plan, _ptr(yt), _ptr(x), int(maxiter), float(rtol), 0.0,
                              _ptr(ws), _ptr(iters), _ptr(resid), stream)
    _lib.check(rc)
    torch.cuda.synchronize(A.device)
    return x.cpu().numpy(), int(iters.item()), float(resid.item())


def _same(a, b):
    xa, ia, ra = a
    xb, ib, rb = b
    assert np.all(np.isfinite(xa))
    assert ia == ib and ra == rb, (ia, ib, ra, rb)
    assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32)), \
        np.max(np.abs(xa - xb))


def test_cg_noalloc_matches_workspace_entry(hip):
    from pynufft_amd import _lib
    lib = _lib.load()
    big_w, small_w = W.radial2d(), W.plain2d()
    big, small = _plan(big_w), _plan(small_w)
    need = ctypes.c_size_t()
    _lib.check(lib.pnh_workspace_bytes(big._plan, _lib.OP_SOLVE_CG, _lib.C128,
                                       ctypes.byref(need)))
    print('\nradial 2-D CG scratch: %.1f MiB of %d MiB'
          % (need.value / 2 ** 20, _lib.NOALLOC_SCRATCH_BYTES >> 20))
    assert need.value <= _lib.NOALLOC_SCRATCH_BYTES
    yb, ys = _y(big, big_w, 5), _y(small, small_w, 6)
    # radial: 10 iterations (fixed); plain: to rtol 1e-3 (stops early)
    ref_b = _solve(big, yb, 10, 0.0, noalloc=False)
    ref_s = _solve(small, ys, 500, 1e-3, noalloc=False)
    assert 0 < ref_s[1] < 500
    _same(_solve(big, yb, 10, 0.0, noalloc=True), ref_b)
    _same(_solve(small, ys, 500, 1e-3, noalloc=True), ref_s)
    _same(_solve(big, yb, 10, 0.0, noalloc=True), ref_b)
    _same(_solve(small, ys, 500, 1e-3, noalloc=True), ref_s)


def test_cg_noalloc_refuses_oversized_solve(hip):
    """A 3-D 64^3 / 128^3 plan needs 4 x 32 MiB of FP64 grid vectors: the
    call fails with a clear error and launches nothing."""
    from pynufft_amd import _lib
    from pynufft_amd.nufft import NUFFT
    A = NUFFT()
    om = W.uniform_om(1000, 3, 7)
    A.plan(om, (64, 64, 64), (128, 128, 128), (6, 6, 6))
    A._ensure_fp64()
    y = np.ones(1000, np.complex64)
    with pytest.raises(ValueError, match='static scratch'):
        _solve(A, y, 3, 0.0, noalloc=True)
