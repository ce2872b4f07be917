"""Accuracy of the HIP path beyond the stage goldens.

* Error against the FP64 direct NUDFT (reference linalg/nudft_cpu.py:43-103)
  of forward and adjoint, next to the reference's own error, on the plain
  2-D config in full and on 8192 samples of the radial 2-D config.
* The full-grid 3-D adjoint (128^3 / 256^3) of 4096 nonzero samples against
  the reference's subset-plan adjoint (tests/golden/uniform3d_adjoint.npz).
* Plans without oversampling (Kd == Nd), which run their complex64
  operators in FP64, against the reference (tests/golden/nooversample.npz),
  including the tile / work-item / workgroup / coil layout that a random
  geometry sweep found at 2.25e-5 with FP32 operators.
* The CG's final residual output and the device tuple's stream.

Each test prints the errors it measures (run with -s to see them).
"""
import numpy as np
import pytest

import oracle
from pynufft_amd import workloads as W
from conftest import golden, rel

TOL = 1e-5
pytestmark = pytest.mark.gpu


def _nufft(om, Nd, Kd, Jd, **kw):
    from pynufft_amd.nufft import NUFFT
    A = NUFFT()
    A.plan(om, tuple(Nd), tuple(Kd), tuple(Jd), **kw)
    return A


def _nudft_adjoint_scale(Nd, Kd):
    # NUFFT adjoint = A^H y / prod(Kd), NUDFT adjoint = F^H y / prod(Nd)
    return float(np.prod(Nd)) / float(np.prod(Kd))


def test_plain2d_error_vs_nudft(hip):
    """64^2 / 128^2, Jd 6^2, M 4096 in full: forward and adjoint against
    the FP64 NUDFT, beside the reference's own errors (its golden outputs)."""
    g = golden('plain2d')
    w = W.plain2d()
    A = _nufft(w['om'], w['Nd'], w['Kd'], w['Jd'])
    D = oracle.OracleNUDFT()
    D.plan(w['om'], w['Nd'])
    x, yin = w['x'], g['adjoint_in']
    yd = D.forward(x)
    xd = D.adjoint(yin) * _nudft_adjoint_scale(w['Nd'], w['Kd'])
    e_fwd, r_fwd = rel(A.forward(x), yd), rel(g['forward'], yd)
    e_adj, r_adj = rel(A.adjoint(yin), xd), rel(g['adjoint'], xd)
    print('\nplain2d vs NUDFT: forward HIP %.3e (reference %.3e), adjoint HIP %.3e '
          '(reference %.3e)' % (e_fwd, r_fwd, e_adj, r_adj))
    assert abs(e_fwd - r_fwd) < 1e-6 and abs(e_adj - r_adj) < 1e-6
    assert e_fwd < 2e-5 and e_adj < 2e-5


def test_radial2d_error_vs_nudft(hip):
    """256^2 / 512^2 radial (M 205 824) on the 8192 samples of the radial
    golden: forward of the phantom and the adjoint of those samples (zero
    elsewhere) against the FP64 NUDFT, beside the reference (forward: its
    golden samples) and the oracle's subset plan (adjoint)."""
    import torch
    g = golden('radial2d')
    w = W.radial2d()
    sel = g['ysel']
    A = _nufft(w['om'], w['Nd'], w['Kd'], w['Jd'])
    D = oracle.OracleNUDFT()
    D.plan(w['om'][sel], w['Nd'])
    yd = D.forward(w['x'])
    e_fwd = rel(A.forward(w['x'])[sel], yd)
    r_fwd = rel(g['forward_sel'], yd)
    ysub = g['y_noisy'][sel]
    yfull = np.zeros(w['om'].shape[0], np.complex64)
    yfull[sel] = ysub
    xd = D.adjoint(ysub) * _nudft_adjoint_scale(w['Nd'], w['Kd'])
    e_adj = rel(A.adjoint(torch.from_numpy(yfull).to(A.device)).cpu().numpy(), xd)
    B = oracle.OracleNUFFT()
    B.plan(w['om'][sel], w['Nd'], w['Kd'], w['Jd'])
    r_adj = rel(B.adjoint(ysub), xd)
    print('\nradial2d (8192 samples) vs NUDFT: forward HIP %.3e (reference %.3e), '
          'adjoint HIP %.3e (oracle %.3e)' % (e_fwd, r_fwd, e_adj, r_adj))
    assert abs(e_fwd - r_fwd) < 1e-6 and abs(e_adj - r_adj) < 1e-6
    assert e_fwd < 2e-5 and e_adj < 2e-5


@pytest.mark.slow
def test_uniform3d_full_grid_adjoint(hip):
    """Full 2 097 152-sample plan, y zero except 4096 samples: the adjoint on
    the full 256^3 grid matches the reference's subset-plan adjoint on 16384
    voxels, in norm and in four projections."""
    import torch
    g = golden('uniform3d_adjoint')
    w = W.uniform3d()
    A = _nufft(w['om'], w['Nd'], w['Kd'], w['Jd'])
    y = torch.zeros(w['om'].shape[0], dtype=torch.complex64)
    y[torch.from_numpy(g['sel'])] = torch.from_numpy(g['y'])
    adj = A.adjoint(y.to(A.device))
    a = adj.cpu().numpy()
    e_sel = rel(a.ravel()[g['xsel']], g['adjoint_sel'])
    e_norm = abs(np.linalg.norm(a) / float(g['adjoint_norm']) - 1)
    e_proj = 0.0
    for i, seed in enumerate(g['proj_seeds']):
        v = torch.from_numpy(W.gaussian_image(w['Nd'], int(seed))).to(A.device)
        p = torch.vdot(v.reshape(-1).to(torch.complex128),
                       adj.reshape(-1).to(torch.complex128)).item()
        scale = float(torch.linalg.vector_norm(v)) * float(g['adjoint_norm'])
        e_proj = max(e_proj, abs(p - g['adjoint_proj'][i]) / scale)
    print('\nuniform3d full-grid adjoint: voxels %.3e, norm %.3e, projections %.3e'
          % (e_sel, e_norm, e_proj))
    assert e_sel < TOL and e_norm < TOL and e_proj < TOL


NOS = {
    # name: (expects FP64 operators, extra plan arguments)
    'nos3d': (True, {}),
    'nos3d_sweep': (True, dict(tile=(18, 7, 7), chunk=17, threads=64)),
    'nos2d': (True, {}),
    'k125': (False, {}),
}


@pytest.mark.parametrize('case', sorted(NOS))
def test_nooversample_plans(hip, case):
    """Kd == Nd (and 1.25 Nd): forward, adjoint, selfadjoint, y2k and a
    10-iteration CG within 1e-5 of the reference, one coil at a time and as
    the two coils of a batch=2 plan.  'nos3d_sweep' is the layout (tile
    18x7x7, 17-row work items, 64-thread workgroups, 2 coils) on which the
    FP32 operators were 2.25e-5 from the oracle's selfadjoint."""
    import torch
    name = case.split('_')[0]
    fp64, kw = NOS[case]
    g = golden('nooversample')
    Nd, Kd, Jd = (tuple(int(v) for v in g[name + '__' + k]) for k in ('Nd', 'Kd', 'Jd'))
    om = g[name + '__om']
    errs = {}
    A = _nufft(om, Nd, Kd, Jd, **kw)
    assert A.fp64_ops == fp64
    for c in range(2):
        p = '%s__%d_' % (name, c)
        x, yin = g[p + 'x'], g[p + 'forward'].astype(np.complex64)
        errs['forward%d' % c] = rel(A.forward(x), g[p + 'forward'])
        errs['adjoint%d' % c] = rel(A.adjoint(yin), g[p + 'adjoint'])
        errs['selfadjoint%d' % c] = rel(A.selfadjoint(x), g[p + 'selfadjoint'])
        errs['y2k%d' % c] = rel(A.y2k(yin), g[p + 'y2k'])
        errs['cg%d' % c] = rel(A.solve(yin, 'cg', maxiter=10), g[p + 'solve_cg10'])
    B = _nufft(om, Nd, Kd, Jd, batch=2, **kw)
    assert B.fp64_ops == fp64
    xs = np.stack([g['%s__%d_x' % (name, c)] for c in range(2)], axis=-1)
    ys = np.stack([g['%s__%d_forward' % (name, c)] for c in range(2)], axis=-1)
    yb = B.forward(xs)
    xa = B.adjoint(ys.astype(np.complex64))
    sa = B.selfadjoint(xs)
    for c in range(2):
        p = '%s__%d_' % (name, c)
        errs['batch_forward%d' % c] = rel(yb[:, c], g[p + 'forward'])
        errs['batch_adjoint%d' % c] = rel(xa[..., c], g[p + 'adjoint'])
        errs['batch_selfadjoint%d' % c] = rel(sa[..., c], g[p + 'selfadjoint'])
    worst = max(errs, key=errs.get)
    print('\n%s (fp64_ops=%s): worst %s %.3e' % (case, fp64, worst, errs[worst]))
    assert errs[worst] < TOL, errs


def test_cg_residual_output(hip):
    """pnh_solve_cg's final ||r|| per coil: ||b|| at maxiter = 0, below the
    tolerance when the stopping test holds, and after a fixed number of
    iterations the FP64 norm of b - P^H P k recomputed from scipy's CG
    iterate (same iterate as the reference's solve)."""
    import scipy.sparse.linalg
    g = golden('tiny2d')
    Nd, Kd, Jd = (tuple(int(v) for v in g[k]) for k in ('Nd', 'Kd', 'Jd'))
    A = _nufft(g['om'], Nd, Kd, Jd)
    yin = g['adjoint_in']
    B = oracle.OracleNUFFT()
    B.plan(g['om'], Nd, Kd, Jd)
    b = B.y2k(yin.astype(np.complex128)).ravel()
    bn = np.linalg.norm(b)
    A.solve(yin, 'cg', maxiter=0)
    assert abs(A.cg_resid / bn - 1) < 1e-12
    op = scipy.sparse.linalg.LinearOperator(
        (b.size, b.size), matvec=lambda v: B.spH.dot(B.sp.dot(v)), dtype=np.complex128)
    for it in (3, 5):
        k, _ = scipy.sparse.linalg.cg(op, b, maxiter=it, rtol=0.0)
        ref = np.linalg.norm(b - op.matvec(k))
        A.solve(yin, 'cg', maxiter=it, rtol=0.0)
        print('\ncg resid after %d: HIP %.6e, scipy iterate %.6e' % (it, A.cg_resid, ref))
        assert abs(A.cg_resid - ref) < 1e-9 * bn
    A.solve(yin, 'cg', maxiter=200, rtol=1e-2)
    assert A.cg_iters < 200 and A.cg_resid < 1e-2 * bn
    # per coil with batch: coil 1 is coil 0 scaled by 4 (a power of two, so
    # the complex64 samples and the shared fixed-point grid scale exactly)
    C = _nufft(g['om'], Nd, Kd, Jd, batch=2)
    C.solve(np.stack([yin, 4 * yin], axis=-1), 'cg', maxiter=4, rtol=0.0)
    r = C.cg_resid
    assert r.shape == (2,) and abs(r[1] / (4 * r[0]) - 1) < 1e-9


def test_device_list_stream(hip):
    """helper.device_list() carries the device's current stream in the
    sixth field (the reference's queue), and a NUFFT built from the tuple
    launches on it."""
    import torch
    from pynufft_amd import helper
    from pynufft_amd.nufft import NUFFT
    dl = helper.device_list()
    assert len(dl) >= 1
    t = dl[0]
    assert t[0] == 'hip' and t[2] == 0 and t[6] == 64
    assert isinstance(t[5], torch.cuda.Stream)
    s = torch.cuda.Stream()
    A = NUFFT(t[:5] + (s,) + t[6:])
    g = golden('tiny2d')
    Nd, Kd, Jd = (tuple(int(v) for v in g[k]) for k in ('Nd', 'Kd', 'Jd'))
    A.plan(g['om'], Nd, Kd, Jd)
    assert A.thr is s
    x = torch.from_numpy(g['x']).to(A.device)
    torch.cuda.synchronize()
    y = A.forward(x)
    s.synchronize()
    assert rel(y.cpu().numpy(), g['forward']) < TOL
