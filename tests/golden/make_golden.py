"""Write golden vectors from the reference pynufft CPU path (run once, here).

Usage (in the container that has the reference checkout)::

    python tests/golden/make_golden.py [--ref /root/reference] [--only NAME]

The reference package is loaded read-only under the name ``pynufft`` straight
from its checkout (no bytecode is written there).  Each case writes one
``<name>.npz`` beside this script holding inputs, expected outputs and the
numpy/scipy versions used.  Nothing from the reference is stored except these
vectors.  Inputs come from ``pynufft_amd.workloads`` so a test elsewhere can
regenerate and fingerprint them.
"""
import argparse
import importlib.util
import os
import sys

sys.dont_write_bytecode = True
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402
import scipy  # noqa: E402

from pynufft_amd import workloads as W  # noqa: E402


def load_reference(path):
    spec = importlib.util.spec_from_file_location(
        'pynufft', os.path.join(path, '__init__.py'),
        submodule_search_locations=[path])
    mod = importlib.util.module_from_spec(spec)
    sys.modules['pynufft'] = mod
    spec.loader.exec_module(mod)
    return mod


def versions():
    return dict(numpy_version=np.__version__, scipy_version=scipy.__version__)


def plan_record(ref, om, Nd, Kd, Jd, full_csr=True, rows=None):
    """Reference helper.plan output: alpha/beta, sn_d, sn, and CSR rows."""
    st = ref.helper.plan(om, Nd, Kd, Jd, ft_axes=None, format='CSR')
    rec = {}
    for d in range(len(Nd)):
        rec['alpha_%d' % d] = np.asarray(st['alpha'][d])
        rec['beta_%d' % d] = np.asarray(st['beta'][d])
        a, b = st['alpha'][d], st['beta'][d]
        rec['snd_%d' % d] = ref.helper.nufft_scale(Nd[d], Kd[d], a, b)
    rec['sn'] = st['sn']
    p = st['p'].tocsr()
    if full_csr:
        rec['p_data'] = p.data
        rec['p_indices'] = p.indices
        rec['p_indptr'] = p.indptr
    if rows is not None:
        P = int(np.prod(Jd))
        sel = (rows[:, None] * P + np.arange(P)[None, :]).ravel()
        rec['p_rows'] = rows
        rec['p_rows_data'] = p.data[sel]
        rec['p_rows_indices'] = p.indices[sel]
        rec['p_fingerprint'] = np.array(W.fingerprint(p.data))
        rec['p_index_sum'] = np.array(int(p.indices.astype(np.int64).sum()))
    return rec


def operator_record(ref, om, Nd, Kd, Jd, x, cg_iters=30, y_cg=None):
    A = ref.NUFFT()
    A.plan(om, Nd, Kd, Jd)
    rec = {}
    rec['sn_c64'] = A.sn
    rec['xx2k'] = A.xx2k(A.x2xx(x))
    y = A.forward(x)
    rec['forward'] = y
    yin = y.astype(np.complex64) if y_cg is None else y_cg
    rec['adjoint_in'] = yin
    rec['adjoint'] = A.adjoint(yin)
    rec['selfadjoint'] = A.selfadjoint(x)
    rec['y2k'] = A.y2k(yin)
    rec['solve_cg'] = A.solve(yin, 'cg', maxiter=cg_iters)
    rec['cg_iters'] = np.array(cg_iters)
    return rec


def save(name, **arrays):
    arrays.update(versions())
    path = os.path.join(HERE, name + '.npz')
    np.savez_compressed(path, **arrays)
    print('wrote', path, '%.1f KB' % (os.path.getsize(path) / 1024))


TINY = {
    # name: (Nd, Kd, Jd, M, seed)
    'tiny1d': ((16,), (32,), (6,), 20, 11),
    'tiny1d_ratio16': ((15,), (24,), (5,), 17, 12),
    'tiny2d': ((16, 12), (32, 24), (6, 5), 30, 13),
    'tiny2d_j3': ((10, 10), (16, 16), (3, 3), 25, 14),
    'tiny2d_j8': ((20, 18), (40, 36), (8, 7), 40, 18),
    'tiny3d': ((8, 8, 8), (16, 16, 16), (6, 6, 6), 10, 15),
    'tiny3d_mixed': ((6, 8, 10), (12, 16, 20), (4, 6, 8), 12, 16),
}


def make_tiny(ref):
    for name, (Nd, Kd, Jd, M, seed) in TINY.items():
        om = W.uniform_om(M, len(Nd), seed)
        x = W.gaussian_image(Nd, seed + 1000)
        rec = plan_record(ref, om, Nd, Kd, Jd)
        rec.update(operator_record(ref, om, Nd, Kd, Jd, x, cg_iters=5))
        save(name, om=om, x=x, Nd=np.array(Nd), Kd=np.array(Kd),
             Jd=np.array(Jd), **rec)


def make_grid_j1(ref):
    """Jd = 1, Kd = Nd, om on the grid: forward equals the NUDFT exactly."""
    Nd = Kd = (8, 8)
    Jd = (1, 1)
    g = np.indices(Kd).reshape(2, -1).T
    rng = np.random.default_rng(17)
    sel = rng.choice(g.shape[0], 24, replace=False)
    om = 2 * np.pi * (g[sel] - np.array(Kd) / 2) / np.array(Kd)
    x = W.gaussian_image(Nd, 1017)
    rec = plan_record(ref, om, Nd, Kd, Jd)
    A = ref.NUFFT()
    A.plan(om, Nd, Kd, Jd)
    rec['forward'] = A.forward(x)
    rec['adjoint'] = A.adjoint(rec['forward'].astype(np.complex64))
    save('grid_j1', om=om, x=x, Nd=np.array(Nd), Kd=np.array(Kd),
         Jd=np.array(Jd), **rec)


def make_plain2d(ref):
    w = W.plain2d()
    om, x, Nd, Kd, Jd = w['om'], w['x'], w['Nd'], w['Kd'], w['Jd']
    rows = np.arange(0, om.shape[0], 64)
    rec = plan_record(ref, om, Nd, Kd, Jd, full_csr=False, rows=rows)
    rec.update(operator_record(ref, om, Nd, Kd, Jd, x, cg_iters=30))
    save('plain2d', om=om, x=x, Nd=np.array(Nd), Kd=np.array(Kd),
         Jd=np.array(Jd), **rec)


def _subset(a, idx):
    return a.ravel()[idx]


def make_radial2d(ref):
    w = W.radial2d()
    om, x, Nd, Kd, Jd = w['om'], w['x'], w['Nd'], w['Kd'], w['Jd']
    M = om.shape[0]
    A = ref.NUFFT()
    A.plan(om, Nd, Kd, Jd)
    y = A.forward(x)
    y_noisy = (y + W.complex_noise(y, 0.01, w['noise_seed'])).astype(np.complex64)
    rng = np.random.default_rng(2024)
    ysel = np.sort(rng.choice(M, 8192, replace=False))
    xsel = np.sort(rng.choice(int(np.prod(Nd)), 8192, replace=False))
    ksel = np.sort(rng.choice(int(np.prod(Kd)), 8192, replace=False))
    adj = A.adjoint(y_noisy)
    sa = A.selfadjoint(x)
    kk = A.xx2k(A.x2xx(x))
    cg = A.solve(y_noisy, 'cg', maxiter=30)
    cg10 = A.solve(y_noisy, 'cg', maxiter=10)
    cg20 = A.solve(y_noisy, 'cg', maxiter=20)
    # the same problem with the samples in another order: the spread of the
    # reference's own 30-iteration result under an FP64 summation reordering
    perm = np.random.default_rng(2025).permutation(M)
    B = ref.NUFFT()
    B.plan(om[perm], Nd, Kd, Jd)
    cg_perm = B.solve(y_noisy[perm], 'cg', maxiter=30)
    save('radial2d', x_fp=np.array(W.fingerprint(x)),
         om_fp=np.array(W.fingerprint(om)), ysel=ysel, xsel=xsel, ksel=ksel,
         forward_sel=y[ysel], forward_fp=np.array(W.fingerprint(y)),
         forward_norm=np.linalg.norm(y), y_noisy=y_noisy,
         adjoint=adj.astype(np.complex64), selfadjoint_sel=_subset(sa, xsel),
         selfadjoint_norm=np.linalg.norm(sa), xx2k_sel=_subset(kk, ksel),
         xx2k_norm=np.linalg.norm(kk), solve_cg=cg.astype(np.complex64),
         solve_cg10=cg10.astype(np.complex64),
         solve_cg20=cg20.astype(np.complex64),
         solve_cg_permuted=cg_perm.astype(np.complex64), cg_iters=np.array(30))


def make_uniform3d(ref):
    """3-D 128^3 / 256^3: per-row exactness lets a subset plan give y[sel]."""
    w = W.uniform3d()
    om, x, Nd, Kd, Jd = w['om'], w['x'], w['Nd'], w['Kd'], w['Jd']
    M = om.shape[0]
    rng = np.random.default_rng(3030)
    ysel = np.sort(rng.choice(M, 4096, replace=False))
    A = ref.NUFFT()
    A.plan(om[ysel], Nd, Kd, Jd)
    kk = A.xx2k(A.x2xx(x))
    ysub = A.k2y(kk)
    ksel = np.sort(rng.choice(int(np.prod(Kd)), 8192, replace=False))
    rec = plan_record(ref, om[ysel[:64]], Nd, Kd, Jd, full_csr=True)
    del rec['sn']
    save('uniform3d', x_fp=np.array(W.fingerprint(x)),
         om_fp=np.array(W.fingerprint(om)), ysel=ysel, forward_sel=ysub,
         ksel=ksel, xx2k_sel=_subset(kk, ksel), xx2k_norm=np.linalg.norm(kk),
         **{'rows64_' + k: v for k, v in rec.items()})


def make_uniform3d_adjoint(ref):
    """3-D 128^3 / 256^3 adjoint of 4096 nonzero samples on the full grid.

    The adjoint is linear and each CSR row depends only on its own sample,
    so a plan of the selected samples gives the full plan's adjoint of a y
    that is zero elsewhere.  The 128^3 image is kept as 16384 voxels, its
    norm and its inner products with four seeded Gaussian images."""
    w = W.uniform3d()
    om, Nd, Kd, Jd = w['om'], w['Nd'], w['Kd'], w['Jd']
    M = om.shape[0]
    rng = np.random.default_rng(3031)
    sel = np.sort(rng.choice(M, 4096, replace=False))
    yv = (rng.standard_normal(4096) + 1j * rng.standard_normal(4096)).astype(np.complex64)
    A = ref.NUFFT()
    A.plan(om[sel], Nd, Kd, Jd)
    adj = A.adjoint(yv)
    xsel = np.sort(rng.choice(int(np.prod(Nd)), 16384, replace=False))
    proj = np.array([np.vdot(W.gaussian_image(Nd, 3100 + i).astype(np.complex128).ravel(),
                             adj.astype(np.complex128).ravel()) for i in range(4)])
    save('uniform3d_adjoint', om_fp=np.array(W.fingerprint(om)), sel=sel, y=yv,
         xsel=xsel, adjoint_sel=adj.ravel()[xsel], adjoint_norm=np.linalg.norm(adj),
         adjoint_proj=proj, proj_seeds=np.arange(3100, 3104))


# Plans without oversampling on some axis: the apodization spans a large
# range and the interpolation cancels large terms (name: Nd, Kd, Jd, M, seed)
NOOVERSAMPLE = {
    'nos3d': ((18, 14, 7), (18, 14, 7), (6, 8, 7), 1313, 122),
    'nos2d': ((32, 32), (32, 32), (6, 6), 3000, 123),
    'k125': ((32, 32), (40, 40), (6, 6), 3000, 124),
}


def make_nooversample(ref):
    """forward/adjoint/selfadjoint/y2k/CG of the reference on plans with
    Kd == Nd (and one at Kd = 1.25 Nd), two images each (the two coils of a
    batch=2 plan are checked against them one by one)."""
    out = {}
    for name, (Nd, Kd, Jd, M, seed) in NOOVERSAMPLE.items():
        om = W.uniform_om(M, len(Nd), seed)
        A = ref.NUFFT()
        A.plan(om, Nd, Kd, Jd)
        out[name + '__om'] = om
        out[name + '__Nd'], out[name + '__Kd'], out[name + '__Jd'] = \
            np.array(Nd), np.array(Kd), np.array(Jd)
        for c in range(2):
            x = W.gaussian_image(Nd, seed * 10 + c)
            y = A.forward(x)
            yin = y.astype(np.complex64)
            p = '%s__%d_' % (name, c)
            out[p + 'x'] = x
            out[p + 'forward'] = y
            out[p + 'adjoint'] = A.adjoint(yin)
            out[p + 'selfadjoint'] = A.selfadjoint(x)
            out[p + 'y2k'] = A.y2k(yin)
            out[p + 'solve_cg10'] = A.solve(yin, 'cg', maxiter=10)
    save('nooversample', **out)


def make_small3d(ref):
    """3-D 32^3 / 64^3, M = 20000: full operator outputs (subset + norms)."""
    Nd, Kd, Jd = (32, 32, 32), (64, 64, 64), (6, 6, 6)
    om = W.uniform_om(20000, 3, 303)
    x = W.gaussian_image(Nd, 304)
    A = ref.NUFFT()
    A.plan(om, Nd, Kd, Jd)
    y = A.forward(x)
    yin = y.astype(np.complex64)
    save('small3d', om_fp=np.array(W.fingerprint(om)),
         x_fp=np.array(W.fingerprint(x)), Nd=np.array(Nd), Kd=np.array(Kd),
         Jd=np.array(Jd), forward=y, adjoint=A.adjoint(yin),
         selfadjoint=A.selfadjoint(x), solve_cg=A.solve(yin, 'cg', maxiter=30),
         cg_iters=np.array(30), **versions())


SOLVER_CASES = ('tiny1d', 'tiny2d', 'tiny3d', 'plain2d')
L1TV_RHO = (0.05, 1.0)


def make_solvers(ref):
    """The Toeplitz weights and AHA of ``_precompute_sp_cpu`` /
    ``_selfadjoint2_cpu`` and the solvers other than 'cg' (``solve_cpu.solve``:
    'dc', 'L1TVOLS', 'lsqr', 'lsmr', 'bicgstab', 'bicg', 'gmres', 'lgmres'),
    on the inputs of existing fixtures; keys are ``<case>__<output>``."""
    out = {}
    for case in SOLVER_CASES:
        g = np.load(os.path.join(HERE, case + '.npz'))
        Nd, Kd, Jd = (tuple(int(v) for v in g[k]) for k in ('Nd', 'Kd', 'Jd'))
        x = g['x']
        yin = g['adjoint_in']
        A = ref.NUFFT()
        A.plan(g['om'], Nd, Kd, Jd)
        A._precompute_sp_cpu()
        out[case + '__W'] = A.W
        out[case + '__selfadjoint2'] = A._selfadjoint2_cpu(x)
        out[case + '__dc3'] = A.solve(yin, 'dc', maxiter=3)
        for rho in L1TV_RHO:
            out[case + '__l1tvols5_rho%g' % rho] = A.solve(yin, 'L1TVOLS', maxiter=5, rho=rho)
        out[case + '__lsqr8'] = A.solve(yin, 'lsqr', iter_lim=8)
        out[case + '__lsmr8'] = A.solve(yin, 'lsmr', maxiter=8)
        for solver in ('bicgstab', 'bicg'):
            out[case + '__%s6' % solver] = A.solve(yin, solver, maxiter=6)
        out[case + '__gmres'] = A.solve(yin, 'gmres', restart=4, maxiter=2)
        out[case + '__lgmres'] = A.solve(yin, 'lgmres', maxiter=2)
    save('solvers', **out)


CASES = {'tiny': make_tiny, 'grid_j1': make_grid_j1, 'plain2d': make_plain2d,
         'radial2d': make_radial2d, 'small3d': make_small3d,
         'uniform3d': make_uniform3d, 'uniform3d_adjoint': make_uniform3d_adjoint,
         'nooversample': make_nooversample, 'solvers': make_solvers}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--ref', default='/root/reference')
    ap.add_argument('--only', default=None)
    args = ap.parse_args()
    ref = load_reference(args.ref)
    for name, fn in CASES.items():
        if args.only and name != args.only:
            continue
        print('case', name, flush=True)
        fn(ref)


if __name__ == '__main__':
    main()
