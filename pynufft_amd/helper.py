"""Host-side FP64 plan construction and device discovery.

The min-max interpolator is designed exactly as pynufft does it (reference
``src/_helper/helper.py:620-802``): a Kaiser-Bessel-fitted cosine series gives
the scaling factor per axis, and each sample gets, per axis, J FP64 min-max
coefficients ``u_d`` and a first support cell.  What differs is the layout
the device needs:

* samples are re-ordered into gridding-tile order (a permutation ``perm``),
  so that neighbouring rows touch neighbouring cells;
* coefficients are kept factorized per axis (``u_d``, M x J_d) in complex64
  and, on demand, complex128; the full prod(Jd) table (the reference's CSR
  data) is only materialized for the ``'CSR'`` format;
* the gridding work list (tile, rows) is built here so that ``y2k`` needs no
  atomics.

Error behaviour follows the reference: non-tuple Nd/Kd/Jd raise TypeError,
length mismatches raise KeyError (helper.py:644-654).
"""
import concurrent.futures
import math
import os

import numpy as np
import scipy.linalg
import scipy.special

# Kaiser-Bessel shape per J for K/N == 2 (reference helper.py:973-988).
_KB_TABLE = {2: 2.5, 3: 2.27, 4: 2.31, 5: 2.34, 6: 2.32, 7: 2.32, 8: 2.35,
             9: 2.34, 10: 2.34, 11: 2.35, 12: 2.34, 13: 2.35, 14: 2.35,
             15: 2.35, 16: 2.33}


def device_list():
    """HIP devices as pynufft-style 7-tuples.

    Reference ``helper.device_list`` (helper.py:1172-1224) returns
    ``(API, api_n, dev_num, platform, device, thread, wavefront)``; here
    ``('hip', 0, index, None, torch.device('cuda', index), None, 64)``.
    """
    import torch
    from . import _lib
    n = _lib.device_count()
    return [('hip', 0, i, None, torch.device('cuda', i), None, 64)
            for i in range(n)]


# ---------------------------------------------------------------- design --
def _kb_alpha(J, ratio):
    """KB shape for support J at oversampling ratio K/N (helper.py:961-997)."""
    if ratio != 2:
        return 2.34 * J
    if J not in _KB_TABLE:
        # the reference fails here with a TypeError (dict_keys - ndarray)
        raise TypeError('Jd=%d has no Kaiser-Bessel entry for K/N = 2 '
                        '(supported: 2..16)' % J)
    return J * _KB_TABLE[J]


def _kb_transform(u, J, a):
    """KB Fourier transform with m = 0, d = 1 (helper.py:1000-1012)."""
    u = u * (1.0 + 0.0j)
    z = np.sqrt((2 * np.pi * (J / 2) * u) ** 2.0 - a ** 2.0)
    nu = 1.0 / 2 + 0
    y = ((2 * np.pi) ** (1.0 / 2)) * ((J / 2) ** 1.0) * (a ** 0) / \
        scipy.special.iv(0, a) * scipy.special.jv(nu, z) / (z ** nu)
    return np.real(y)


def fit_scaling(N, J, K):
    """Cosine-series coefficients (alpha, beta) of 1/KB_ft (helper.py:910-958).

    The least-squares fit runs in complex64 like the reference, so alpha
    carries float32 precision; it is returned as a float32 (L+1, 1) array.
    """
    beta = 1
    centre = (N - 1.0) / 2.0
    L = 13 if N > 40 else int(math.ceil(N / 3))
    n = np.arange(0, N) * 1.0 - centre
    a = _kb_alpha(J, K / N)
    if J > 1:
        rhs = (1 / _kb_transform(n / K, J, a)).reshape((N, 1))
    else:
        rhs = np.ones((N, 1))
    basis = np.cos(np.dot((beta * (2 * np.pi / K) * n).reshape((N, 1)),
                          np.arange(0, L + 1).reshape((1, L + 1))))
    sol = np.linalg.lstsq(np.nan_to_num(basis.astype(np.complex64)),
                          np.nan_to_num(rhs.conj().astype(np.complex64)),
                          rcond=-1)[0]
    if J > 1:
        sol[1:] = sol[1:] / 2.0
    else:
        sol[0] = 1.0
        sol[1:] = 0.0
    return np.real(sol), beta


def scaling_1d(N, K, alpha, beta):
    """Per-axis scaling factor s[n] = sum_l alpha_|l| e^{i 2pi/K (n-c) beta l}.

    complex128 (N,), helper.py:1015-1050.
    """
    alpha = np.real(alpha)
    L = len(alpha) - 1
    if L == 0:
        return np.dot(alpha, np.ones((N, 1)))[:, 0]
    acc = np.zeros((N, 1))
    expo = 1j * (2 * np.pi / K) * (np.arange(0, N).reshape((N, 1)) -
                                   (N - 1) / 2.0) * beta
    for l in range(-L, L + 1):
        acc = acc + alpha[abs(l)] * np.exp(expo * l)
    return acc[:, 0]


def overlap_pinv(N, J, K, alpha, beta):
    """Pseudo-inverse of the J x J overlap matrix (helper.py:1053-1083)."""
    L = np.size(alpha) - 1
    jj = np.arange(1, J + 1)
    lag = jj[None, :] - jj[:, None]
    ratio = 1.0 * K / N
    acc = np.zeros((J, J))
    for l1 in range(-L, L + 1):
        for l2 in range(-L, L + 1):
            acc = acc + (alpha[abs(l1)] * alpha[abs(l2)]) * \
                np.sinc(1.0 * (lag + beta * (l1 - l2)) / ratio)
    return scipy.linalg.pinv(acc)


def _dsinc(x):
    """d/dx sin(pi x)/(pi x) = (cos(pi x) - sinc(x)) / x, 0 at x = 0."""
    out = np.zeros_like(x)
    nz = x != 0
    xn = x[nz]
    out[nz] = (np.cos(np.pi * xn) - np.sinc(xn)) / xn
    return out


def _minmax_chunk(o, N, J, K, alpha, beta, Tinv, grad=False):
    gam = 2.0 * np.pi / (K * 1.0)
    L = np.size(alpha) - 1
    ratio = 1.0 * K / N
    koff = np.floor(1.0 * o / gam - 1.0 * J / 2.0)
    arg = np.add.outer(-np.arange(1, J + 1) * 1.0, 1.0 * o / gam - koff)  # (J, m)
    r = np.zeros(arg.shape, dtype=np.float32)
    for l in range(-L, L + 1):
        r = r + (alpha[abs(l)] * 1.0) * np.sinc((arg + 1.0 * l * beta) / ratio)
    c = Tinv.dot(r)
    ph = np.exp(-1.0j * o * N / 2.0) * np.exp(1.0j * gam * (N * 1.0 - 1.0) / 2.0 * arg)
    v = np.exp(1.0j * gam * (N * 1.0 - 1.0) / 2.0 * arg) * c
    v = np.exp(-1.0j * o * N / 2.0) * v
    base = np.mod(koff.astype(np.int64) + 1, K)
    if not grad:
        return v.T.conj(), base
    # d/d(om): arg moves by 1/gam (koff is piecewise constant), the two
    # phases contribute i(N-1)/2 - iN/2 = -i/2, and dr = sum_l alpha sinc'.
    dr = np.zeros(arg.shape)
    for l in range(-L, L + 1):
        dr = dr + (alpha[abs(l)] * 1.0) * _dsinc((arg + 1.0 * l * beta) / ratio)
    dr = dr / (gam * ratio)
    dv = ph * (Tinv.dot(dr) - 0.5j * c)
    return v.T.conj(), base, dv.T.conj()


def minmax_axis(om, N, J, K, alpha, beta, Tinv, chunk=1 << 16, pool=None,
                grad=False):
    """Per-sample min-max factors along one axis.

    Returns u (M, J) complex128 -- the reference's ``ud[d]`` (helper.py:606-618
    with OMEGA_u :148-162) -- and base (M,) int64, the first support cell
    mod(floor(om/gam - J/2) + 1, K) (OMEGA_k :164-210); with ``grad`` also
    du/d(om) (M, J) complex128.  Sample chunks are independent and run on
    ``pool`` (numpy releases the GIL); the arithmetic per sample is the same
    as unchunked.
    """
    M = om.shape[0]
    u = np.empty((M, J), dtype=np.complex128)
    base = np.empty(M, dtype=np.int64)
    du = np.empty((M, J), dtype=np.complex128) if grad else None

    def job(s):
        res = _minmax_chunk(om[s:s + chunk], N, J, K, alpha, beta, Tinv, grad)
        u[s:s + chunk] = res[0]
        base[s:s + chunk] = res[1]
        if grad:
            du[s:s + chunk] = res[2]

    starts = range(0, M, chunk)
    if pool is None:
        for s in starts:
            job(s)
    else:
        list(pool.map(job, starts))
    return (u, base, du) if grad else (u, base)


def factor_gradients(st):
    """d u_d / d om_d (M, J_d) complex128 for every axis of a plan."""
    om = st['om']
    out = []
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for a in range(len(st['Nd'])):
            N, J, K = st['Nd'][a], st['Jd'][a], st['Kd'][a]
            alpha, beta = st['alpha'][a], st['beta'][a]
            Tinv = overlap_pinv(N, J, K, alpha, beta)
            res = minmax_axis(om[:, a], N, J, K, alpha, beta, Tinv, pool=pool,
                              grad=True)
            out.append(res[2])
    return out


def tile_shape(Kd):
    """Gridding tile per axis: about 1024 cells, fastest axis longest."""
    d = len(Kd)
    pref = {1: (1024,), 2: (32, 32), 3: (8, 8, 16)}[d]
    return tuple(int(min(p, k)) for p, k in zip(pref, Kd))


def sort_and_bin(base, Kd, Jd, tile, chunk_entries=None):
    """Trajectory sort and the deterministic gridding work list.

    Rows are samples sorted by (tile of first support cell, sample index).
    Every tile met by a row's support gets that row in its bin; bins are
    ordered by original sample index, so each grid cell sums its samples in
    ascending sample order (the order of the reference's transposed-CSR
    matvec, helper.py:285-311 + _nufft_class_methods_cpu.py:110-112).
    Bins longer than ``chunk_entries`` are split into chunks that write
    partial tiles, summed in chunk order afterwards.

    Returns perm, items (n,4) int32, rows int32, merge (n,3) int32, n_slots.
    """
    d = len(Kd)
    M = base.shape[1]
    tgrid = [(Kd[a] + tile[a] - 1) // tile[a] for a in range(d)]
    n_tiles = int(np.prod(tgrid))
    # sort key: tile of the first support cell (C order over the tile grid)
    tkey = np.zeros(M, dtype=np.int64)
    for a in range(d):
        tkey = tkey * tgrid[a] + base[a] // tile[a]
    perm = np.argsort(tkey, kind='stable').astype(np.int64)
    inv = np.empty(M, dtype=np.int64)
    inv[perm] = np.arange(M)
    if M == 0:
        items = np.stack([np.arange(n_tiles), np.zeros(n_tiles), np.zeros(n_tiles),
                          -np.ones(n_tiles)], axis=1).astype(np.int32)
        return (perm, items, np.zeros(0, np.int32), np.zeros((0, 3), np.int32), 0)
    # distinct tiles touched per axis, per sample: cells (b + j) mod K
    per_axis = []
    for a in range(d):
        cells = np.mod(base[a][:, None] + np.arange(Jd[a])[None, :], Kd[a])
        t = np.sort(cells // tile[a], axis=1)              # (M, J)
        new = np.ones_like(t, dtype=bool)
        new[:, 1:] = t[:, 1:] != t[:, :-1]
        width = int(new.sum(axis=1).max()) if M else 1
        out = np.full((M, width), -1, dtype=np.int64)
        pos = np.cumsum(new, axis=1) - 1
        rr, cc = np.nonzero(new)
        out[rr, pos[rr, cc]] = t[rr, cc]
        per_axis.append(out)
    # Cartesian product of per-axis tile sets -> (sample, tile) pairs
    tid = np.zeros((M, 1), dtype=np.int64)
    valid = np.ones((M, 1), dtype=bool)
    for a in range(d):
        ta = per_axis[a]
        tid = (tid[:, :, None] * tgrid[a] + ta[:, None, :]).reshape(M, -1)
        valid = (valid[:, :, None] & (ta[:, None, :] >= 0)).reshape(M, -1)
    samp = np.broadcast_to(np.arange(M, dtype=np.int64)[:, None], tid.shape)[valid]
    key = tid[valid] * max(M, 1) + samp          # (tile, sample) lexicographic
    key.sort()
    ptile = key // max(M, 1)
    rows = inv[key - ptile * max(M, 1)].astype(np.int32)
    counts = np.bincount(ptile, minlength=n_tiles).astype(np.int64)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    total = int(counts.sum())
    if chunk_entries is None:
        chunk_entries = int(max(64, min(1024, math.ceil(total / 8192))))
    n_chunks = np.maximum(1, (counts + chunk_entries - 1) // chunk_entries)
    split = n_chunks > 1
    slot0 = np.zeros(n_tiles, dtype=np.int64)
    slot0[split] = np.concatenate([[0], np.cumsum(n_chunks[split])[:-1]]) \
        if split.any() else []
    n_slots = int(n_chunks[split].sum())
    tiles_rep = np.repeat(np.arange(n_tiles), n_chunks)
    k = np.arange(tiles_rep.size) - np.repeat(np.cumsum(n_chunks) - n_chunks, n_chunks)
    st = starts[tiles_rep] + k * chunk_entries
    cnt = np.minimum(chunk_entries, starts[tiles_rep] + counts[tiles_rep] - st)
    cnt = np.maximum(cnt, 0)
    slot = np.where(split[tiles_rep], slot0[tiles_rep] + k, -1)
    items = np.stack([tiles_rep, st, cnt, slot], axis=1)
    # heaviest items first (longest-processing-time order)
    items = items[np.argsort(-items[:, 2], kind='stable')].astype(np.int32)
    mt = np.nonzero(split)[0]
    merge = np.stack([mt, slot0[mt], n_chunks[mt]], axis=1).astype(np.int32) \
        if mt.size else np.zeros((0, 3), dtype=np.int32)
    return perm, items, rows, merge, n_slots


def plan(om, Nd, Kd, Jd, ft_axes=None, tile=None, chunk_entries=None):
    """FP64 host plan: per-axis design, per-sample factors, sort and bins.

    Mirrors ``helper.plan(om, Nd, Kd, Jd, ft_axes)`` (reference
    helper.py:620-802).  Returns a dict; arrays are numpy.
    """
    if type(Nd) != tuple:
        raise TypeError('Nd must be tuple, e.g. (256, 256)')
    if type(Kd) != tuple:
        raise TypeError('Kd must be tuple, e.g. (512, 512)')
    if type(Jd) != tuple:
        raise TypeError('Jd must be tuple, e.g. (6, 6)')
    if (len(Nd) != len(Kd)) | (len(Nd) != len(Jd)) | (len(Kd) != len(Jd)):
        raise KeyError('Nd, Kd, Jd must be in the same length, e.g. '
                       'Nd=(256,256),Kd=(512,512),Jd=(6,6)')
    dd = len(Nd)
    if ft_axes is not None and tuple(sorted(ft_axes)) != tuple(range(dd)):
        raise NotImplementedError('only ft_axes over all axes are supported')
    if dd < 1 or dd > 3:
        raise NotImplementedError('1 to 3 dimensions are supported (got %d)' % dd)
    om = np.asarray(om, dtype=np.float64)
    if om.ndim == 1 and dd == 1:
        om = om[:, None]
    if om.ndim != 2 or om.shape[1] != dd:
        raise ValueError('om must have shape (M, %d)' % dd)
    for a in range(dd):
        if Kd[a] < Nd[a]:
            raise ValueError('Kd must be >= Nd on every axis')
        if Jd[a] > Kd[a]:
            raise ValueError('Jd must be <= Kd on every axis')
    M = om.shape[0]
    st = {'Nd': Nd, 'Kd': Kd, 'Jd': Jd, 'M': np.int32(M), 'om': om,
          'tol': 0, 'alpha': [], 'beta': []}
    snd, ud, based = [], [], []
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for a in range(dd):
            alpha, beta = fit_scaling(Nd[a], Jd[a], Kd[a])
            st['alpha'].append(alpha)
            st['beta'].append(beta)
            snd.append(scaling_1d(Nd[a], Kd[a], alpha, beta))
            Tinv = overlap_pinv(Nd[a], Jd[a], Kd[a], alpha, beta)
            u, b = minmax_axis(om[:, a], Nd[a], Jd[a], Kd[a], alpha, beta, Tinv,
                               pool=pool if M > (1 << 17) else None)
            ud.append(u)
            based.append(b)
    # full scaling factor = real part of the outer product (helper.py:558-584)
    sn = np.reshape(1.0, (1,) * dd)
    for a in range(dd):
        shp = [1] * dd
        shp[a] = Nd[a]
        sn = sn * np.reshape(snd[a], shp)
    st['sn'] = sn.real
    st['snd'] = snd
    base = np.stack(based, axis=0) if M else np.zeros((dd, 0), np.int64)
    tile = tuple(tile) if tile is not None else tile_shape(Kd)
    perm, items, rows, merge, n_slots = sort_and_bin(base, Kd, Jd, tile,
                                                     chunk_entries)
    st.update(ud=ud, base=base, perm=perm, tile=tile, items=items,
              bin_rows=rows, merge=merge, n_slots=n_slots)
    return st


def twiddles(n):
    """exp(-2 pi i t / n), t < n, complex128 (rounded per dtype by the caller)."""
    t = np.arange(n)
    return np.exp(-2j * np.pi * t / n)
