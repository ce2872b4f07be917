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
import hashlib
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
    ``(API, api_n, dev_num, platform, device, thread, wavefront)``, whose
    ``thread`` is the queue the device path launches on (helper.py:1195,
    1216); here ``('hip', 0, index, None, torch.device('cuda', index),
    stream, 64)`` with ``stream`` the device's current torch stream, which a
    ``NUFFT`` built from the tuple launches on.
    """
    import torch
    from . import _lib
    n = _lib.device_count()
    return [('hip', 0, i, None, torch.device('cuda', i),
             torch.cuda.current_stream(torch.device('cuda', i)), 64)
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


GRID_LDS_MAX = 163840        # bytes: int64 (re, im) accumulator + 256 B scratch
MAX_ITEM_ROWS = 4096         # rows per gridding work item (fixed-point headroom)


def tile_valid(t, K, J):
    """A tile of extent t on an axis of K cells is usable when it spans the
    axis or when every tile, the last (partial) one included, is at least
    J - 1 cells wide, so that a tile's padding reaches only the next tile."""
    if t >= K:
        return True
    if t < max(1, J - 1):
        return False
    return K - (-(-K // t) - 1) * t >= J - 1


def tile_pad(tile, Kd, Jd):
    """Accumulator extent per axis: tile + J - 1, or K for a spanning tile."""
    return tuple(K if t >= K else t + J - 1 for t, K, J in zip(tile, Kd, Jd))


def _valid_near(pref, K, J):
    """Largest valid tile <= pref, else the smallest valid one above it."""
    for t in range(min(pref, K), 0, -1):
        if tile_valid(t, K, J):
            return t
    for t in range(pref + 1, K + 1):
        if tile_valid(t, K, J):
            return t
    return K


def tile_shape(Kd, Jd=None):
    """Gridding tile per axis: 1024 cells (1-D), 16 x 16 (2-D) or
    16 x 16 x 16 (3-D) when valid for (Kd, Jd); then, while the padded
    accumulator does not fit in LDS, the axis with the widest accumulator
    steps down to its next valid tile."""
    d = len(Kd)
    Jd = tuple(Jd) if Jd is not None else (6,) * d
    pref = {1: (1024,), 2: (16, 16), 3: (16, 16, 16)}[d]
    tile = [_valid_near(p, K, J) for p, K, J in zip(pref, Kd, Jd)]

    def fits(t):
        return int(np.prod(tile_pad(t, Kd, Jd))) * 16 + 256 <= GRID_LDS_MAX

    while not fits(tile):
        pads = tile_pad(tile, Kd, Jd)
        for a in np.argsort(pads, kind='stable')[::-1]:
            t = tile[a] - 1
            while t >= 1 and not tile_valid(t, Kd[a], Jd[a]):
                t -= 1
            if t >= 1:
                tile[a] = t
                break
        else:
            raise ValueError('Jd %s is too large for the LDS gridding accumulator' % (Jd,))
    return tuple(int(t) for t in tile)


def check_tile(tile, Kd, Jd):
    """ValueError unless ``tile`` is valid on every axis and fits in LDS."""
    tile = tuple(int(t) for t in tile)
    if len(tile) != len(Kd):
        raise ValueError('tile must have one extent per axis')
    for t, K, J in zip(tile, Kd, Jd):
        if t < 1 or not tile_valid(t, K, J):
            raise ValueError('gridding tile %d on an axis of %d cells leaves a tile '
                             'narrower than Jd - 1 = %d' % (t, K, J - 1))
    if int(np.prod(tile_pad(tile, Kd, Jd))) * 16 + 256 > GRID_LDS_MAX:
        raise ValueError('gridding tile %s needs more LDS than a workgroup has' % (tile,))
    return tuple(min(t, K) for t, K in zip(tile, Kd))


def sort_and_bin(base, Kd, Jd, tile, chunk_rows=None):
    """Row order and the gridding work list (tile ownership).

    A row belongs to the tile that holds its first support cell.  Rows are
    the samples sorted by that tile (C order over the tile grid), ties in
    sample order, so each tile owns one contiguous run of rows; runs longer
    than ``chunk_rows`` (at most 4096) are cut into chunks.  Every chunk is
    one work item with its own partial slot; a tile's slots are consecutive
    and in row order, which fixes the order in which the gather pass adds
    them.

    Returns perm (M,) int64; items (n, 4) int32 rows [tile id, first row,
    rows, slot], heaviest first; tile_slots (n_tiles, 2) int32 rows
    [first slot, slot count].
    """
    d = len(Kd)
    M = base.shape[1]
    tgrid = [(Kd[a] + tile[a] - 1) // tile[a] for a in range(d)]
    n_tiles = int(np.prod(tgrid))
    tkey = np.zeros(M, dtype=np.int64)
    for a in range(d):
        tkey = tkey * tgrid[a] + base[a] // tile[a]
    perm = np.argsort(tkey, kind='stable').astype(np.int64)
    counts = np.bincount(tkey, minlength=n_tiles).astype(np.int64)
    starts = np.cumsum(counts) - counts
    if chunk_rows is None:
        chunk_rows = int(min(MAX_ITEM_ROWS, max(256, math.ceil(M / 2048))))
    chunk_rows = int(chunk_rows)
    if not 1 <= chunk_rows <= MAX_ITEM_ROWS:
        raise ValueError('chunk rows must be in 1..%d' % MAX_ITEM_ROWS)
    n_chunks = (counts + chunk_rows - 1) // chunk_rows
    slot0 = np.cumsum(n_chunks) - n_chunks
    tiles_rep = np.repeat(np.arange(n_tiles), n_chunks)
    k = np.arange(tiles_rep.size) - slot0[tiles_rep]
    first = starts[tiles_rep] + k * chunk_rows
    cnt = np.minimum(chunk_rows, starts[tiles_rep] + counts[tiles_rep] - first)
    items = np.stack([tiles_rep, first, cnt, np.arange(tiles_rep.size)], axis=1)
    items = items[np.argsort(-cnt, kind='stable')].astype(np.int32).reshape(-1, 4)
    tile_slots = np.stack([slot0, n_chunks], axis=1).astype(np.int32)
    return perm, items, tile_slots


def coefficient_bound(ud):
    """Upper bound of |coefficient| = prod_d |u_d[j_d]| over all rows and
    offsets, with a margin for rounding the factors to complex64."""
    if not len(ud) or ud[0].shape[0] == 0:
        return 1.0
    prod = np.ones(ud[0].shape[0])
    for u in ud:
        prod = prod * np.abs(u).max(axis=1)
    return float(prod.max()) * (1.0 + 2.0 ** -20)


def row_sum_bound(ud):
    """Upper bound of sum_j |coefficient(row, j)| = prod_d sum_j |u_d[row, j]|
    over all rows, with the same margin as coefficient_bound."""
    if not len(ud) or ud[0].shape[0] == 0:
        return 1.0
    prod = np.ones(ud[0].shape[0])
    for u in ud:
        prod = prod * np.abs(u).sum(axis=1)
    return float(prod.max()) * (1.0 + 2.0 ** -20)


_CACHE_VERSION = 1


def _cache_file(cache_dir, om, Nd, Kd, Jd):
    h = hashlib.sha256()
    h.update(repr((_CACHE_VERSION, tuple(Nd), tuple(Kd), tuple(Jd), om.shape)).encode())
    h.update(np.ascontiguousarray(om).tobytes())
    return os.path.join(cache_dir, 'plan_%s.npz' % h.hexdigest()[:32])


def _design_load(path, dd):
    """The geometry part of a plan (alpha, beta, sn_d, factors, first cells)
    from a cache file, or None when it is missing or unreadable."""
    try:
        with np.load(path, allow_pickle=False) as f:
            return {'alpha': [f['alpha_%d' % a] for a in range(dd)],
                    'beta': [int(f['beta_%d' % a]) for a in range(dd)],
                    'snd': [f['snd_%d' % a] for a in range(dd)],
                    'ud': [f['ud_%d' % a] for a in range(dd)],
                    'base': f['base']}
    except (OSError, KeyError, ValueError):
        return None


def _design_store(path, design):
    arrays = {'base': design['base']}
    for a in range(len(design['ud'])):
        arrays['alpha_%d' % a] = np.asarray(design['alpha'][a])
        arrays['beta_%d' % a] = np.asarray(design['beta'][a])
        arrays['snd_%d' % a] = np.asarray(design['snd'][a])
        arrays['ud_%d' % a] = design['ud'][a]
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = '%s.tmp%d.npz' % (path[:-4], os.getpid())
        np.savez(tmp, **arrays)
        os.replace(tmp, path)
    except OSError:
        pass


def _design(om, Nd, Kd, Jd):
    """Per-axis min-max design in FP64: alpha/beta, sn_d, factors u_d and
    first support cells (reference helper.py:620-802 up to the CSR build)."""
    dd = len(Nd)
    M = om.shape[0]
    out = {'alpha': [], 'beta': [], 'snd': [], 'ud': []}
    based = []
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for a in range(dd):
            alpha, beta = fit_scaling(Nd[a], Jd[a], Kd[a])
            out['alpha'].append(alpha)
            out['beta'].append(beta)
            out['snd'].append(scaling_1d(Nd[a], Kd[a], alpha, beta))
            Tinv = overlap_pinv(Nd[a], Jd[a], Kd[a], alpha, beta)
            u, b = minmax_axis(om[:, a], Nd[a], Jd[a], Kd[a], alpha, beta, Tinv,
                               pool=pool if M > (1 << 17) else None)
            out['ud'].append(u)
            based.append(b)
    out['base'] = np.stack(based, axis=0) if M else np.zeros((dd, 0), np.int64)
    return out


def plan(om, Nd, Kd, Jd, ft_axes=None, tile=None, chunk_rows=None, cache_dir=None):
    """FP64 host plan: per-axis design, per-sample factors, sort and bins.

    Mirrors ``helper.plan(om, Nd, Kd, Jd, ft_axes)`` (reference
    helper.py:620-802).  Returns a dict; arrays are numpy.

    ``cache_dir`` (default: the ``PYNUFFT_AMD_PLAN_CACHE`` environment
    variable, unset = no cache) keeps the FP64 design of each geometry
    (trajectory bytes, Nd, Kd, Jd) in an ``.npz`` file there and reuses it
    on the next plan of the same geometry; the row sort and work list are
    always rebuilt.
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
    st = {'Nd': Nd, 'Kd': Kd, 'Jd': Jd, 'M': np.int32(M), 'om': om, 'tol': 0}
    if cache_dir is None:
        cache_dir = os.environ.get('PYNUFFT_AMD_PLAN_CACHE') or None
    design = None
    if cache_dir is not None:
        path = _cache_file(cache_dir, om, Nd, Kd, Jd)
        design = _design_load(path, dd)
    if design is None:
        design = _design(om, Nd, Kd, Jd)
        if cache_dir is not None:
            _design_store(path, design)
    st['alpha'], st['beta'] = design['alpha'], design['beta']
    snd, ud, base = design['snd'], design['ud'], design['base']
    # full scaling factor = real part of the outer product (helper.py:558-584)
    sn = np.reshape(1.0, (1,) * dd)
    for a in range(dd):
        shp = [1] * dd
        shp[a] = Nd[a]
        sn = sn * np.reshape(snd[a], shp)
    st['sn'] = sn.real
    st['snd'] = snd
    tile = check_tile(tile, Kd, Jd) if tile is not None else tile_shape(Kd, Jd)
    perm, items, tile_slots = sort_and_bin(base, Kd, Jd, tile, chunk_rows)
    st.update(ud=ud, base=base, perm=perm, tile=tile, items=items,
              tile_slots=tile_slots, u_bound=coefficient_bound(ud),
              row_bound=row_sum_bound(ud),
              fp64_ops=needs_fp64_operators(st['sn']))
    return st


# Apodization spread max|sn| / min|sn| above which the complex64 operators
# run in FP64.  The interpolation undoes the apodization, so x * sn and its
# spectrum are up to the spread larger than the samples, and the products
# cancel: FP32 arithmetic then loses about spread * 1e-8 of relative
# accuracy (measured on uniform 1-3-D plans with K/N from 1 to 2), while the
# reference's complex128 CSR products and ifftn keep it.  Plans with K >= 2N
# and J <= 8 stay below 8; Kd == Nd reaches 500 in 2-D and 36 000 in 3-D.
FP64_SPREAD = 64.0


def needs_fp64_operators(sn):
    """True when the apodization spread calls for FP64 operators."""
    a = np.abs(np.asarray(sn, dtype=np.float64))
    if a.size == 0 or not np.all(np.isfinite(a)) or a.min() <= 0.0:
        return True
    return bool(a.max() / a.min() > FP64_SPREAD)


def laplacian_spectrum(Kd):
    """The Laplacian kernel of the L1-TV solve in k-space, computed as the
    reference computes it (src/_helper/helper.py:11-45,
    create_laplacian_kernel): the periodic stencil (-2 ndim at the origin, 1
    at the two neighbours along each axis) in complex64, transformed by
    numpy's fftn (complex64 out).  Its rounding (imaginary parts of order
    1e-7 where the exact spectrum is real) moves the split-Bregman result by
    up to 1e-5, so the solve uses these values rather than the closed form."""
    d = len(Kd)
    stencil = np.zeros(tuple(Kd), dtype=np.complex64)
    origin = (0,) * d
    stencil[origin] = -2.0 * d
    for p in range(d):
        for step in (1, -1):
            at = list(origin)
            at[p] = step
            stencil[tuple(at)] = 1
    return np.fft.fftn(stencil)


def twiddles(n):
    """exp(-2 pi i t / n), t < n, complex128 (rounded per dtype by the caller)."""
    t = np.arange(n)
    return np.exp(-2j * np.pi * t / n)
