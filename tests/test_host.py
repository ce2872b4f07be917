"""Host logic of the product (no GPU): the FP64 plan builder, the row sort
and the gridding work list, checked by a numpy emulation of the y2k
kernels' algorithm (tile ownership, padded accumulators, fixed point,
gather order) against the oracle's transposed-CSR matvec."""
import numpy as np
import pytest

import oracle
from pynufft_amd import helper
from pynufft_amd import workloads as W
from conftest import TINY_CASES, golden, geometry, rel


def tile_frame(tid, tile, tgrid, Kd):
    """Tile origin per axis and the wrap index (K for a spanning tile)."""
    D = len(Kd)
    T0, wrap = [0] * D, [0] * D
    for d in range(D - 1, -1, -1):
        T0[d] = (tid % tgrid[d]) * tile[d]
        tid //= tgrid[d]
        wrap[d] = Kd[d] if tgrid[d] == 1 else 1 << 30
    return T0, wrap


def gather_sources(c, tile, tgrid, Kd, Jd):
    """(tile id, accumulator index) pairs the gather pass adds for grid cell
    c, in its order: neighbour masks 0 .. 2^D - 1 (bit d = the tile below
    along axis d)."""
    D = len(Kd)
    pad = helper.tile_pad(tile, Kd, Jd)
    out = []
    for mask in range(1 << D):
        ok, tix, pidx = True, 0, 0
        for d in range(D):
            t, x = c[d] // tile[d], c[d] % tile[d]
            if tgrid[d] == 1:
                ok &= not (mask >> d) & 1
                x = c[d]
            elif (mask >> d) & 1:
                ok &= x < Jd[d] - 1
                t = tgrid[d] - 1 if t == 0 else t - 1
                x += min(tile[d], Kd[d] - t * tile[d])
            tix = tix * tgrid[d] + t
            pidx = pidx * pad[d] + x
        if ok:
            out.append((tix, pidx))
    return out


def emulate_y2k(st, Kd, Jd, y):
    """The y2k kernels' algorithm in numpy: per work item, every support box
    of the tile's rows into the padded accumulator (asserting it lies
    inside), 64-bit fixed point with the item's power-of-two scale, the
    partial slot scaled back; then the gather pass in its fixed order."""
    D = len(Kd)
    tile = st['tile']
    tgrid = [(Kd[a] + tile[a] - 1) // tile[a] for a in range(D)]
    pad = helper.tile_pad(tile, Kd, Jd)
    perm = st['perm']
    base = st['base'][:, perm]
    u = [st['ud'][a][perm] for a in range(D)]
    yrow = np.asarray(y)[perm]
    slots = {}
    for tid, first, cnt, slot in st['items']:
        assert 1 <= cnt <= helper.MAX_ITEM_ROWS
        T0, wrap = tile_frame(int(tid), tile, tgrid, Kd)
        rows = range(first, first + cnt)
        ymax = max(abs(yrow[r].real) + abs(yrow[r].imag) for r in rows)
        B = ymax * st['u_bound']
        e = max(np.frexp(B)[1], -960) if B > 0 else 0
        scale = 2.0 ** (50 - e)
        acc_re = np.zeros(int(np.prod(pad)), dtype=np.int64)
        acc_im = np.zeros_like(acc_re)
        for r in rows:
            for d in range(D):           # the row belongs to this tile
                assert T0[d] <= base[d][r] < T0[d] + tile[d]
            for j in np.ndindex(*Jd):
                x = [base[d][r] - T0[d] + j[d] for d in range(D)]
                x = [x[d] - Kd[d] if x[d] >= wrap[d] else x[d] for d in range(D)]
                assert all(0 <= x[d] < pad[d] for d in range(D))
                w = 1.0
                for d in range(D):
                    w = w * u[d][r][j[d]]
                v = np.conj(w) * yrow[r]
                assert abs(v.real) * scale < 2 ** 50 and abs(v.imag) * scale < 2 ** 50
                i = int(np.ravel_multi_index(x, pad))
                acc_re[i] += int(np.rint(v.real * scale))
                acc_im[i] += int(np.rint(v.imag * scale))
        slots[int(slot)] = (acc_re + 1j * acc_im) / scale
    return gather(st, Kd, Jd, slots)


def gather(st, Kd, Jd, slots):
    """The gather pass: every grid cell sums, in the kernel's order, the
    slots of the tiles whose padded box holds it."""
    tile = st['tile']
    tgrid = [(Kd[a] + tile[a] - 1) // tile[a] for a in range(len(Kd))]
    out = np.zeros(Kd, dtype=complex)
    for c in np.ndindex(*Kd):
        for tix, pidx in gather_sources(c, tile, tgrid, Kd, Jd):
            s0, ns = st['tile_slots'][tix]
            for k in range(ns):
                out[c] += slots[int(s0 + k)][pidx]
    return out


def item_bounds(st, Kd, Jd):
    """The item-bound kernel's quantity: per work item, the largest over
    the cells of its padded tile of the sum over its rows of the product of
    the complex64 factor magnitudes."""
    D = len(Kd)
    tile = st['tile']
    tgrid = [(Kd[a] + tile[a] - 1) // tile[a] for a in range(D)]
    pad = helper.tile_pad(tile, Kd, Jd)
    perm = st['perm']
    base = st['base'][:, perm]
    mag = [np.abs(st['ud'][a][perm].astype(np.complex64)).astype(np.float64) for a in range(D)]
    out = []
    for tid, first, cnt, slot in st['items']:
        T0, wrap = tile_frame(int(tid), tile, tgrid, Kd)
        acc = np.zeros(int(np.prod(pad)))
        for r in range(first, first + cnt):
            for j in np.ndindex(*Jd):
                x = [base[d][r] - T0[d] + j[d] for d in range(D)]
                x = [x[d] - Kd[d] if x[d] >= wrap[d] else x[d] for d in range(D)]
                w = 1.0
                for d in range(D):
                    w *= mag[d][r][j[d]]
                acc[int(np.ravel_multi_index(x, pad))] += w
        out.append(acc.max())
    return np.array(out)


def emulate_y2k_packed(st, Kd, Jd, y):
    """The packed complex64 gridding in numpy: complex64 factors, scale
    2^(30 - e) with 2^e > max(|Re y| + |Im y|) * item bound, each rounded
    contribution added as the 64-bit word re + im * 2^32 (modulo 2^64), the
    real and imaginary sums unpacked from the total.  Asserts that the sums
    stay below 2^31 and that unpacking recovers them exactly."""
    D = len(Kd)
    tile = st['tile']
    tgrid = [(Kd[a] + tile[a] - 1) // tile[a] for a in range(D)]
    pad = helper.tile_pad(tile, Kd, Jd)
    perm = st['perm']
    base = st['base'][:, perm]
    u = [st['ud'][a][perm].astype(np.complex64).astype(np.complex128) for a in range(D)]
    yrow = np.asarray(y).astype(np.complex64).astype(np.complex128)[perm]
    bounds = item_bounds(st, Kd, Jd)
    slots = {}
    for n_item, (tid, first, cnt, slot) in enumerate(st['items']):
        T0, wrap = tile_frame(int(tid), tile, tgrid, Kd)
        rows = range(first, first + cnt)
        ymax = max(abs(yrow[r].real) + abs(yrow[r].imag) for r in rows)
        B = ymax * bounds[n_item]
        e = max(np.frexp(B)[1], -960) if B > 0 else 0
        scale = 2.0 ** (30 - e)
        n = int(np.prod(pad))
        packed = [0] * n
        acc_re = np.zeros(n, dtype=np.int64)
        acc_im = np.zeros(n, dtype=np.int64)
        for r in rows:
            for j in np.ndindex(*Jd):
                x = [base[d][r] - T0[d] + j[d] for d in range(D)]
                x = [x[d] - Kd[d] if x[d] >= wrap[d] else x[d] for d in range(D)]
                w = 1.0
                for d in range(D):
                    w = w * u[d][r][j[d]]
                v = np.conj(w) * yrow[r]
                re, im = int(np.rint(v.real * scale)), int(np.rint(v.imag * scale))
                i = int(np.ravel_multi_index(x, pad))
                packed[i] = (packed[i] + re + (im << 32)) % (1 << 64)
                acc_re[i] += re
                acc_im[i] += im
        assert np.abs(acc_re).max() < 2 ** 31 and np.abs(acc_im).max() < 2 ** 31
        lo = np.array([(v & 0xffffffff) for v in packed], dtype=np.int64)
        re = np.where(lo >= 2 ** 31, lo - 2 ** 32, lo)
        im = np.array([((v - int(r)) % (1 << 64)) >> 32 for v, r in zip(packed, re)],
                      dtype=np.int64)
        im = np.where(im >= 2 ** 31, im - 2 ** 32, im)
        assert np.array_equal(re, acc_re) and np.array_equal(im, acc_im)
        slots[int(slot)] = (re + 1j * im) / scale
    return gather(st, Kd, Jd, slots)


def emulate_k2y(st, Kd, Jd, k):
    """The tile k2y kernel's algorithm in numpy: per work item, the padded
    box of the tile copied from the grid ((T0 + x) wrapped once), every row
    gathering its support box from it."""
    D = len(Kd)
    tile = st['tile']
    tgrid = [(Kd[a] + tile[a] - 1) // tile[a] for a in range(D)]
    pad = helper.tile_pad(tile, Kd, Jd)
    perm = st['perm']
    base = st['base'][:, perm]
    u = [st['ud'][a][perm] for a in range(D)]
    y = np.full(perm.size, np.nan, dtype=complex)
    for tid, first, cnt, slot in st['items']:
        T0, wrap = tile_frame(int(tid), tile, tgrid, Kd)
        box = np.zeros(pad, dtype=complex)
        for x in np.ndindex(*pad):
            c = [T0[d] + x[d] for d in range(D)]
            c = tuple(c[d] - Kd[d] if c[d] >= Kd[d] else c[d] for d in range(D))
            assert all(0 <= c[d] < Kd[d] for d in range(D))
            box[x] = k[c]
        for r in range(first, first + cnt):
            acc = 0j
            for j in np.ndindex(*Jd):
                x = [base[d][r] - T0[d] + j[d] for d in range(D)]
                x = tuple(x[d] - Kd[d] if x[d] >= wrap[d] else x[d] for d in range(D))
                w = 1.0
                for d in range(D):
                    w = w * u[d][r][j[d]]
                acc += w * box[x]
            y[perm[r]] = acc
    return y


@pytest.mark.parametrize('name', TINY_CASES)
@pytest.mark.parametrize('mode', ['default', 'small_tiles'])
def test_tile_k2y_emulation(name, mode):
    g = golden(name)
    Nd, Kd, Jd = geometry(g)
    tile = _small_tiles(Kd, Jd) if mode == 'small_tiles' else None
    st = helper.plan(g['om'], Nd, Kd, Jd, tile=tile, chunk_rows=3)
    y = emulate_k2y(st, Kd, Jd, g['xx2k'])
    assert rel(y, g['forward']) < 1e-6      # xx2k is stored in complex64


def test_tile_validity_and_gather_sources():
    """For every valid tile on small axes, the gather's neighbour rule finds
    exactly the (tile, accumulator index) pairs whose padded box holds the
    cell, and each padded box reaches only its own and the next tile."""
    for K in range(1, 30):
        for J in range(1, min(K, 8) + 1):
            for t in range(1, K + 2):
                if not helper.tile_valid(t, K, J):
                    continue
                tt = min(t, K)
                tg = -(-K // tt)
                pad = helper.tile_pad((tt,), (K,), (J,))[0]
                cover = {c: [] for c in range(K)}
                for td in range(tg):
                    T0 = td * tt
                    E = min(tt, K - T0)
                    span = E + J - 1 if tg > 1 else K
                    for x in range(span):
                        c = (T0 + x) % K
                        assert x < pad
                        if tg > 1:
                            assert c // tt in (td, (td + 1) % tg)
                        cover[c].append((td, x))
                for c in range(K):
                    got = sorted(gather_sources((c,), (tt,), (tg,), (K,), (J,)))
                    assert got == sorted(cover[c]), (K, J, t, c)


@pytest.mark.parametrize('name', TINY_CASES + ['plain2d'])
def test_plan_factors_bit_exact(name):
    g = golden(name)
    Nd, Kd, Jd = geometry(g)
    st = helper.plan(g['om'], Nd, Kd, Jd)
    ref = oracle.plan(g['om'], Nd, Kd, Jd)
    kst = [int(np.prod(Kd[a + 1:])) for a in range(len(Nd))]
    for a in range(len(Nd)):
        assert np.array_equal(st['alpha'][a], ref['alpha'][a])
        assert np.array_equal(st['snd'][a], ref['snd'][a][:, 0])
        assert np.array_equal(st['ud'][a], ref['ud'][a])
        assert np.array_equal(st['base'][a] * kst[a], ref['kd'][a][:, 0])
    assert np.array_equal(st['sn'], ref['sn'])
    # Kronecker products in the reference's order rebuild its CSR exactly
    M = g['om'].shape[0]
    uu = st['ud'][0]
    for a in range(1, len(Nd)):
        uu = np.einsum('mi,mj->mij', uu, st['ud'][a]).reshape(M, -1)
    assert np.array_equal(uu.ravel(), ref['p'].data)


def _small_tiles(Kd, Jd):
    """Valid tiles that split every axis that can be split (wrap paths)."""
    out = []
    for K, J in zip(Kd, Jd):
        ts = [t for t in range(max(1, J - 1), K) if helper.tile_valid(t, K, J)]
        out.append(ts[0] if ts else K)
    return tuple(out)


@pytest.mark.parametrize('name', TINY_CASES)
@pytest.mark.parametrize('mode', ['default', 'small_tiles', 'chunk3'])
def test_gridding_emulation(name, mode):
    g = golden(name)
    Nd, Kd, Jd = geometry(g)
    tile = _small_tiles(Kd, Jd) if mode == 'small_tiles' else None
    chunk = 3 if mode == 'chunk3' else None
    st = helper.plan(g['om'], Nd, Kd, Jd, tile=tile, chunk_rows=chunk)
    k = emulate_y2k(st, Kd, Jd, g['adjoint_in'])
    assert rel(k, g['y2k']) < 1e-13


@pytest.mark.parametrize('name', TINY_CASES)
@pytest.mark.parametrize('mode', ['default', 'chunk3'])
def test_packed_gridding_emulation(name, mode):
    """Packed accumulators: no real or imaginary sum reaches 2^31, unpacking
    the 64-bit totals recovers both sums exactly, and the result is as close
    to the reference's complex128 y2k as the complex64 factors allow (the
    same products summed exactly in FP64)."""
    g = golden(name)
    Nd, Kd, Jd = geometry(g)
    st = helper.plan(g['om'], Nd, Kd, Jd, chunk_rows=3 if mode == 'chunk3' else None)
    k = emulate_y2k_packed(st, Kd, Jd, g['adjoint_in'])
    A = oracle.OracleNUFFT()
    A.plan(g['om'], Nd, Kd, Jd)
    # the same complex64 factors summed exactly (FP64 products, no rounding)
    perm = st['perm']
    D = len(Kd)
    u = [st['ud'][a].astype(np.complex64).astype(np.complex128) for a in range(D)]
    exact = np.zeros(Kd, dtype=complex)
    y = np.asarray(g['adjoint_in']).astype(np.complex128)
    for m in range(len(perm)):
        for j in np.ndindex(*Jd):
            w = 1.0
            for d in range(D):
                w = w * u[d][m][j[d]]
            c = tuple((st['base'][d][m] + j[d]) % Kd[d] for d in range(D))
            exact[c] += np.conj(w) * y[m]
    e_fix = rel(k, exact)
    assert e_fix < 2e-8, e_fix
    assert rel(k, g['y2k']) < rel(exact, g['y2k']) + 2e-8


def test_work_list_invariants():
    w = W.radial2d()
    sel = np.arange(0, w['om'].shape[0], 7)
    Nd, Kd, Jd = w['Nd'], w['Kd'], w['Jd']
    st = helper.plan(w['om'][sel], Nd, Kd, Jd, chunk_rows=64)
    M = sel.size
    perm = st['perm']
    assert np.array_equal(np.sort(perm), np.arange(M))
    items = st['items'].astype(np.int64)
    tile = st['tile']
    tgrid = [(Kd[a] + tile[a] - 1) // tile[a] for a in range(2)]
    base = st['base'][:, perm]
    owner = (base[0] // tile[0]) * tgrid[1] + base[1] // tile[1]
    # rows are sorted by owning tile, ties in sample order
    assert np.all(np.diff(owner) >= 0)
    same = np.diff(owner) == 0
    assert np.all(np.diff(perm)[same] > 0)
    # items cover every row once, each within its tile, at most chunk rows
    covered = np.zeros(M, dtype=int)
    for tid, first, cnt, slot in items:
        assert 1 <= cnt <= 64
        assert np.all(owner[first:first + cnt] == tid)
        covered[first:first + cnt] += 1
    assert np.all(covered == 1)
    # one slot per item; a tile's slots are consecutive and in row order
    assert sorted(items[:, 3]) == list(range(len(items)))
    ts = st['tile_slots'].astype(np.int64)
    assert ts.shape == (int(np.prod(tgrid)), 2)
    assert ts[:, 1].sum() == len(items)
    by_slot = items[np.argsort(items[:, 3])]
    for t in np.nonzero(ts[:, 1])[0]:
        blk = by_slot[ts[t, 0]:ts[t, 0] + ts[t, 1]]
        assert np.all(blk[:, 0] == t)
        assert np.all(blk[1:, 1] == blk[:-1, 1] + blk[:-1, 2])
    # heaviest work first
    assert np.all(np.diff(items[:, 2]) <= 0)


def test_tile_choice():
    assert helper.tile_shape((256, 256, 256), (6, 6, 6)) == (16, 16, 16)
    assert helper.tile_shape((512, 512), (6, 6)) == (16, 16)
    # shrinks to fit LDS with a wide interpolator
    t = helper.tile_shape((256, 256, 256), (8, 8, 8))
    assert np.prod(helper.tile_pad(t, (256,) * 3, (8,) * 3)) * 16 + 256 <= helper.GRID_LDS_MAX
    # every default tile is valid
    for K in range(1, 200, 7):
        for J in (1, 2, 4, 6, 8):
            if J > K:
                continue
            for d in (1, 2, 3):
                t = helper.tile_shape((K,) * d, (J,) * d)
                assert all(helper.tile_valid(x, K, J) for x in t)
    with pytest.raises(ValueError):
        helper.check_tile((6, 6), (16, 16), (6, 6))       # last tile 4 < 5


def test_autotune_candidates_are_valid():
    from pynufft_amd import autotune
    for Kd, Jd in [((256, 256, 256), (6, 6, 6)), ((512, 512), (6, 6)), ((384, 384), (6, 6)),
                   ((64,), (6,)), ((100, 90), (8, 5)), ((24, 16, 20), (4, 6, 8))]:
        cands = autotune.candidates(Kd, Jd)
        assert cands and len(set(cands)) == len(cands)
        for tile, threads in cands:
            assert helper.check_tile(tile, Kd, Jd) == tile
            assert threads in (256, 512, 1024)
    assert helper.coefficient_bound([np.array([[3 + 4j, 1]]), np.array([[2, -1]])]) >= 10


def test_row_sum_bound_brute_force():
    """row_sum_bound bounds sum_j |coefficient(row, j)| over every row: the
    full Kronecker coefficients of random 3-D factors, summed directly."""
    rng = np.random.default_rng(5)
    ud = [rng.standard_normal((7, J)) + 1j * rng.standard_normal((7, J)) for J in (3, 4, 2)]
    full = np.einsum('ma,mb,mc->mabc', *ud).reshape(7, -1)
    exact = np.abs(full).sum(axis=1).max()
    b = helper.row_sum_bound(ud)
    assert exact <= b <= exact * (1 + 1e-5)
    # one row, closed form: (|3+4i| + 1) * (2 + 1)
    assert abs(helper.row_sum_bound([np.array([[3 + 4j, 1]]), np.array([[2, -1]])]) - 18) < 1e-4


def test_plan_errors_follow_reference():
    om = np.zeros((4, 2))
    with pytest.raises(TypeError):
        helper.plan(om, [8, 8], (16, 16), (4, 4))
    with pytest.raises(TypeError):
        helper.plan(om, (8, 8), [16, 16], (4, 4))
    with pytest.raises(TypeError):
        helper.plan(om, (8, 8), (16, 16), [4, 4])
    with pytest.raises(KeyError):
        helper.plan(om, (8, 8), (16, 16), (4,))
    with pytest.raises(TypeError):
        helper.plan(om, (20, 20), (40, 40), (17, 17))
    with pytest.raises(ValueError):
        helper.plan(om, (8, 8), (4, 4), (2, 2))


def test_samples_on_grid_edges():
    """om = -pi and om just below pi wrap to the first/last grid cells."""
    Nd, Kd, Jd = (8, 8), (16, 16), (6, 6)
    om = np.array([[-np.pi, -np.pi], [np.pi - 1e-12, 0.0], [0.0, 0.0],
                   [np.pi - 1e-12, np.pi - 1e-12]])
    st = helper.plan(om, Nd, Kd, Jd)
    ref = oracle.plan(om, Nd, Kd, Jd)
    for a in range(2):
        assert np.array_equal(st['ud'][a], ref['ud'][a])
        assert np.array_equal(st['base'][a] * (16 if a == 0 else 1), ref['kd'][a][:, 0])
    y = np.arange(4) + 1j
    A = oracle.OracleNUFFT()
    A.plan(om, Nd, Kd, Jd)
    assert rel(emulate_y2k(st, Kd, Jd, y), A.y2k(y)) < 1e-14


def _factor(n):
    """The device FFT's radix schedule (16 first, then 8, 4, 2, 3, 5, 7, 11, 13)."""
    out = []
    m = n
    while m > 1:
        for R in (16, 8, 4, 2, 3, 5, 7, 11, 13):
            if m % R == 0:
                out.append(R)
                m //= R
                break
        else:
            raise ValueError(n)
    return out


def _dft16(v, sign):
    """The kernel's 16-point DFT: 4-point DFTs over n1 of x[4 n1 + n2],
    twiddles W16^(n2 k1), 4-point DFTs over n2."""
    W4 = np.exp(sign * 2j * np.pi * np.outer(np.arange(4), np.arange(4)) / 4)
    a = np.array([W4 @ v[n2::4] for n2 in range(4)])            # a[n2][k1]
    a = a * np.exp(sign * 2j * np.pi * np.outer(np.arange(4), np.arange(4)) / 16)
    out = np.empty(16, complex)
    for k1 in range(4):
        out[k1::4] = W4 @ a[:, k1]
    return out


@pytest.mark.parametrize('n', [16, 64, 128, 256, 384, 512, 1000, 4096, 13 * 16])
@pytest.mark.parametrize('sign', [-1, 1])
def test_fft_stockham_schedule(n, sign):
    """The device FFT's Stockham passes (butterfly (j) reads j + r n/R, twiddle
    w^(r k n/(ns R)) with k = j % ns, writes (j // ns) ns R + k + r ns)
    reproduce numpy's FFT for every radix schedule it uses."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    tw = np.exp(-2j * np.pi * np.arange(n) / n)
    v = x.copy()
    ns = 1
    for R in _factor(n):
        nb = n // R
        out = np.empty(n, complex)
        for j in range(nb):
            k = j % ns
            e = v[j + np.arange(R) * nb].copy()
            if ns > 1:
                w = tw[np.arange(R) * k * (n // (ns * R))]
                e = e * (np.conj(w) if sign > 0 else w)
            if R == 16:
                e = _dft16(e, sign)
            else:
                e = np.exp(sign * 2j * np.pi * np.outer(np.arange(R), np.arange(R)) / R) @ e
            out[(j // ns) * ns * R + k + np.arange(R) * ns] = e
        v = out
        ns *= R
    ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
    assert np.allclose(v, ref, rtol=1e-12, atol=1e-10 * np.abs(ref).max())


def test_plan_cache_round_trip(tmp_path, monkeypatch):
    """helper.plan(cache_dir=...) stores the FP64 design per geometry and a
    second plan of the same geometry reuses it (no design computed) with
    identical results; another trajectory or shape misses; an unreadable
    file is recomputed."""
    from pynufft_amd import helper
    g = golden('tiny2d')
    Nd, Kd, Jd = geometry(g)
    a = helper.plan(g['om'], Nd, Kd, Jd, cache_dir=str(tmp_path))
    files = list(tmp_path.iterdir())
    assert len(files) == 1

    def no_design(*args):
        raise AssertionError('design recomputed')
    monkeypatch.setattr(helper, '_design', no_design)
    b = helper.plan(g['om'], Nd, Kd, Jd, cache_dir=str(tmp_path))
    for k in ('sn', 'base', 'perm', 'items', 'tile_slots'):
        assert np.array_equal(a[k], b[k]), k
    for d in range(len(Nd)):
        assert np.array_equal(a['ud'][d], b['ud'][d])
        assert np.array_equal(a['alpha'][d], b['alpha'][d]) and a['beta'][d] == b['beta'][d]
    assert a['u_bound'] == b['u_bound']
    with pytest.raises(AssertionError):
        helper.plan(g['om'][:-1], Nd, Kd, Jd, cache_dir=str(tmp_path))
    with pytest.raises(AssertionError):
        helper.plan(g['om'], Nd, tuple(k + 2 for k in Kd), Jd, cache_dir=str(tmp_path))
    monkeypatch.undo()
    files[0].write_bytes(b'not a plan')
    c = helper.plan(g['om'], Nd, Kd, Jd, cache_dir=str(tmp_path))
    assert np.array_equal(a['ud'][0], c['ud'][0])


def test_source_hash_ignores_comments():
    """The profiles' source hash covers the kernel sources' code: comments and
    whitespace runs are dropped, string literals (which may hold comment
    markers) are kept."""
    import bench
    code = 'int a = 1;  // note\n/* block\n comment */ int b;\nconst char* s = "// x /* y */";'
    assert bench._strip_comments(code) == 'int a = 1; int b; const char* s = "// x /* y */";'
    assert bench._strip_comments("char q = '\\'';  int c;") == "char q = '\\''; int c;"
    assert len(bench.source_hash()) == 16


def test_one_pass_exponent_bound():
    """The one-pass FP64 P^H P fixes each work item's fixed-point exponent
    before forming any sample: 2^e > 1.5 row_bound u_bound m, m the largest
    |Re k| + |Im k| over the item's padded input box.  Every sample the item
    forms then has |Re y| + |Im y| <= sqrt(2) row_bound m, and every
    contribution conj(coef) y stays inside the 2^50 fixed-point range, on
    random (white-noise) grids, the case that loosens the bound most."""
    rng = np.random.default_rng(11)
    for Nd, Kd, Jd, M in [((12, 10), (24, 20), (6, 6), 300), ((8, 6, 6), (16, 12, 12), (4, 5, 6), 200)]:
        D = len(Nd)
        om = rng.uniform(-np.pi, np.pi, (M, D))
        st = helper.plan(om, Nd, Kd, Jd, chunk_rows=17)
        k = rng.standard_normal(Kd) + 1j * rng.standard_normal(Kd)
        tile = st['tile']
        tgrid = [(Kd[a] + tile[a] - 1) // tile[a] for a in range(D)]
        pad = helper.tile_pad(tile, Kd, Jd)
        perm = st['perm']
        base = st['base'][:, perm]
        u = [st['ud'][a][perm] for a in range(D)]
        worst = 0.0
        for tid, first, cnt, slot in st['items']:
            T0, wrap = tile_frame(int(tid), tile, tgrid, Kd)
            box = np.array([k[tuple((T0[d] + x[d]) % Kd[d] for d in range(D))]
                            for x in np.ndindex(*pad)])
            m = np.max(np.abs(box.real) + np.abs(box.imag))
            e = np.frexp(1.5 * st['row_bound'] * st['u_bound'] * m)[1]
            for r in range(first, first + cnt):
                coefs, y = [], 0.0
                for j in np.ndindex(*Jd):
                    c = [(base[d][r] + j[d]) % Kd[d] for d in range(D)]
                    w = 1.0
                    for d in range(D):
                        w = w * u[d][r][j[d]]
                    coefs.append(w)
                    y = y + w * k[tuple(c)]
                assert abs(y.real) + abs(y.imag) <= np.sqrt(2) * st['row_bound'] * m
                for w in coefs:
                    v = np.conj(w) * y
                    worst = max(worst, abs(v.real) * 2.0 ** -e, abs(v.imag) * 2.0 ** -e)
        assert worst < 1.0
