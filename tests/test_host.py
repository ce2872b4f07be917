"""Host logic of the product (no GPU): the FP64 plan builder, the trajectory
sort and the gridding work list (checked by a numpy emulation of the y2k
kernel's algorithm against the oracle's transposed-CSR matvec)."""
import numpy as np
import pytest

import oracle
from pynufft_amd import helper
from pynufft_amd import workloads as W
from conftest import TINY_CASES, golden, geometry, rel


def _wrap(b, j, K):
    c = b + j
    return c - K if c >= K else c


def clip_axis(b, J, K, T0, E):
    """Mirror of the kernel's clip_axis (csrc/interp.hip)."""
    if E == K:
        return (1 if b + J <= K else 0), 0, J
    T1 = T0 + E
    l1, h1 = max(0, T0 - b), min(min(J, K - b), T1 - b)
    l2, h2 = max(K - b, T0 - b + K), min(J, T1 - b + K)
    a1, a2 = h1 > l1, h2 > l2
    if a1 and a2:
        return 0, 0, 0
    if a1:
        return 1, l1, h1 - l1
    if a2:
        return 1, l2, h2 - l2
    return 1, 0, 0


def emulate_y2k(st, Kd, Jd, y):
    """The y2k kernel's algorithm in numpy, step for step: one accumulator per
    work item, entries in bin order, support boxes clipped to the tile
    (packed start cell, 16-bit reciprocal division) or, for a wrapped support
    in a tile that (nearly) spans an axis, the masked walk; split tiles summed
    in chunk order."""
    D = len(Kd)
    tile = st['tile']
    tgrid = [(Kd[a] + tile[a] - 1) // tile[a] for a in range(D)]
    perm = st['perm']
    base = st['base'][:, perm]
    u = [st['ud'][a][perm] for a in range(D)]
    cells = int(np.prod(tile))
    J = list(Jd) + [1] * (3 - D)
    K = list(Kd) + [1] * (3 - D)
    out = np.full(Kd, np.nan, dtype=complex)
    parts = {}
    for tid, s, cnt, slot in st['items']:
        t, T0, E = tid, [0, 0, 0], [1, 1, 1]
        for d in range(D - 1, -1, -1):
            T0[d] = (t % tgrid[d]) * tile[d]
            t //= tgrid[d]
            E[d] = min(tile[d], Kd[d] - T0[d])
        E1, E2 = E[1], E[2]
        acc = np.zeros(cells, complex)
        for r in st['bin_rows'][s:s + cnt]:
            bb = [int(base[d][r]) for d in range(D)] + [0] * (3 - D)
            lo, n, st0, single = [0, 0, 0], [1, 1, 1], [0, 0, 0], 1
            for d in range(D):
                ok, lo[d], n[d] = clip_axis(bb[d], J[d], K[d], T0[d], E[d])
                single &= ok
                st0[d] = _wrap(bb[d], lo[d], K[d]) - T0[d]
            yv = y[perm[r]]
            if single:
                tot = n[0] * n[1] * n[2]
                start = (st0[0] * E1 + st0[1]) * E2 + st0[2]
                m2 = (65536 + n[2] - 1) // max(n[2], 1)
                m1 = (65536 + n[1] - 1) // max(n[1], 1)
                for q in range(tot):
                    if tot <= 256:
                        tt = (q * m2) >> 16
                        t2 = q - tt * n[2]
                        t0 = (tt * m1) >> 16
                        t1 = tt - t0 * n[1]
                    else:
                        t2, tt = q % n[2], q // n[2]
                        t1, t0 = tt % n[1], tt // n[1]
                    w = u[0][r][lo[0] + t0]
                    if D > 1:
                        w = w * u[1][r][lo[1] + t1]
                    if D > 2:
                        w = w * u[2][r][lo[2] + t2]
                    acc[start + (t0 * E1 + t1) * E2 + t2] += np.conj(w) * yv
            else:
                for j in np.ndindex(*Jd):
                    l = [_wrap(bb[d], j[d], K[d]) - T0[d] for d in range(D)]
                    if all(0 <= l[d] < E[d] for d in range(D)):
                        l = l + [0] * (3 - D)
                        w = 1
                        for d in range(D):
                            w = w * u[d][r][j[d]]
                        acc[(l[0] * E1 + l[1]) * E2 + l[2]] += np.conj(w) * yv
        blk = acc[:E[0] * E1 * E2].reshape([E[d] for d in range(D)])
        sl = tuple(slice(T0[d], T0[d] + E[d]) for d in range(D))
        if slot < 0:
            out[sl] = blk
        else:
            parts[slot] = (sl, blk)
    for tid, s0, nn in st['merge']:
        sl = parts[s0][0]
        out[sl] = sum(parts[s0 + k][1] for k in range(nn))
    return out


def test_clip_axis_exhaustive():
    """clip_axis returns exactly the hit offsets, as one contiguous run of
    cells, or flags the case (two runs) for the masked walk."""
    for K in range(1, 26):
        for J in range(1, min(K, 8) + 1):
            for tile in range(1, K + 1):
                for T0 in range(0, K, tile):
                    E = min(tile, K - T0)
                    for b in range(K):
                        hits = [j for j in range(J) if T0 <= (b + j) % K < T0 + E]
                        single, lo, n = clip_axis(b, J, K, T0, E)
                        if single:
                            assert hits == list(range(lo, lo + n))
                            cells = [(b + j) % K for j in hits]
                            assert cells == list(range(cells[0], cells[0] + n)) if n else True
                        else:
                            runs = [(b + j) % K for j in hits]
                            assert any(runs[i + 1] != runs[i] + 1 for i in range(len(runs) - 1))


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


@pytest.mark.parametrize('name', TINY_CASES)
@pytest.mark.parametrize('chunk', [None, 3])
def test_binned_gridding_emulation(name, chunk):
    g = golden(name)
    Nd, Kd, Jd = geometry(g)
    st = helper.plan(g['om'], Nd, Kd, Jd, chunk_entries=chunk)
    k = emulate_y2k(st, Kd, Jd, g['adjoint_in'])
    assert not np.isnan(k).any(), 'some grid cells were never written'
    assert rel(k, g['y2k']) < 1e-14


def test_work_list_invariants():
    w = W.radial2d()
    sel = np.arange(0, w['om'].shape[0], 7)
    Nd, Kd, Jd = w['Nd'], w['Kd'], w['Jd']
    st = helper.plan(w['om'][sel], Nd, Kd, Jd, chunk_entries=64)
    M = sel.size
    perm = st['perm']
    assert np.array_equal(np.sort(perm), np.arange(M))
    items = st['items'].astype(np.int64)
    tgrid = [(Kd[a] + st['tile'][a] - 1) // st['tile'][a] for a in range(2)]
    # every tile is written by at least one item; chunks tile each bin exactly
    assert set(items[:, 0]) == set(range(int(np.prod(tgrid))))
    total = 0
    for t in np.unique(items[:, 0]):
        it = items[items[:, 0] == t]
        it = it[np.argsort(it[:, 1])]
        assert np.all(it[1:, 1] == it[:-1, 1] + it[:-1, 2])
        assert (len(it) > 1) == (it[0, 3] >= 0)
        total += it[:, 2].sum()
    assert total == st['bin_rows'].size
    # (row, tile) pairs are unique and each bin is in ascending sample order
    for t0 in range(0, items.shape[0], 97):
        tid, s, c, _ = items[t0]
        rows = st['bin_rows'][s:s + c]
        assert len(np.unique(rows)) == len(rows)
        assert np.all(np.diff(perm[rows]) > 0)
    # heaviest work first
    assert np.all(np.diff(items[:, 2]) <= 0)


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
