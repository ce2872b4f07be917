"""Host logic of the product (no GPU): the FP64 plan builder, the trajectory
sort and the gridding work list (checked by a numpy emulation of the y2k
kernel's algorithm against the oracle's transposed-CSR matvec)."""
import numpy as np
import pytest

import oracle
from pynufft_amd import helper
from pynufft_amd import workloads as W
from conftest import TINY_CASES, golden, geometry, rel


def emulate_y2k(st, Kd, Jd, y):
    """The binned gridding algorithm in numpy: one accumulator per work item,
    samples visited in bin order, support cells clipped to the tile, split
    tiles summed in chunk order."""
    D = len(Kd)
    tile = st['tile']
    tgrid = [(Kd[a] + tile[a] - 1) // tile[a] for a in range(D)]
    perm = st['perm']
    base = st['base'][:, perm]
    u = [st['ud'][a][perm] for a in range(D)]
    out = np.full(Kd, np.nan, dtype=complex)
    part = {}
    for tid, s, cnt, slot in st['items']:
        t, T0 = tid, [0] * D
        for a in range(D - 1, -1, -1):
            T0[a] = (t % tgrid[a]) * tile[a]
            t //= tgrid[a]
        E = [min(tile[a], Kd[a] - T0[a]) for a in range(D)]
        acc = np.zeros(E, complex)
        for r in st['bin_rows'][s:s + cnt]:
            for j in np.ndindex(*Jd):
                l = [(base[a][r] + j[a]) % Kd[a] - T0[a] for a in range(D)]
                if all(0 <= l[a] < E[a] for a in range(D)):
                    w = 1
                    for a in range(D):
                        w = w * u[a][r][j[a]]
                    acc[tuple(l)] += np.conj(w) * y[perm[r]]
        sl = tuple(slice(T0[a], T0[a] + E[a]) for a in range(D))
        if slot < 0:
            out[sl] = acc
        else:
            part[slot] = (sl, acc)
    for tid, s0, n in st['merge']:
        sl = part[s0][0]
        out[sl] = sum(part[s0 + k][1] for k in range(n))
    return out


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
