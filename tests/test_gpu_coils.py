"""Multi-coil operators (reference ``linalg/nufft_cpu.py:163-203``):
forward_one2many broadcasts one image over the coil sensitivities and runs
the batched forward; adjoint_many2one combines the coil adjoints with the
conjugate sensitivities and averages over coils (``numpy.mean``).  Checked
against the oracle's single-coil operators, relative L2 <= 1e-5."""
import numpy as np
import pytest

import oracle
from pynufft_amd import workloads as W
from conftest import golden, geometry, rel

pytestmark = pytest.mark.gpu


def _sense(Nd, nb, seed):
    rng = np.random.default_rng(seed)
    S = rng.standard_normal(Nd + (nb,)) + 1j * rng.standard_normal(Nd + (nb,))
    return (S / np.sqrt((np.abs(S) ** 2).sum(-1, keepdims=True))).astype(np.complex64)


@pytest.mark.parametrize('nb', [1, 3, 8])
@pytest.mark.parametrize('name', ['tiny1d', 'tiny2d', 'tiny3d_mixed', 'plain2d'])
def test_one2many_many2one_match_oracle(hip, name, nb):
    from pynufft_amd.nufft import NUFFT
    g = golden(name)
    Nd, Kd, Jd = geometry(g)
    S = _sense(Nd, nb, 11 + nb)
    x = g['x'].astype(np.complex64)
    A = NUFFT()
    A.plan(g['om'], Nd, Kd, Jd, batch=nb)
    A.set_sense(S)
    B = oracle.OracleNUFFT()
    B.plan(g['om'], Nd, Kd, Jd)
    y = A.forward_one2many(x)
    assert y.shape == (g['om'].shape[0], nb)
    for c in range(nb):
        xs = (x * S[..., c]).astype(np.complex64)
        assert rel(y[:, c], B.forward(xs)) < 1e-5
    xa = A.adjoint_many2one(y)
    assert xa.shape == Nd
    ref = np.mean(np.stack([B.adjoint(y[:, c]) * S[..., c].conj() for c in range(nb)], -1), -1)
    assert rel(xa, ref) < 1e-5
    xs2 = A.selfadjoint_one2many2one(x)
    assert rel(xs2, A.adjoint_many2one(A.forward_one2many(x))) < 1e-6


def test_coil_operator_arguments(hip):
    import torch
    from pynufft_amd.nufft import NUFFT
    g = golden('tiny2d')
    Nd, Kd, Jd = geometry(g)
    A = NUFFT()
    A.plan(g['om'], Nd, Kd, Jd)
    with pytest.raises(ValueError):
        A.set_sense(np.ones(Nd + (1,), np.complex64))     # no batch
    A.plan(g['om'], Nd, Kd, Jd, batch=2)
    with pytest.raises(ValueError):
        A.forward_one2many(g['x'])                         # no sense yet
    with pytest.raises(ValueError):
        A.set_sense(np.ones(Nd + (3,), np.complex64))     # wrong coil count
    A.reset_sense()
    x = torch.tensor(g['x'].astype(np.complex64), device=A.device)
    y = A.forward_one2many(x)                              # torch in, torch out
    assert isinstance(y, torch.Tensor) and tuple(y.shape) == (g['om'].shape[0], 2)
    assert torch.equal(y[:, 0], y[:, 1])
    assert rel(y[:, 0].cpu().numpy(), NUFFT_single(g).forward(g['x'])) < 1e-6
    xa = A.adjoint_many2one(y)
    assert rel(xa.cpu().numpy(), NUFFT_single(g).adjoint(y[:, 0].cpu().numpy())) < 1e-6
    A.plan(g['om'], Nd, Kd, Jd, batch=2)                   # re-plan drops the sense
    with pytest.raises(ValueError):
        A.adjoint_many2one(y)


def NUFFT_single(g):
    from pynufft_amd.nufft import NUFFT
    A = NUFFT()
    A.plan(g['om'], *geometry(g))
    return A
