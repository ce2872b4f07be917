"""CPU oracle for the min-max NUFFT hot path -- TEST INFRASTRUCTURE ONLY.

This package is a from-scratch numpy/scipy restatement of pynufft's CPU
``NUFFT()`` path (plan construction in FP64, the six operator stages, the
Toeplitz ``selfadjoint2``, the solvers of ``solve_cpu.solve`` and the FP64
direct NUDFT).  It is the checker that the HIP
product path is compared against.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or the timed CPU
baseline.  The product package ``pynufft_amd`` never imports it and has no CPU
fallback: when the HIP library is missing it raises.

Pinning (what checks each function against something other than itself) is
listed per function in ``nufft_oracle.py`` and enforced by
``tests/test_oracle_golden.py`` and ``tests/test_oracle_math.py``.
"""
from .nufft_oracle import (  # noqa: F401
    OracleNUFFT, OracleNUDFT, plan, fit_alpha_beta, kb_shape, kb_ft,
    scale_factor_1d, interp_T, interp_r, interp_u, interp_cols, offset,
    pad_index_maps, cg_scipy, laplacian_spectrum, diff_back, diff_fwd, l1tvols,
)
