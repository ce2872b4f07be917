"""The C ABI library: it loads, exports every symbol include/pynufft_hip.h
declares, its descriptor layout matches the ctypes mirror, and argument errors
come back as status codes with a message (no GPU needed for any of this)."""
import ctypes
import os
import re
import shutil
import subprocess
import tempfile

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, 'include', 'pynufft_hip.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(pnh_[a-z0-9_]+)\s*\(', text)))


@pytest.fixture(scope='module')
def lib():
    from pynufft_amd import _lib
    return _lib.load()


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 17
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from pynufft_amd import _lib
    assert set(names) == set(_lib.EXPORTS), 'ctypes prototypes out of sync'


def test_version_and_error_convention(lib):
    from pynufft_amd import _lib
    assert _lib.version().startswith('pynufft_hip')
    d = _lib.PlanDesc()
    h = ctypes.c_void_p()
    rc = lib.pnh_plan_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == -1
    assert b'ndim' in lib.pnh_last_error()
    with pytest.raises(ValueError):
        _lib.check(rc)
    n = ctypes.c_size_t()
    assert lib.pnh_workspace_bytes(None, 0, 0, ctypes.byref(n)) == -1
    assert lib.pnh_forward(None, None, None, None, None) == -1


def test_plan_desc_layout_matches_header():
    """Compile a C probe against the header; compare sizeof/offsetof."""
    from pynufft_amd import _lib
    cc = shutil.which('gcc') or shutil.which('cc')
    if cc is None:
        pytest.skip('no C compiler')
    fields = [f for f, _ in _lib.PlanDesc._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "pynufft_hip.h"',
           'int main(void){', 'printf("%zu\\n", sizeof(pnh_plan_desc));']
    src += ['printf("%%zu\\n", offsetof(pnh_plan_desc, %s));' % f for f in fields]
    src += ['return 0;}']
    with tempfile.TemporaryDirectory() as tmp:
        c = os.path.join(tmp, 'probe.c')
        exe = os.path.join(tmp, 'probe')
        with open(c, 'w') as fh:
            fh.write('\n'.join(src))
        subprocess.check_call([cc, '-I', os.path.dirname(HEADER), c, '-o', exe])
        out = [int(v) for v in subprocess.check_output([exe]).split()]
    assert out[0] == ctypes.sizeof(_lib.PlanDesc)
    for f, off in zip(fields, out[1:]):
        assert getattr(_lib.PlanDesc, f).offset == off, f


def test_operator_requires_gpu_or_raises_loudly():
    """No silent CPU fallback: without a device the operator refuses."""
    import torch
    from pynufft_amd.nufft import NUFFT
    if torch.cuda.is_available():
        pytest.skip('device present')
    with pytest.raises(RuntimeError):
        NUFFT()


def _fake_desc(ndim, Jd, tile, K=64, N=32, M=100):
    """A descriptor whose pointers are non-null placeholders: only host-side
    validation runs (the size query touches no device memory)."""
    from pynufft_amd import _lib
    d = _lib.PlanDesc()
    d.ndim = ndim
    d.batch = 1
    d.M = M
    for a in range(ndim):
        d.Nd[a] = N
        d.Kd[a] = K
        d.Jd[a] = Jd
        d.tile[a] = tile
        d.tw_c64[a] = 0x1000
        d.tw_c128[a] = 0x1000
    for f in ('perm', 'base', 'u_c64', 'u_c128', 'sn_f32', 'items', 'tile_slots'):
        setattr(d, f, 0x1000)
    d.n_items = 4
    d.n_slots = 4
    d.max_tile_slots = 1
    d.n_tiles = (-(-K // tile)) ** ndim if tile < K else 1
    d.u_bound = 1.0
    return d


def test_batch_validation_on_host(lib):
    from pynufft_amd import _lib
    n = ctypes.c_size_t()
    ok = (_lib.PlanDesc * 2)(_fake_desc(2, 6, 32), _fake_desc(2, 6, 32, K=96, N=48))
    assert lib.pnh_batch_table_bytes(ok, 2, ctypes.byref(n)) == 0 and n.value > 0
    mixed = (_lib.PlanDesc * 2)(_fake_desc(2, 6, 32), _fake_desc(3, 6, 8))
    assert lib.pnh_batch_table_bytes(mixed, 2, ctypes.byref(n)) == -1
    assert b'ndim' in lib.pnh_last_error()
    jd = (_lib.PlanDesc * 2)(_fake_desc(2, 6, 32), _fake_desc(2, 4, 32))
    assert lib.pnh_batch_table_bytes(jd, 2, ctypes.byref(n)) == -1
    assert lib.pnh_batch_table_bytes(ok, 0, ctypes.byref(n)) == -1


def test_gridding_tile_validation_on_host(lib):
    """The library refuses tiles whose padding would skip a neighbour, tiles
    whose accumulator exceeds LDS, and inconsistent work-list sizes."""
    from pynufft_amd import _lib
    n = ctypes.c_size_t()
    bad = (_lib.PlanDesc * 1)(_fake_desc(2, 6, 12))      # 64 = 5 * 12 + 4, 4 < 5
    assert lib.pnh_batch_table_bytes(bad, 1, ctypes.byref(n)) == -1
    assert b'narrower than Jd - 1' in lib.pnh_last_error()
    big = (_lib.PlanDesc * 1)(_fake_desc(3, 6, 32))       # 37^3 cells
    assert lib.pnh_batch_table_bytes(big, 1, ctypes.byref(n)) == -2
    assert b'LDS' in lib.pnh_last_error()
    d = _fake_desc(2, 6, 32)
    d.n_tiles = 3
    assert lib.pnh_batch_table_bytes((_lib.PlanDesc * 1)(d), 1, ctypes.byref(n)) == -1
    d = _fake_desc(2, 6, 32)
    d.threads = 100
    assert lib.pnh_batch_table_bytes((_lib.PlanDesc * 1)(d), 1, ctypes.byref(n)) == -1


def test_noalloc_scratch_size_matches_header():
    """The ctypes mirror's static-scratch size is the header's."""
    from pynufft_amd import _lib
    m = re.search(r'#define PNH_NOALLOC_SCRATCH_BYTES \((\d+)u << (\d+)\)', open(HEADER).read())
    assert m and int(m.group(1)) << int(m.group(2)) == _lib.NOALLOC_SCRATCH_BYTES
