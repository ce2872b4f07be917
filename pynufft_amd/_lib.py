"""ctypes binding of libpynufft_hip.so (C ABI declared in include/pynufft_hip.h).

torch is imported first so that the library binds to the HIP runtime torch
already loaded (both carry the soname libamdhip64.so.7).  There is no CPU
fallback: if the library is missing, every operator raises.
"""
import ctypes
import os

import torch  # noqa: F401  (must precede the library load)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libpynufft_hip.so')

MAXD = 3
C64, C128 = 0, 1
(OP_FORWARD, OP_ADJOINT, OP_SELFADJOINT, OP_Y2K, OP_K2Y2K, OP_SOLVE_CG,
 OP_FORWARD_ONE2MANY, OP_ADJOINT_MANY2ONE, OP_PRECOMPUTE_SP, OP_SELFADJOINT2,
 OP_SOLVE_DC, OP_SOLVE_L1TVOLS, OP_SOLVE_LSQR, OP_SOLVE_LSMR, OP_SOLVE_BICG) = range(15)
BICGSTAB, BICG = 0, 1
NOALLOC_SCRATCH_BYTES = 32 << 20   # PNH_NOALLOC_SCRATCH_BYTES

_vp = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int32)


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ('ndim', ctypes.c_int32),
        ('batch', ctypes.c_int32),
        ('M', ctypes.c_int64),
        ('Nd', ctypes.c_int64 * MAXD),
        ('Kd', ctypes.c_int64 * MAXD),
        ('Jd', ctypes.c_int32 * MAXD),
        ('perm', _vp),
        ('base', _vp),
        ('u_c64', _vp),
        ('u_c128', _vp),
        ('du_c64', _vp),
        ('u_off', ctypes.c_int64 * MAXD),
        ('csr_c64', _vp),
        ('csr_col', _vp),
        ('sn_f32', _vp),
        ('sn_f64', _vp),
        ('tw_c64', _vp * MAXD),
        ('tw_c128', _vp * MAXD),
        ('tile', ctypes.c_int32 * MAXD),
        ('u_bound', ctypes.c_double),
        ('n_items', ctypes.c_int64),
        ('items', _vp),
        ('n_tiles', ctypes.c_int64),
        ('tile_slots', _vp),
        ('n_slots', ctypes.c_int64),
        ('max_tile_slots', ctypes.c_int32),
        ('threads', ctypes.c_int32),
        ('fp64_ops', ctypes.c_int32),
        ('item_bound', _vp),
        ('row_bound', ctypes.c_double),
    ]


class HipError(RuntimeError):
    """A HIP runtime failure reported by the library."""


_lib = None

EXPORTS = {
    'pnh_version': (ctypes.c_char_p, []),
    'pnh_last_error': (ctypes.c_char_p, []),
    'pnh_device_count': (ctypes.c_int, []),
    'pnh_plan_create': (ctypes.c_int, [ctypes.POINTER(PlanDesc), ctypes.POINTER(_vp)]),
    'pnh_plan_destroy': (ctypes.c_int, [_vp]),
    'pnh_workspace_bytes': (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_size_t)]),
    'pnh_item_bounds': (ctypes.c_int, [_vp, _vp, _vp]),
    'pnh_x2xx': (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    'pnh_xx2x': (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    'pnh_xx2k': (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    'pnh_k2xx': (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    'pnh_k2y': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp]),
    'pnh_y2k': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp]),
    'pnh_k2y2k': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp]),
    'pnh_forward': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    'pnh_adjoint': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    'pnh_selfadjoint': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    'pnh_forward_grad_om': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    'pnh_forward_one2many': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    'pnh_adjoint_many2one': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    'pnh_precompute_sp': (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    'pnh_selfadjoint2': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    'pnh_solve_dc': (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp, _vp]),
    'pnh_solve_l1tvols': (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double,
                                         _vp, _vp, _vp]),
    'pnh_solve_lsqr': (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      _vp, _vp, _vp]),
    'pnh_solve_bicg': (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double, _vp, _vp, _vp]),
    'pnh_gmres_workspace_bytes': (ctypes.c_int, [_vp, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_size_t)]),
    'pnh_lgmres_workspace_bytes': (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int,
                                                  ctypes.POINTER(ctypes.c_size_t)]),
    'pnh_solve_lgmres': (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        _vp, _vp, _vp]),
    'pnh_solve_gmres': (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double, _vp, _vp, _vp]),
    'pnh_solve_lsmr': (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      _vp, _vp, _vp]),
    'pnh_solve_cg': (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, _vp, _vp, _vp, _vp]),
    'pnh_solve_cg_noalloc': (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, _vp, _vp, _vp]),
    'pnh_batch_table_bytes': (ctypes.c_int, [ctypes.POINTER(PlanDesc), ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_size_t)]),
    'pnh_batch_create': (ctypes.c_int, [ctypes.POINTER(PlanDesc), ctypes.c_int, _vp, _vp,
                                        ctypes.POINTER(_vp)]),
    'pnh_batch_destroy': (ctypes.c_int, [_vp]),
    'pnh_batch_sizes': (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_int64)]),
    'pnh_batch_workspace_bytes': (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_size_t)]),
    'pnh_batch_forward': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    'pnh_batch_adjoint': (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    'pnh_batch_solve_cg': (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double, _vp, _vp, _vp, _vp]),
    'pnh_batch_cg_graph_create': (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double,
                                                 ctypes.c_double, _vp, _vp, _vp,
                                                 ctypes.POINTER(_vp)]),
    'pnh_cg_graph_create': (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, _vp, _vp, _vp,
                                           ctypes.POINTER(_vp)]),
    'pnh_graph_launch': (ctypes.c_int, [_vp, _vp]),
    'pnh_graph_destroy': (ctypes.c_int, [_vp]),
}


def load(path=LIB_PATH):
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError('libpynufft_hip.so not found at %s; build it with '
                          '`python -m pynufft_amd.build`' % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc == 0:
        return
    msg = load().pnh_last_error().decode(errors='replace')
    if rc == -1:
        raise ValueError(msg)
    if rc == -2:
        raise NotImplementedError(msg)
    raise HipError(msg)


def device_count():
    n = load().pnh_device_count()
    if n < 0:
        check(n)
    return n


def version():
    return load().pnh_version().decode()
