"""NUFFT: pynufft's operator class on MI355X through the HIP C ABI.

Same public surface as the reference's ``NUFFT`` (``nufft/__init__.py:
496-728``): ``plan``, ``forward``, ``adjoint``, ``selfadjoint``, ``solve``
and the stage methods ``x2xx, xx2k, k2y, y2k, k2xx, xx2x, k2y2k,
xx2k_one2one, k2xx_one2one``.  numpy inputs give numpy outputs (like the
reference's ``_*_host`` wrappers, ``_nufft_class_methods_cpu.py:365-445``);
torch tensors give torch tensors on the plan's device.

Precision: forward/adjoint/selfadjoint and the stages run on complex64 data
with fp32 arithmetic (as the reference device path); ``k2y``/``y2k``/
``k2y2k`` given complex128 input run in FP64, and ``solve(..., 'cg')`` runs
the whole CG in FP64 as the reference CPU path does (scipy on complex128).
Plans whose apodization spans more than ``helper.FP64_SPREAD`` (max|sn| /
min|sn|, e.g. Kd == Nd) set ``fp64_ops``: their complex64 operators and
interpolation stages then run in FP64 inside (complex64 in and out), because
the interpolation cancels terms up to that spread larger than its result.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import helper

_c64 = torch.complex64
_c128 = torch.complex128
# current HIP stream of a device as an integer handle, without building a
# Stream object per call (torch's own fast path; absent -> public API)
_raw_stream = getattr(torch._C, '_cuda_getCurrentRawStream', None)


def lanes_per_sample(J_last):
    """Lanes per sample of the k2y kernel (fixes the CSR table layout)."""
    if J_last <= 4:
        return J_last
    if J_last == 6:
        return 6
    if J_last <= 8:
        return 8
    if J_last <= 16:
        return 16
    raise NotImplementedError('Jd above 16 on the last axis is not supported')


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class NUFFT:
    """Min-max NUFFT operator on one HIP device.

    ``NUFFT()`` uses the current HIP device; ``NUFFT(helper.device_list()[i])``
    selects device i.  ``legacy=True`` stores the full prod(Jd) coefficient
    table (the reference's CSR device path, ``nufft/__init__.py:538-540``);
    the default keeps per-axis factors (the reference's pELL device path).
    """

    def __init__(self, device_indx=None, legacy=None):
        lib = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError('pynufft_amd needs a HIP device (none visible)')
        stream = None
        if device_indx is None:
            dev = torch.device('cuda', torch.cuda.current_device())
        elif isinstance(device_indx, tuple):
            dev = torch.device(device_indx[4]) if isinstance(device_indx[4], torch.device) \
                else torch.device('cuda', int(device_indx[2]))
            # the tuple's queue (helper.device_list): launch on that stream
            if len(device_indx) > 5 and isinstance(device_indx[5], torch.cuda.Stream):
                stream = device_indx[5]
        elif isinstance(device_indx, torch.device):
            dev = device_indx
        else:
            dev = torch.device('cuda', int(device_indx))
        if dev.index is None:
            dev = torch.device('cuda', torch.cuda.current_device())
        self._libh = lib
        self.device = dev
        self._dev_index = dev.index
        self.thr = stream
        self._stream_handle = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        self.processor = 'hip'
        self.legacy = bool(legacy)
        self.dtype = np.complex64
        self.wavefront = 64
        self.Nd = ()
        self.Kd = ()
        self.Jd = ()
        self.ndims = 0
        self.ft_axes = ()
        self.batch = None
        self.fp64_ops = False
        self._plan = None
        self._ws = {}
        self._sense = None
        self.W = None
        self._graphs = {}

    # ------------------------------------------------------------ plan --
    def plan(self, om, Nd, Kd, Jd, ft_axes=None, batch=None, radix=None,
             tile=None, chunk=None, host_plan=None, threads=0, tune=False, cache_dir=None):
        """Design the interpolator on the host in FP64 and upload it.

        Reference ``_plan_cpu`` (``_nufft_class_methods_cpu.py:27-124``) and
        ``_plan_device`` (``_nufft_class_methods_device.py:98-181``); ``batch``
        follows the deprecated ``NUFFT_hsa.plan`` (``linalg/nufft_hsa.py:
        152-242``): coils share the trajectory, coil index fastest.
        ``radix == len(Nd)`` stores the full coefficient table like
        ``legacy=True``.  ``tile`` (gridding tile per axis), ``chunk`` (rows
        per gridding work item, at most 4096) and ``threads`` (gridding
        workgroup size) are tuning knobs; the defaults come from
        ``helper.tile_shape``, ``helper.sort_and_bin`` and the library;
        ``tune=True`` measures candidates on the device and keeps the fastest
        (``pynufft_amd.autotune``).
        ``host_plan`` may pass a ``helper.plan`` result computed elsewhere
        (e.g. in a thread pool).  ``cache_dir`` (or the
        ``PYNUFFT_AMD_PLAN_CACHE`` environment variable) keeps the FP64
        design per geometry on disk (``helper.plan``).  Returns 0.
        """
        self.release()
        st = host_plan if host_plan is not None else \
            helper.plan(om, Nd, Kd, Jd, ft_axes=ft_axes, tile=tile, chunk_rows=chunk,
                        cache_dir=cache_dir)
        if tune:
            from . import autotune
            cfg = autotune.tune(om, Nd, Kd, Jd, batch=batch, host_plan=st, device=self.device)
            perm, items, ts = helper.sort_and_bin(st['base'], Kd, Jd, cfg['tile'], cfg['chunk'])
            st = dict(st)
            st.update(perm=perm, items=items, tile_slots=ts, tile=cfg['tile'])
            threads = cfg['threads']
            self.tuned = cfg
        dd = len(Nd)
        self.ndims = dd
        self.ft_axes = tuple(range(dd)) if ft_axes is None else tuple(ft_axes)
        self.Nd, self.Kd, self.Jd = tuple(Nd), tuple(Kd), tuple(Jd)
        if batch is not None and int(batch) < 1:
            raise ValueError('batch must be >= 1')
        self.batch = None if batch is None else int(batch)
        self._nb = 1 if batch is None else int(batch)
        self._sense = None
        self.W = None
        M = int(st['M'])
        self.M = (M,)
        self.Kdprod = int(np.prod(Kd))
        self.Jdprod = int(np.prod(Jd))
        self.Ndprod = int(np.prod(Nd))
        self.prodKd = (self.Kdprod,)
        self.sn = np.asarray(st['sn'].astype(np.complex64), order='C')
        self.st = {k: st[k] for k in ('Nd', 'Kd', 'Jd', 'M', 'om', 'tol',
                                       'alpha', 'beta')}
        self._host = st
        self._threads = int(threads)
        self._format = 'CSR' if (self.legacy or (radix is not None and radix >= dd)) \
            else 'factor'
        dev = self.device
        perm = st['perm']
        self._t = {}
        t = self._t
        # (at least one element each, so that M = 0 still passes valid pointers)
        t['perm'] = torch.from_numpy(
            perm.astype(np.int32) if M else np.zeros(1, np.int32)).to(dev)
        t['base'] = torch.from_numpy(
            np.ascontiguousarray(st['base'][:, perm]).astype(np.int32)
            if M else np.zeros(dd, np.int32)).to(dev)
        u_sorted = [np.ascontiguousarray(st['ud'][a][perm]) for a in range(dd)]
        self._u_off = [0] * _lib.MAXD
        off = 0
        for a in range(dd):
            self._u_off[a] = off
            off += u_sorted[a].size
        u128 = np.concatenate([u.ravel() for u in u_sorted]) if M else \
            np.zeros(1, np.complex128)
        t['u_c64'] = torch.from_numpy(u128.astype(np.complex64)).to(dev)
        self._u128_host = u128
        # apodization spread too large for FP32 operators (Kd close to Nd):
        # forward/adjoint/selfadjoint and the complex64 stages run in FP64
        self.fp64_ops = bool(st.get('fp64_ops', False))
        if self.fp64_ops:
            t['u_c128'] = torch.from_numpy(u128).to(dev)
        t['sn_f32'] = torch.from_numpy(st['sn'].astype(np.float32).ravel()).to(dev)
        t['sn_f64'] = torch.from_numpy(st['sn'].astype(np.float64).ravel()).to(dev)
        for a in range(dd):
            tw = helper.twiddles(Kd[a])
            t['tw64_%d' % a] = torch.from_numpy(tw.astype(np.complex64)).to(dev)
            t['tw128_%d' % a] = torch.from_numpy(tw).to(dev)
        t['items'] = torch.from_numpy(np.ascontiguousarray(
            st['items'] if st['items'].size else np.zeros((1, 4), np.int32))).to(dev)
        t['tile_slots'] = torch.from_numpy(np.ascontiguousarray(st['tile_slots'])).to(dev)
        if self._format == 'CSR':
            self._build_csr()
        self._create()
        if st['items'].shape[0]:
            self._item_bounds()
        return 0

    def _item_bounds(self):
        """Per-work-item bounds of the gridding sums (a constant of the
        trajectory and work list, computed on the device once); the plan is
        re-created with them so that the complex64 gridding packs each cell's
        fixed-point real and imaginary parts into one 64-bit word."""
        n = int(self._host['items'].shape[0])
        b = torch.empty(n, dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            self._launch(self._libh.pnh_item_bounds, self._plan, _ptr(b), self._stream())
        self._t['item_bound'] = b
        self._create()

    def _ensure_du(self):
        """Upload d u_d / d om_d (FP64 on the host, stored complex64)."""
        if 'du_c64' not in self._t:
            st = self._host
            perm = st['perm']
            du = helper.factor_gradients(st)
            flat = np.concatenate([np.ascontiguousarray(du[a][perm]).ravel()
                                   for a in range(self.ndims)]) if self.M[0] else \
                np.zeros(1, np.complex128)
            self._t['du_c64'] = torch.from_numpy(flat.astype(np.complex64)).to(self.device)
            self._create()

    def _ensure_fp64(self):
        if 'u_c128' not in self._t:
            self._t['u_c128'] = torch.from_numpy(self._u128_host).to(self.device)
            self._create()

    def _build_csr(self):
        """Full prod(Jd) coefficient table from the FP64 factors (rounded to
        complex64) and its int32 grid cells, laid out for k2y's lanes: slices
        of S rows, [slice][g][row in slice][j_last] with g the flat index over
        all but the last axis and S = 64 / lanes_per_sample(J_last)."""
        st = self._host
        dd = self.ndims
        M = self.M[0]
        JL = self.Jd[-1]
        G = self.Jdprod // JL
        S = 64 // lanes_per_sample(JL)
        perm = st['perm']
        dev = self.device
        ns = (M + S - 1) // S
        val = torch.zeros((ns, G, S, JL), dtype=_c64, device=dev)
        col = torch.zeros((ns, G, S, JL), dtype=torch.int32, device=dev)
        kstride = [int(np.prod(self.Kd[a + 1:])) for a in range(dd)]
        step = 1 << 16
        for s in range(0, M, step):
            e = min(M, s + step)
            rows = perm[s:e]
            coef = torch.ones((e - s, 1), dtype=_c128, device=dev)
            cell = torch.zeros((e - s, 1), dtype=torch.int64, device=dev)
            for a in range(dd):
                ua = torch.from_numpy(np.ascontiguousarray(st['ud'][a][rows])).to(dev)
                ca = torch.from_numpy(np.mod(st['base'][a][rows][:, None] +
                                             np.arange(self.Jd[a])[None, :],
                                             self.Kd[a]) * kstride[a]).to(dev)
                coef = (coef[:, :, None] * ua[:, None, :]).reshape(e - s, -1)
                cell = (cell[:, :, None] + ca[:, None, :]).reshape(e - s, -1)
            r = torch.arange(s, e, device=dev)
            val[r // S, :, r % S, :] = coef.to(_c64).reshape(e - s, G, JL)
            col[r // S, :, r % S, :] = cell.to(torch.int32).reshape(e - s, G, JL)
        self._t['csr_val'] = val
        self._t['csr_col'] = col

    def _create(self):
        if self._plan is not None:
            self._libh.pnh_plan_destroy(self._plan)
            self._plan = None
        t = self._t
        st = self._host
        d = _lib.PlanDesc()
        d.ndim = self.ndims
        d.batch = self._nb
        d.M = self.M[0]
        for a in range(self.ndims):
            d.Nd[a] = self.Nd[a]
            d.Kd[a] = self.Kd[a]
            d.Jd[a] = self.Jd[a]
            d.u_off[a] = self._u_off[a]
            d.tw_c64[a] = t['tw64_%d' % a].data_ptr()
            d.tw_c128[a] = t['tw128_%d' % a].data_ptr()
            d.tile[a] = st['tile'][a]
        d.perm = t['perm'].data_ptr()
        d.base = t['base'].data_ptr()
        d.u_c64 = t['u_c64'].data_ptr()
        d.u_c128 = t['u_c128'].data_ptr() if 'u_c128' in t else None
        d.du_c64 = t['du_c64'].data_ptr() if 'du_c64' in t else None
        d.csr_c64 = t['csr_val'].data_ptr() if 'csr_val' in t else None
        d.csr_col = t['csr_col'].data_ptr() if 'csr_col' in t else None
        d.sn_f32 = t['sn_f32'].data_ptr()
        d.sn_f64 = t['sn_f64'].data_ptr()
        d.u_bound = float(st['u_bound'])
        d.row_bound = float(st['row_bound'])
        d.n_items = int(st['items'].shape[0])
        d.items = t['items'].data_ptr()
        d.n_tiles = int(st['tile_slots'].shape[0])
        d.tile_slots = t['tile_slots'].data_ptr()
        d.n_slots = int(st['items'].shape[0])
        d.max_tile_slots = int(st['tile_slots'][:, 1].max()) if st['tile_slots'].size else 0
        d.threads = self._threads
        d.fp64_ops = 1 if getattr(self, 'fp64_ops', False) else 0
        d.item_bound = t['item_bound'].data_ptr() if 'item_bound' in t else None
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self._libh.pnh_plan_create(ctypes.byref(d), ctypes.byref(h)))
        self._desc = d
        self._plan = h
        self._ws = {}

    def release(self):
        """Free the device plan and buffers (reference ``release``,
        ``_nufft_class_methods_device.py:653-683``)."""
        for g in getattr(self, '_graphs', {}).values():
            self._libh.pnh_graph_destroy(g[0])
        self._graphs = {}
        if getattr(self, '_plan', None) is not None:
            self._libh.pnh_plan_destroy(self._plan)
            self._plan = None
        self._t = {}
        self._ws = {}

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    # ------------------------------------------------------- plumbing --
    def _need_plan(self):
        if self._plan is None:
            raise RuntimeError('call plan() first')

    def _stream(self):
        if self._stream_handle is not None:
            return self._stream_handle
        if _raw_stream is not None:
            return ctypes.c_void_p(_raw_stream(self._dev_index))
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _workspace(self, op, dtype):
        key = (op, dtype)
        ws = self._ws.get(key)
        if ws is None:
            n = ctypes.c_size_t()
            _lib.check(self._libh.pnh_workspace_bytes(self._plan, op, dtype,
                                                      ctypes.byref(n)))
            ws = torch.empty(max(int(n.value), 256), dtype=torch.uint8,
                             device=self.device)
            self._ws[key] = ws
        return ws

    def _shape(self, kind):
        extra = () if self.batch is None else (self.batch,)
        return {'x': self.Nd, 'k': self.Kd, 'y': self.M}[kind] + extra

    def _in(self, a, kind, allow128=False, copy=False):
        """Device tensor (contiguous complex) for a numpy/torch input."""
        shape = self._shape(kind)
        is_np = not isinstance(a, torch.Tensor)
        if is_np:
            a = np.asarray(a)
            want = np.complex128 if (allow128 and a.dtype == np.complex128) \
                else np.complex64
            if a.size != int(np.prod(shape)):
                raise ValueError('expected %s of shape %s, got %s'
                                 % (kind, shape, a.shape))
            t = torch.from_numpy(np.ascontiguousarray(a.reshape(shape), dtype=want))
            t = t.to(self.device)
        else:
            want = _c128 if (allow128 and a.dtype == _c128) else _c64
            if a.numel() != int(np.prod(shape)):
                raise ValueError('expected %s of shape %s, got %s'
                                 % (kind, shape, tuple(a.shape)))
            t = a.to(device=self.device, dtype=want).reshape(shape)
            t = t.clone() if copy else t.contiguous()
        return t, is_np

    @staticmethod
    def _out(t, is_np):
        return t.cpu().numpy() if is_np else t

    def _empty(self, kind, dtype=_c64):
        return torch.empty(self._shape(kind), dtype=dtype, device=self.device)

    def _call(self, fn, *args):
        # the library launches on the current device: switch only when needed
        if torch.cuda.current_device() == self._dev_index:
            self._launch(fn, self._plan, *args)
        else:
            with torch.cuda.device(self.device):
                self._launch(fn, self._plan, *args)

    def _launch(self, fn, *args):
        """Call into the library.  A plan bound to the device tuple's stream
        launches there, ordered after torch's current stream (inputs) and
        before it (outputs, and the reuse of freed memory)."""
        if self.thr is None:
            _lib.check(fn(*args))
            return
        cur = torch.cuda.current_stream(self.device)
        self.thr.wait_stream(cur)
        _lib.check(fn(*args))
        cur.wait_stream(self.thr)

    # --------------------------------------------------------- stages --
    def x2xx(self, x):
        """xx = x * sn (reference ``_x2xx_cpu``, ``_nufft_class_methods_cpu.py:224-230``)."""
        self._need_plan()
        xt, is_np = self._in(x, 'x')
        out = self._empty('x')
        self._call(self._libh.pnh_x2xx, _ptr(xt), _ptr(out), self._stream())
        return self._out(out, is_np)

    def xx2x(self, xx):
        """x = xx * sn (reference ``_xx2x_cpu``, ``:346-351``)."""
        self._need_plan()
        xt, is_np = self._in(xx, 'x')
        out = self._empty('x')
        self._call(self._libh.pnh_xx2x, _ptr(xt), _ptr(out), self._stream())
        return self._out(out, is_np)

    def xx2k(self, xx):
        """k = fftn(zero-pad(xx)) (reference ``_xx2k_cpu``, ``:232-249``)."""
        self._need_plan()
        xt, is_np = self._in(xx, 'x')
        out = self._empty('k')
        self._call(self._libh.pnh_xx2k, _ptr(xt), _ptr(out), self._stream())
        return self._out(out, is_np)

    def k2xx(self, k):
        """xx = crop(ifftn(k)) (reference ``_k2xx_cpu``, ``:316-329``)."""
        self._need_plan()
        kt, is_np = self._in(k, 'k', copy=True)
        out = self._empty('x')
        self._call(self._libh.pnh_k2xx, _ptr(kt), _ptr(out), self._stream())
        return self._out(out, is_np)

    def xx2k_one2one(self, xx):
        """Reference ``_xx2k_one2one_cpu`` (``:251-267``)."""
        return self.xx2k(xx)

    def k2xx_one2one(self, k):
        """Reference ``_k2xx_one2one_cpu`` (``:331-344``)."""
        return self.k2xx(k)

    def _interp_input(self, a, kind):
        """Device input of an interpolation stage and whether it runs in FP64:
        complex128 input, or any input on a plan with ``fp64_ops`` (then
        upcast here and the result rounded back to complex64)."""
        t, is_np = self._in(a, kind, allow128=True)
        fp64 = t.dtype == _c128 or self.fp64_ops
        out_dtype = t.dtype
        if fp64:
            self._ensure_fp64()
            t = t.to(_c128)
        return t, is_np, fp64, out_dtype

    def k2y(self, k):
        """y = P k (reference ``_k2y_cpu``, ``:282-288``); complex128 in -> FP64."""
        self._need_plan()
        kt, is_np, fp64, odt = self._interp_input(k, 'k')
        out = self._empty('y', kt.dtype)
        self._call(self._libh.pnh_k2y, _lib.C128 if fp64 else _lib.C64,
                   _ptr(kt), _ptr(out), self._stream())
        return self._out(out.to(odt), is_np)

    def y2k(self, y):
        """k = P^H y (reference ``_y2k_cpu``, ``:308-314``); complex128 in -> FP64."""
        self._need_plan()
        yt, is_np, fp64, odt = self._interp_input(y, 'y')
        dt = _lib.C128 if fp64 else _lib.C64
        out = self._empty('k', yt.dtype)
        ws = self._workspace(_lib.OP_Y2K, dt)
        self._call(self._libh.pnh_y2k, dt, _ptr(yt), _ptr(out), _ptr(ws),
                   self._stream())
        return self._out(out.to(odt), is_np)

    def k2y2k(self, k):
        """P^H P k (reference ``_k2y2k_cpu``, ``:353-362``); complex128 in -> FP64."""
        self._need_plan()
        kt, is_np, fp64, odt = self._interp_input(k, 'k')
        dt = _lib.C128 if fp64 else _lib.C64
        out = self._empty('k', kt.dtype)
        ws = self._workspace(_lib.OP_K2Y2K, dt)
        self._call(self._libh.pnh_k2y2k, dt, _ptr(kt), _ptr(out), _ptr(ws),
                   self._stream())
        return self._out(out.to(odt), is_np)

    def forward_grad_om(self, k, g):
        """d(loss)/d(om) for y = k2y(k): Re(conj(g) dy/dom) summed over coils.

        ``k`` is the oversampled spectrum ``xx2k(x2xx(x))``, ``g`` the incoming
        gradient dL/d conj(y) (torch's convention).  Returns (M, ndim) float32.
        """
        self._need_plan()
        self._ensure_du()
        kt, is_np = self._in(k, 'k')
        gt, _ = self._in(g, 'y')
        out = torch.empty((self.M[0], self.ndims), dtype=torch.float32,
                          device=self.device)
        self._call(self._libh.pnh_forward_grad_om, _ptr(kt), _ptr(gt), _ptr(out),
                   self._stream())
        return self._out(out, is_np)

    # ------------------------------------------------------- operators --
    def forward(self, x):
        """y = k2y(xx2k(x2xx(x))) (reference ``_forward_cpu``, ``:168-179``)."""
        self._need_plan()
        xt, is_np = self._in(x, 'x')
        out = self._empty('y')
        ws = self._workspace(_lib.OP_FORWARD, _lib.C64)
        self._call(self._libh.pnh_forward, _ptr(xt), _ptr(out), _ptr(ws),
                   self._stream())
        return self._out(out, is_np)

    def adjoint(self, y):
        """x = xx2x(k2xx(y2k(y))) (reference ``_adjoint_cpu``, ``:181-193``)."""
        self._need_plan()
        yt, is_np = self._in(y, 'y')
        out = self._empty('x')
        ws = self._workspace(_lib.OP_ADJOINT, _lib.C64)
        self._call(self._libh.pnh_adjoint, _ptr(yt), _ptr(out), _ptr(ws),
                   self._stream())
        return self._out(out, is_np)

    def selfadjoint(self, x):
        """A^H A x / prod(Kd) (reference ``_selfadjoint_cpu``, ``:201-214``)."""
        self._need_plan()
        xt, is_np = self._in(x, 'x')
        out = self._empty('x')
        ws = self._workspace(_lib.OP_SELFADJOINT, _lib.C64)
        self._call(self._libh.pnh_selfadjoint, _ptr(xt), _ptr(out), _ptr(ws),
                   self._stream())
        return self._out(out, is_np)

    def solve(self, y, solver=None, *args, **kwargs):
        """Solve for the image (reference ``linalg/solve_cpu.py:185-288``).

        * ``'cg'``: CG on the k-space normal equations P^H P k = P^H y with
          scipy's semantics (x0 = 0, stop when ||r|| < max(rtol ||b||, atol) at
          the start of an iteration, ``maxiter`` default 10 * prod(Kd)), all in
          FP64 on the device; returns crop(ifftn(k)) / sn.  After the call
          ``self.cg_iters`` holds the iterations used (per coil with
          ``batch``).
        * ``'bicgstab'`` / ``'bicg'``: scipy's BiCGSTAB / BiCG on the same
          normal equations in FP64 (``rtol``, ``atol``, ``maxiter``; plans
          without ``batch``); ``self.bicg_info`` then holds (scipy's info,
          iterations).
        * ``'gmres'``: scipy's restarted GMRES on the same normal equations
          in FP64 (``restart`` <= 64, ``maxiter`` cycles, ``rtol``,
          ``atol``; plans without ``batch``); ``self.bicg_info`` then holds
          (scipy's info, inner iterations).
        * ``'lgmres'``: scipy's LGMRES on the same normal equations in FP64
          (``maxiter`` outer iterations, ``inner_m``, ``outer_k``, ``rtol``,
          ``atol``; plans without ``batch``); ``self.bicg_info`` then holds
          (scipy's info, inner iterations).
        * ``'dc'``: density-compensated adjoint, ``maxiter`` pipe-density
          iterations (``solve_cpu.py:165-183``).
        * ``'L1TVOLS'``: L1 total-variation least squares by split Bregman,
          ``maxiter`` outer iterations, regularisation ``rho``
          (``solve_cpu.py:42-163``; plans without ``batch``).
        * ``'lsqr'`` / ``'lsmr'``: scipy's LSQR / LSMR on P with k-space
          unknowns in FP64 (``solve_cpu.py:236-256``; options ``damp, atol,
          btol, conlim`` and ``iter_lim`` / ``maxiter``; plans without
          ``batch``).  ``self.lsqr_info`` then holds (istop, iterations).
        """
        if solver is None:
            raise KeyError('solver must be one of: cg, bicgstab, bicg, gmres, lgmres, dc, '
                           'L1TVOLS, lsqr, lsmr')
        if solver in ('cgs', 'qmr', 'minres'):
            raise TypeError(solver + ' requires real symmetric matrix')
        if solver == 'dc':
            return self._solve_dc(y, *args, **kwargs)
        if solver == 'L1TVOLS':
            return self._solve_l1tvols(y, *args, **kwargs)
        if solver == 'lsqr':
            return self._solve_lsqr(y, *args, **kwargs)
        if solver == 'lsmr':
            return self._solve_lsmr(y, *args, **kwargs)
        if solver in ('bicgstab', 'bicg'):
            return self._solve_bicg(y, solver, *args, **kwargs)
        if solver == 'gmres':
            return self._solve_gmres(y, *args, **kwargs)
        if solver == 'lgmres':
            return self._solve_lgmres(y, *args, **kwargs)
        if solver != 'cg':
            # the reference looks the name up in its table of scipy solvers
            raise KeyError("unknown solver %r; use 'cg', 'bicgstab', 'bicg', 'gmres', "
                           "'lgmres', 'dc', 'L1TVOLS', 'lsqr' or 'lsmr'" % (solver,))
        if args:
            raise TypeError('pass solver options as keywords (maxiter=..., rtol=...)')
        maxiter = kwargs.pop('maxiter', None)
        rtol = kwargs.pop('rtol', kwargs.pop('tol', 1e-5))
        atol = kwargs.pop('atol', 0.0)
        if kwargs:
            raise TypeError('unsupported cg options: %s' % sorted(kwargs))
        if maxiter is None:
            maxiter = 10 * self.Kdprod
        self._need_plan()
        self._ensure_fp64()
        yt, is_np = self._in(y, 'y')
        out = self._empty('x')
        ws = self._workspace(_lib.OP_SOLVE_CG, _lib.C128)
        iters = torch.zeros(self._nb, dtype=torch.int32, device=self.device)
        resid = torch.zeros(self._nb, dtype=torch.float64, device=self.device)
        self._call(self._libh.pnh_solve_cg, _ptr(yt), _ptr(out), int(maxiter),
                   float(rtol), float(atol), _ptr(ws), _ptr(iters), _ptr(resid),
                   self._stream())
        self._cg_iters_dev = iters
        self._cg_resid_dev = resid
        return self._out(out, is_np)

    def solve_graph(self, y, maxiter=30, rtol=1e-5, atol=0.0):
        """``solve(y, 'cg', maxiter=..., rtol=..., atol=...)`` through a
        captured HIP graph, built on first use for these settings (input and
        output go through buffers the graph owns); for repeated solves on one
        trajectory without per-kernel launch cost.  Same result as
        ``solve``."""
        self._need_plan()
        self._ensure_fp64()
        key = (int(maxiter), float(rtol), float(atol))
        g = self._graphs.get(key)
        if g is None:
            yb = self._empty('y')
            xb = self._empty('x')
            it = torch.zeros(self._nb, dtype=torch.int32, device=self.device)
            res = torch.zeros(self._nb, dtype=torch.float64, device=self.device)
            ws = torch.empty_like(self._workspace(_lib.OP_SOLVE_CG, _lib.C128))
            h = ctypes.c_void_p()
            torch.cuda.synchronize(self.device)
            with torch.cuda.device(self.device):
                _lib.check(self._libh.pnh_cg_graph_create(
                    self._plan, _ptr(yb), _ptr(xb), key[0], key[1], key[2], _ptr(ws),
                    _ptr(it), _ptr(res), ctypes.byref(h)))
            g = (h, yb, xb, it, res, ws)
            self._graphs[key] = g
        h, yb, xb, it, res, ws = g
        yt, is_np = self._in(y, 'y')
        yb.copy_(yt)
        with torch.cuda.device(self.device):
            self._launch(self._libh.pnh_graph_launch, h, self._stream())
        self._cg_iters_dev = it
        self._cg_resid_dev = res
        return self._out(xb.clone(), is_np)

    def _solve_dc(self, y, maxiter):
        """x = adjoint(W y), W from ``maxiter`` pipe-density iterations
        (complex128 weights, FP64 interpolation, as the reference)."""
        self._need_plan()
        self._ensure_fp64()
        if int(maxiter) < 0:
            raise ValueError('maxiter must be >= 0')
        yt, is_np = self._in(y, 'y')
        out = self._empty('x')
        ws = self._workspace(_lib.OP_SOLVE_DC, _lib.C64)
        self._call(self._libh.pnh_solve_dc, _ptr(yt), _ptr(out), int(maxiter), _ptr(ws),
                   self._stream())
        return self._out(out, is_np)

    def _solve_l1tvols(self, y, maxiter, rho):
        """Split-Bregman L1-TV least squares (mu = 1, lambda = rho)."""
        self._need_plan()
        if self.batch is not None:
            raise NotImplementedError('L1TVOLS solves plans without batch')
        if int(maxiter) < 0:
            raise ValueError('maxiter must be >= 0')
        if not float(rho) > 0:
            raise ValueError('rho must be > 0')
        yt, is_np = self._in(y, 'y')
        out = self._empty('x')
        lap = self._t.get('laplacian')
        if lap is None:
            # the reference's Laplacian spectrum (a constant of Kd)
            lap = torch.from_numpy(helper.laplacian_spectrum(self.Kd)).to(self.device)
            self._t['laplacian'] = lap
        ws = self._workspace(_lib.OP_SOLVE_L1TVOLS, _lib.C64)
        self._call(self._libh.pnh_solve_l1tvols, _ptr(yt), _ptr(out), int(maxiter),
                   float(rho), _ptr(lap), _ptr(ws), self._stream())
        return self._out(out, is_np)

    def _solve_lsqr(self, y, damp=0.0, atol=1e-6, btol=1e-6, conlim=1e8, iter_lim=None,
                    show=False, calc_var=False, x0=None):
        """scipy.sparse.linalg.lsqr semantics on P (k2y / y2k), x0 = 0."""
        self._need_plan()
        if self.batch is not None:
            raise NotImplementedError('lsqr solves plans without batch')
        if calc_var or x0 is not None:
            raise NotImplementedError('lsqr: calc_var and x0 are not supported')
        self._ensure_fp64()
        yt, is_np = self._in(y, 'y')
        out = self._empty('x')
        ws = self._workspace(_lib.OP_SOLVE_LSQR, _lib.C128)
        info = torch.zeros(2, dtype=torch.int32, device=self.device)
        self._call(self._libh.pnh_solve_lsqr, _ptr(yt), _ptr(out),
                   int(iter_lim) if iter_lim is not None else 0, float(damp), float(atol),
                   float(btol), float(conlim), _ptr(ws), _ptr(info), self._stream())
        self._lsqr_info = info
        return self._out(out, is_np)

    def _solve_bicg(self, y, solver, maxiter=None, rtol=None, atol=0.0, tol=None, x0=None,
                    M=None, callback=None):
        """scipy.sparse.linalg.bicgstab / bicg semantics on P^H P, x0 = 0."""
        self._need_plan()
        if self.batch is not None:
            raise NotImplementedError('%s solves plans without batch' % solver)
        if x0 is not None or M is not None or callback is not None:
            raise NotImplementedError('%s: x0, M and callback are not supported' % solver)
        rtol = rtol if rtol is not None else (tol if tol is not None else 1e-5)
        self._ensure_fp64()
        yt, is_np = self._in(y, 'y')
        out = self._empty('x')
        ws = self._workspace(_lib.OP_SOLVE_BICG, _lib.C128)
        info = torch.zeros(2, dtype=torch.int32, device=self.device)
        kind = _lib.BICGSTAB if solver == 'bicgstab' else _lib.BICG
        self._call(self._libh.pnh_solve_bicg, int(kind), _ptr(yt), _ptr(out),
                   int(maxiter) if maxiter is not None else 0, float(rtol), float(atol),
                   _ptr(ws), _ptr(info), self._stream())
        self._bicg_info = info
        return self._out(out, is_np)

    def _solve_gmres(self, y, restart=None, maxiter=None, rtol=None, atol=0.0, tol=None,
                     x0=None, M=None, callback=None, callback_type=None):
        """scipy.sparse.linalg.gmres semantics on P^H P, x0 = 0."""
        self._need_plan()
        if self.batch is not None:
            raise NotImplementedError('gmres solves plans without batch')
        if x0 is not None or M is not None or callback is not None:
            raise NotImplementedError('gmres: x0, M and callback are not supported')
        rtol = rtol if rtol is not None else (tol if tol is not None else 1e-5)
        restart = int(restart) if restart is not None else 0
        self._ensure_fp64()
        yt, is_np = self._in(y, 'y')
        out = self._empty('x')
        n = ctypes.c_size_t()
        with torch.cuda.device(self.device):
            _lib.check(self._libh.pnh_gmres_workspace_bytes(self._plan, restart,
                                                            ctypes.byref(n)))
        ws = torch.empty(max(int(n.value), 256), dtype=torch.uint8, device=self.device)
        info = torch.zeros(2, dtype=torch.int32, device=self.device)
        self._call(self._libh.pnh_solve_gmres, _ptr(yt), _ptr(out), restart,
                   int(maxiter) if maxiter is not None else 0, float(rtol), float(atol),
                   _ptr(ws), _ptr(info), self._stream())
        self._bicg_info = info
        return self._out(out, is_np)

    def _solve_lgmres(self, y, maxiter=1000, inner_m=30, outer_k=3, rtol=None, atol=0.0,
                      tol=None, x0=None, M=None, callback=None, outer_v=None,
                      store_outer_Av=True, prepend_outer_v=False):
        """scipy.sparse.linalg.lgmres semantics on P^H P, x0 = 0."""
        self._need_plan()
        if self.batch is not None:
            raise NotImplementedError('lgmres solves plans without batch')
        if (x0 is not None or M is not None or callback is not None or outer_v is not None
                or not store_outer_Av or prepend_outer_v):
            raise NotImplementedError('lgmres: x0, M, callback, outer_v, store_outer_Av=False '
                                      'and prepend_outer_v are not supported')
        rtol = rtol if rtol is not None else (tol if tol is not None else 1e-5)
        self._ensure_fp64()
        yt, is_np = self._in(y, 'y')
        out = self._empty('x')
        n = ctypes.c_size_t()
        with torch.cuda.device(self.device):
            _lib.check(self._libh.pnh_lgmres_workspace_bytes(self._plan, int(inner_m),
                                                             int(outer_k), ctypes.byref(n)))
        ws = torch.empty(max(int(n.value), 256), dtype=torch.uint8, device=self.device)
        info = torch.zeros(2, dtype=torch.int32, device=self.device)
        self._call(self._libh.pnh_solve_lgmres, _ptr(yt), _ptr(out), int(maxiter),
                   int(inner_m), int(outer_k), float(rtol), float(atol), _ptr(ws), _ptr(info),
                   self._stream())
        self._bicg_info = info
        return self._out(out, is_np)

    @property
    def bicg_info(self):
        """(scipy's info, iterations) of the last 'bicgstab', 'bicg' or 'gmres'
        solve."""
        return tuple(int(v) for v in self._bicg_info.cpu().numpy())

    def _solve_lsmr(self, y, damp=0.0, atol=1e-6, btol=1e-6, conlim=1e8, maxiter=None,
                    show=False, x0=None):
        """scipy.sparse.linalg.lsmr semantics on P (k2y / y2k), x0 = 0."""
        self._need_plan()
        if self.batch is not None:
            raise NotImplementedError('lsmr solves plans without batch')
        if x0 is not None:
            raise NotImplementedError('lsmr: x0 is not supported')
        self._ensure_fp64()
        yt, is_np = self._in(y, 'y')
        out = self._empty('x')
        ws = self._workspace(_lib.OP_SOLVE_LSMR, _lib.C128)
        info = torch.zeros(2, dtype=torch.int32, device=self.device)
        self._call(self._libh.pnh_solve_lsmr, _ptr(yt), _ptr(out),
                   int(maxiter) if maxiter is not None else 0, float(damp), float(atol),
                   float(btol), float(conlim), _ptr(ws), _ptr(info), self._stream())
        self._lsqr_info = info
        return self._out(out, is_np)

    @property
    def lsqr_info(self):
        """(istop, iterations) of the last 'lsqr' or 'lsmr' solve."""
        return tuple(int(v) for v in self._lsqr_info.cpu().numpy())

    def precompute_sp(self):
        """Toeplitz weights W = |xx2k(adjoint(1))| on the Kd grid (float32,
        first coil), kept as ``self.W`` (reference ``_precompute_sp_cpu``,
        ``_nufft_class_methods_cpu.py:126-145``)."""
        self._need_plan()
        W = torch.empty(self.Kd, dtype=torch.float32, device=self.device)
        ws = self._workspace(_lib.OP_PRECOMPUTE_SP, _lib.C64)
        self._call(self._libh.pnh_precompute_sp, _ptr(W), _ptr(ws), self._stream())
        self.W = W
        return W

    def selfadjoint2(self, x):
        """k2xx(W * xx2k(x)) with the Toeplitz weights, no apodization
        (reference ``_selfadjoint2_cpu``, ``:216-222``); computes W first
        when missing."""
        self._need_plan()
        if getattr(self, 'W', None) is None:
            self.precompute_sp()
        xt, is_np = self._in(x, 'x')
        out = self._empty('x')
        ws = self._workspace(_lib.OP_SELFADJOINT2, _lib.C64)
        self._call(self._libh.pnh_selfadjoint2, _ptr(xt), _ptr(self.W), _ptr(out), _ptr(ws),
                   self._stream())
        return self._out(out, is_np)

    @property
    def cg_iters(self):
        it = self._cg_iters_dev.cpu().numpy()
        return int(it[0]) if self.batch is None else it

    @property
    def cg_resid(self):
        """||r|| of the k-space normal equations at the end of the last CG
        solve (per coil with ``batch``): the value of the stopping test when
        it held, else the residual after the last iteration."""
        r = self._cg_resid_dev.cpu().numpy()
        return float(r[0]) if self.batch is None else r

    # ------------------------------------------------------ multi-coil --
    def set_sense(self, coil_profile):
        """Coil sensitivities S, shape Nd + (batch,) (reference
        ``linalg/nufft_cpu.py:166-175``, ``set_sense``)."""
        if self.batch is None:
            raise ValueError('plan with batch=<number of coils> first')
        S = torch.as_tensor(coil_profile).to(self.device, torch.complex64)
        if tuple(S.shape) != self.Nd + (self.batch,):
            raise ValueError('coil profile must have shape %s' % (self.Nd + (self.batch,),))
        self._sense = S.contiguous()

    def reset_sense(self):
        """All coil sensitivities set to 1 (reference ``reset_sense``,
        ``linalg/nufft_cpu.py:163-164``)."""
        if self.batch is None:
            raise ValueError('plan with batch=<number of coils> first')
        self._sense = torch.ones(self.Nd + (self.batch,), dtype=_c64, device=self.device)

    def _need_sense(self):
        self._need_plan()
        if self._sense is None:
            raise ValueError('set_sense(coil_profile) first')

    def forward_one2many(self, x):
        """y[:, c] = forward(x * S[..., c]) for one image x of shape Nd
        (reference ``linalg/nufft_cpu.py:177-189``; coil broadcast by the
        library's coil-expand kernel, then the batched forward)."""
        self._need_sense()
        is_np = not isinstance(x, torch.Tensor)
        if is_np:
            x = np.asarray(x)
        if int(np.prod(x.shape)) != int(np.prod(self.Nd)):
            raise ValueError('expected an image of shape %s, got %s' % (self.Nd, tuple(x.shape)))
        xt = torch.as_tensor(x).to(self.device, _c64).reshape(self.Nd).contiguous()
        out = self._empty('y')
        ws = self._workspace(_lib.OP_FORWARD_ONE2MANY, _lib.C64)
        self._call(self._libh.pnh_forward_one2many, _ptr(xt), _ptr(self._sense), _ptr(out),
                   _ptr(ws), self._stream())
        return self._out(out, is_np)

    def adjoint_many2one(self, y):
        """mean over coils of adjoint(y)[..., c] * conj(S[..., c]) (reference
        ``linalg/nufft_cpu.py:191-203``, ``numpy.mean`` over the coil axis)."""
        self._need_sense()
        yt, is_np = self._in(y, 'y')
        out = torch.empty(self.Nd, dtype=_c64, device=self.device)
        ws = self._workspace(_lib.OP_ADJOINT_MANY2ONE, _lib.C64)
        self._call(self._libh.pnh_adjoint_many2one, _ptr(yt), _ptr(self._sense), _ptr(out),
                   _ptr(ws), self._stream())
        return self._out(out, is_np)

    def selfadjoint_one2many2one(self, x):
        """adjoint_many2one(forward_one2many(x))."""
        is_np = not isinstance(x, torch.Tensor)
        xi = torch.as_tensor(np.asarray(x) if is_np else x).to(self.device, _c64)
        out = self.adjoint_many2one(self.forward_one2many(xi))
        return self._out(out, is_np)

    # ----------------------------------------------------------- misc --
    def to_device(self, x, shape=None):
        """complex64 copy on the plan's device (reference ``to_device``)."""
        a = np.asarray(x) if not isinstance(x, torch.Tensor) else x
        if isinstance(a, np.ndarray):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).to(self.device)
        return a.to(device=self.device, dtype=_c64).contiguous()

    def to_host(self, data):
        """numpy copy of a device tensor (reference ``to_host``)."""
        return data.cpu().numpy()
