"""Many independent NUFFT problems per launch.

``BatchNUFFT`` holds problems of uneven Nd, Kd and M (same ndim and Jd), each
planned exactly like a single ``NUFFT``.  Every operator stage runs for all
problems in one kernel launch, driven by per-problem descriptors and
per-stage block tables on the device; ``solve`` runs the FP64 CG of all
problems together (per-problem scalars and stopping) and can be captured once
into a HIP graph and replayed.  The reference has one problem per NUFFT
object; looping ``NUFFT.solve`` over problems is the baseline this replaces.

Vectors are packed back to back in problem order; ``pack``/``unpack`` convert
from/to lists.

``solve`` and ``solve_graph`` can split the problems into groups of about
equal work, each its own native batch on its own stream, so that one
group's gridding (VALU- and LDS-bound) overlaps another group's CG vector
steps (HBM-bound).  Each problem's arithmetic, and so its result, does not
depend on the grouping.
"""
import concurrent.futures
import ctypes

import numpy as np
import torch

from . import _lib
from . import helper
from .nufft import NUFFT, _ptr


class BatchNUFFT:
    def __init__(self, device=None):
        self._libh = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError('pynufft_amd needs a HIP device (none visible)')
        self.device = device if device is not None else \
            torch.device('cuda', torch.cuda.current_device())
        self._h = None
        self._graphs = {}
        self._ws = {}
        self._groups = []

    def plan(self, problems, tile=None, chunk=None, threads=0, tune=False, streams=8):
        """``problems``: sequence of (om, Nd, Kd, Jd) (or dicts with those keys).
        ``tile``/``chunk``/``threads``: gridding knobs applied to every
        problem (default: each problem's own tile, see ``NUFFT.plan``).
        ``tune=True`` picks them by timing a short batched CG over a list of
        candidates (``pynufft_amd.autotune.tune_batch``).  ``streams`` > 1
        splits the CG solves into that many problem groups of about equal
        work (at most one per problem), run concurrently on their own
        streams; on the 64-problem 2-D workload 1, 4, 8 and 16 groups gave
        about 3 180, 3 420, 3 420-3 600 and 3 210 solves/s.  With ``tune``
        the groups' streams are also picked by timing a few stream sets
        (``pynufft_amd.autotune.tune_streams``)."""
        self.release()
        probs = [(p['om'], tuple(p['Nd']), tuple(p['Kd']), tuple(p['Jd']))
                 if isinstance(p, dict) else
                 (p[0], tuple(p[1]), tuple(p[2]), tuple(p[3])) for p in problems]
        if not probs:
            raise ValueError('need at least one problem')
        # host FP64 designs in parallel (numpy releases the GIL)
        with concurrent.futures.ThreadPoolExecutor(8) as pool:
            sts = list(pool.map(
                lambda p: helper.plan(p[0], p[1], p[2], p[3], tile=tile, chunk_rows=chunk),
                probs))
        if tune:
            from . import autotune
            cfg = autotune.tune_batch(self, probs, sts)
            tile, chunk, threads = cfg['tile'], cfg['chunk'], cfg['threads']
            sts = [autotune.resort(st, p[2], p[3], tile, chunk) for p, st in zip(probs, sts)]
            self.tuned = cfg
        self._build(probs, sts, threads)
        self._build_groups(int(streams))
        if tune and self._groups:
            from . import autotune
            self.tuned['stream_sets_ms'] = autotune.tune_streams(self)
        return 0

    def _build(self, probs, sts, threads):
        """Upload every problem and create the batch tables."""
        self.release()
        self.ops = []
        for (om, Nd, Kd, Jd), st in zip(probs, sts):
            op = NUFFT(self.device)
            op.plan(om, Nd, Kd, Jd, host_plan=st, threads=threads)
            op._ensure_fp64()
            self.ops.append(op)
        self._create()

    def _create(self):
        """Batch tables and native handle over the planned problems in
        ``self.ops``."""
        n = len(self.ops)
        self.n = n
        self.Nd = [op.Nd for op in self.ops]
        self.Kd = [op.Kd for op in self.ops]
        self.M = [op.M[0] for op in self.ops]
        self._descs = (_lib.PlanDesc * n)(*[op._desc for op in self.ops])
        nbytes = ctypes.c_size_t()
        _lib.check(self._libh.pnh_batch_table_bytes(self._descs, n, ctypes.byref(nbytes)))
        self._table = torch.empty(max(256, nbytes.value), dtype=torch.uint8,
                                  device=self.device)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self._libh.pnh_batch_create(self._descs, n, _ptr(self._table),
                                                   self._stream(), ctypes.byref(h)))
        self._h = h
        ni, ns, ng = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self._libh.pnh_batch_sizes(h, ctypes.byref(ni), ctypes.byref(ns),
                                              ctypes.byref(ng)))
        self.n_image, self.n_samples, self.n_grid = ni.value, ns.value, ng.value
        self._x_off = np.concatenate([[0], np.cumsum([int(np.prod(d)) for d in self.Nd])])
        self._y_off = np.concatenate([[0], np.cumsum(self.M)])

    def _build_groups(self, streams):
        """Contiguous problem groups of about equal work (samples times
        support plus grid cells), each a native batch sharing this batch's
        planned problems, with one side stream each."""
        self._groups = []
        if streams <= 1 or self.n < 2:
            return
        ng = min(streams, self.n)
        jp = int(np.prod(self.ops[0].Jd))
        w = np.array([m * jp + int(np.prod(k)) for m, k in zip(self.M, self.Kd)], dtype=float)
        cuts = np.searchsorted(np.cumsum(w), w.sum() * np.arange(1, ng) / ng, side='right')
        bounds = [0] + [int(c) for c in cuts] + [self.n]
        for a, b in zip(bounds[:-1], bounds[1:]):
            if b <= a:
                continue
            sub = BatchNUFFT(self.device)
            sub.ops = self.ops[a:b]
            sub._create()
            self._groups.append((a, b, sub, torch.cuda.Stream(self.device)))

    def release(self):
        for a, b, sub, s in getattr(self, '_groups', []):
            sub.release()
        self._groups = []
        for g in getattr(self, '_graphs', {}).values():
            self._libh.pnh_graph_destroy(g[0])
        self._graphs = {}
        if getattr(self, '_h', None) is not None:
            self._libh.pnh_batch_destroy(self._h)
            self._h = None
        self._ws = {}

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    # ----------------------------------------------------------- packing --
    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def pack_images(self, xs):
        return torch.cat([torch.as_tensor(x).to(self.device, torch.complex64).reshape(-1)
                          for x in xs])

    def pack_samples(self, ys):
        return torch.cat([torch.as_tensor(y).to(self.device, torch.complex64).reshape(-1)
                          for y in ys])

    def unpack_images(self, x):
        return [x[self._x_off[i]:self._x_off[i + 1]].reshape(self.Nd[i])
                for i in range(self.n)]

    def unpack_samples(self, y):
        return [y[self._y_off[i]:self._y_off[i + 1]] for i in range(self.n)]

    def _workspace(self, op):
        ws = self._ws.get(op)
        if ws is None:
            nb = ctypes.c_size_t()
            _lib.check(self._libh.pnh_batch_workspace_bytes(self._h, op, 0, ctypes.byref(nb)))
            ws = torch.empty(max(256, nb.value), dtype=torch.uint8, device=self.device)
            self._ws[op] = ws
        return ws

    def _call(self, fn, *args):
        with torch.cuda.device(self.device):
            _lib.check(fn(self._h, *args))

    # --------------------------------------------------------- operators --
    def forward(self, x):
        """Packed images (or a list) -> packed samples, all problems at once."""
        xt = self.pack_images(x) if isinstance(x, (list, tuple)) else \
            x.to(self.device, torch.complex64).contiguous()
        y = torch.empty(self.n_samples, dtype=torch.complex64, device=self.device)
        ws = self._workspace(_lib.OP_FORWARD)
        self._call(self._libh.pnh_batch_forward, _ptr(xt), _ptr(y), _ptr(ws), self._stream())
        return y

    def adjoint(self, y):
        yt = self.pack_samples(y) if isinstance(y, (list, tuple)) else \
            y.to(self.device, torch.complex64).contiguous()
        x = torch.empty(self.n_image, dtype=torch.complex64, device=self.device)
        ws = self._workspace(_lib.OP_ADJOINT)
        self._call(self._libh.pnh_batch_adjoint, _ptr(yt), _ptr(x), _ptr(ws), self._stream())
        return x

    def _solve_groups(self, method, y, maxiter, rtol, atol):
        """Each problem group's ``method`` on its own stream, after the
        caller's stream has produced y; the caller's stream then waits for
        all of them."""
        yt = self.pack_samples(y) if isinstance(y, (list, tuple)) else \
            y.to(self.device, torch.complex64).contiguous()
        cur = torch.cuda.current_stream(self.device)
        x = torch.empty(self.n_image, dtype=torch.complex64, device=self.device)
        iters = torch.zeros(self.n, dtype=torch.int32, device=self.device)
        resid = torch.zeros(self.n, dtype=torch.float64, device=self.device)
        start = torch.cuda.Event()
        start.record(cur)
        ends = []
        for a, b, sub, s in self._groups:
            s.wait_event(start)
            with torch.cuda.stream(s):
                xg = getattr(sub, method)(yt[self._y_off[a]:self._y_off[b]], maxiter, rtol,
                                          atol)
                x[self._x_off[a]:self._x_off[b]].copy_(xg)
                iters[a:b].copy_(sub.cg_iters)
                resid[a:b].copy_(sub.cg_resid)
            for t in (yt, x, iters, resid):
                t.record_stream(s)
            e = torch.cuda.Event()
            e.record(s)
            ends.append(e)
        for e in ends:
            cur.wait_event(e)
        self.cg_iters = iters
        self.cg_resid = resid
        return x

    def solve(self, y, maxiter=20, rtol=1e-5, atol=0.0):
        """FP64 CG of every problem (scipy semantics per problem)."""
        if self._groups:
            return self._solve_groups('solve', y, maxiter, rtol, atol)
        yt = self.pack_samples(y) if isinstance(y, (list, tuple)) else \
            y.to(self.device, torch.complex64).contiguous()
        x = torch.empty(self.n_image, dtype=torch.complex64, device=self.device)
        iters = torch.zeros(self.n, dtype=torch.int32, device=self.device)
        resid = torch.zeros(self.n, dtype=torch.float64, device=self.device)
        ws = self._workspace(_lib.OP_SOLVE_CG)
        self._call(self._libh.pnh_batch_solve_cg, _ptr(yt), _ptr(x), int(maxiter),
                   float(rtol), float(atol), _ptr(ws), _ptr(iters), _ptr(resid),
                   self._stream())
        self.cg_iters = iters
        self.cg_resid = resid
        return x

    def solve_graph(self, y, maxiter=20, rtol=1e-5, atol=0.0):
        """Same as ``solve`` through a captured HIP graph (built on first use
        for these settings; input/output go through buffers the graph owns)."""
        if self._groups:
            return self._solve_groups('solve_graph', y, maxiter, rtol, atol)
        key = (int(maxiter), float(rtol), float(atol))
        g = self._graphs.get(key)
        if g is None:
            yb = torch.empty(self.n_samples, dtype=torch.complex64, device=self.device)
            xb = torch.empty(self.n_image, dtype=torch.complex64, device=self.device)
            it = torch.zeros(self.n, dtype=torch.int32, device=self.device)
            res = torch.zeros(self.n, dtype=torch.float64, device=self.device)
            ws = torch.empty_like(self._workspace(_lib.OP_SOLVE_CG))
            h = ctypes.c_void_p()
            torch.cuda.synchronize(self.device)
            with torch.cuda.device(self.device):
                _lib.check(self._libh.pnh_batch_cg_graph_create(
                    self._h, _ptr(yb), _ptr(xb), key[0], key[1], key[2], _ptr(ws),
                    _ptr(it), _ptr(res), ctypes.byref(h)))
            g = (h, yb, xb, it, res, ws)
            self._graphs[key] = g
        h, yb, xb, it, res, ws = g
        yt = self.pack_samples(y) if isinstance(y, (list, tuple)) else y
        yb.copy_(yt.to(self.device, torch.complex64).reshape(-1))
        with torch.cuda.device(self.device):
            _lib.check(self._libh.pnh_graph_launch(h, self._stream()))
        self.cg_iters = it
        self.cg_resid = res
        return xb.clone()
