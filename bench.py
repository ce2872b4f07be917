"""Benchmark: 3-D NUFFT forward+adjoint pairs/s on one MI355X.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]

Workload (default ``uniform3d``): Nd 128^3, Kd 256^3, Jd 6^3, M 2 097 152
coordinates uniform in [-pi, pi)^3 (seed 301), complex64 image with N(0,1)
parts (seed 302).  A step is one ``forward`` followed by one ``adjoint`` with
the image, the samples and the plan resident in HBM.  Prints one JSON line.

``roofline`` times the dominant kernel (k2y or y2k, whichever takes longer)
with HIP events on the stream it is launched on and divides the bytes the
running format must move per launch (``physical_bytes``) by the mean launch
time.  The coefficient-streaming (CSR) formula's bytes, which count
prod(Jd) coefficients per sample that the factor format never reads, are
reported separately as ``csr_equivalent``; y2k also carries its LDS-atomic
ceiling.  ``traffic`` is the measured HBM traffic per launch from
``profiles/traffic.json``, reported only when that file was measured on the
same kernel sources.  ``cpu_baseline`` times the CPU oracle (a restatement
of pynufft's NUFFT() path) on a bounded sample.
"""
import hashlib
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = ("3D 128³ NUFFT fwd+adj pairs/s (% HBM peak); "
          "batched vs plain CG solves/s")
HBM_PEAK_GBS = 8000.0
# LDS bandwidth per CU (bytes per clock): a wave-wide 64-bit atomic moves
# 64 * 8 B, so it occupies the LDS for at least 4 clocks; 256 CUs at the
# 2.4 GHz peak engine clock
LDS_BYTES_PER_CLK = 128.0
N_CU = 256
CLOCK_HZ = 2.4e9


def progress(msg):
    print('[bench] %s' % msg, file=sys.stderr, flush=True)


def algorithmic_bytes(kernel, fmt, M, Jd, Kd, nb=1):
    """Algorithmic bytes of one interpolation launch (k2y or y2k):
    B = M*P*8 (complex64 coefficients) + I + M*8*nb (samples) + K*8*nb (the
    grid once), P = prod(Jd), I = the index stream the format reads: M*P*4
    (int32 cells, CSR table) or M*d*4 (int32 first cells, factor format)."""
    d = len(Kd)
    K = int(np.prod(Kd))
    P = int(np.prod(Jd))
    index = M * P * 4 if fmt == 'CSR' else M * d * 4
    return M * P * 8 + index + M * 8 * nb + K * 8 * nb


def physical_bytes(fmt, M, Jd, Kd, nb=1):
    """Bytes the format must at least move per launch: the factor format
    reads sum(Jd) complex64 factors per sample instead of prod(Jd)
    coefficients, plus d first cells, the row->sample index, the samples
    and the grid once."""
    d = len(Kd)
    K = int(np.prod(Kd))
    P = int(np.prod(Jd))
    coef = P * 12 if fmt == 'CSR' else int(np.sum(Jd)) * 8 + d * 4
    return M * (coef + 4 + 8 * nb) + K * 8 * nb


def lds_atomic_floor_ms(M, Jd, nb=1, per_coef=1):
    """Lower bound of y2k's gridding time from the LDS bandwidth its atomics
    need, without bank conflicts: per_coef ds_add_u64 per coefficient (1
    with packed accumulators, else 2), 8 bytes per lane."""
    nbytes = per_coef * 8.0 * M * int(np.prod(Jd)) * nb
    return 1e3 * nbytes / (LDS_BYTES_PER_CLK * N_CU * CLOCK_HZ)


def _strip_comments(text):
    """C/C++ source without its comments and with runs of whitespace
    collapsed (string literals kept), so that comment edits leave the hash of
    the code unchanged."""
    out, i, n = [], 0, len(text)
    while i < n:
        c = text[i]
        if c == '"' or c == "'":
            j = i + 1
            while j < n and text[j] != c:
                j += 2 if text[j] == '\\' else 1
            out.append(text[i:j + 1])
            i = j + 1
        elif text.startswith('//', i):
            j = text.find('\n', i)
            i = n if j < 0 else j
        elif text.startswith('/*', i):
            j = text.find('*/', i + 2)
            i = n if j < 0 else j + 2
            out.append(' ')
        else:
            out.append(c)
            i += 1
    return ' '.join(''.join(out).split())


def source_hash():
    """SHA-1 (16 hex digits) of the kernel sources' code (comments and
    whitespace ignored); profiles record the hash of the build they were
    measured on."""
    h = hashlib.sha1()
    csrc = os.path.join(HERE, 'pynufft_amd', 'csrc')
    for name in sorted(os.listdir(csrc)):
        with open(os.path.join(csrc, name), encoding='utf-8') as f:
            code = _strip_comments(f.read())
        h.update(name.encode() + b'\0' + code.encode())
    return h.hexdigest()[:16]


def fft_bytes(Nd, Kd, nb=1):
    """Pruned axis-by-axis FFT traffic with pad/crop fused (one direction)."""
    d = len(Nd)
    tot = 0
    for a in range(d):
        pre = int(np.prod(Nd[:a]))
        post = int(np.prod(Kd[a + 1:]))
        tot += pre * Nd[a] * post + pre * Kd[a] * post
    return 8 * nb * tot


def time_kernel(fn, reps, torch):
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        fn()
    stop.record()
    stop.synchronize()
    return start.elapsed_time(stop) / reps  # ms


def cpu_baseline(w, m_sample, seed=7):
    """Oracle pair time on the full grid with a sample subset, scaled to M.

    t_pair = t(x2xx+xx2k) + t(k2xx+xx2x) + (t(k2y) + t(y2k)) * M / m_sample.
    """
    import oracle
    om, x = w['om'], w['x']
    M = om.shape[0]
    rng = np.random.default_rng(seed)
    sel = np.sort(rng.choice(M, m_sample, replace=False))
    A = oracle.OracleNUFFT()
    t0 = time.perf_counter()
    A.plan(om[sel], w['Nd'], w['Kd'], w['Jd'])
    t_plan = time.perf_counter() - t0
    t0 = time.perf_counter()
    k = A.xx2k(A.x2xx(x))
    t1 = time.perf_counter()
    y = A.k2y(k)
    t2 = time.perf_counter()
    k2 = A.y2k(y)
    t3 = time.perf_counter()
    A.xx2x(A.k2xx(k2))
    t4 = time.perf_counter()
    scale = M / m_sample
    t_pair = (t1 - t0) + (t4 - t3) + ((t2 - t1) + (t3 - t2)) * scale
    return {
        'value': 1.0 / t_pair, 'unit': 'pairs/s', 'cores': 1, 'kind': 'port',
        'sample': ('oracle forward+adjoint on the full %s->%s grid with %d of %d '
                   'samples; interpolation time scaled by M/%d (plan %.1f s '
                   'excluded); numpy pocketfft and scipy CSR matvec are '
                   'single-threaded' % (w['Nd'], w['Kd'], m_sample, M, m_sample,
                                        t_plan)),
    }


def bench_batched(torch, W, maxiter=20, reps=3):
    """Batched uneven 2-D CG (64 problems, gridding knobs autotuned on the
    batched CG, the solve in BatchNUFFT's default stream groups) vs looping
    the plain operator, each problem planned with the defaults as a user of
    the reference would (one NUFFT per problem)."""
    from pynufft_amd.batch import BatchNUFFT
    from pynufft_amd.nufft import NUFFT
    ps = W.batched2d()
    t0 = time.perf_counter()
    B = BatchNUFFT()
    B.plan(ps, tune=True)
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0
    plain_ops = []
    for p in ps:
        op = NUFFT()
        op.plan(p['om'], p['Nd'], p['Kd'], p['Jd'])
        plain_ops.append(op)
    xs = [torch.from_numpy(p['x']).to(B.device) for p in ps]
    y = B.forward(xs)
    ys = B.unpack_samples(y)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps, out

    progress('batched: planned %d problems in %.1f s (tuned %s)' % (len(ps), t_plan, B.tuned))
    t_plain, plain = timed(lambda: [op.solve(ys[i], 'cg', maxiter=maxiter, rtol=0.0)
                                    for i, op in enumerate(plain_ops)])
    t_batch, xb = timed(lambda: B.solve(y, maxiter=maxiter, rtol=0.0))
    res_plain = np.array([op.cg_resid for op in plain_ops])
    t_graph, xg = timed(lambda: B.solve_graph(y, maxiter=maxiter, rtol=0.0))
    res_batch = B.cg_resid.cpu().numpy()
    xgl = B.unpack_images(xg)
    diff = max(float(torch.linalg.vector_norm(xgl[i] - plain[i]) /
                     torch.linalg.vector_norm(plain[i])) for i in range(B.n))
    n = B.n
    return {
        'problems': n, 'cg_iterations': maxiter, 'samples_total': int(B.n_samples),
        'plain_loop_solves_per_s': n / t_plain,
        'batched_solves_per_s': n / t_batch,
        'batched_graph_solves_per_s': n / t_graph,
        'speedup_graph_vs_plain': t_plain / t_graph,
        'max_rel_diff_vs_plain': diff,
        'final_resid_max_rel_diff_vs_plain': float(np.max(np.abs(res_batch - res_plain) /
                                                          res_plain)),
        'final_resid_median': float(np.median(res_batch)), 'plan_s': t_plan,
        'tuned': {'tile': B.tuned['tile'], 'chunk': B.tuned['chunk'],
                  'threads': B.tuned['threads']},
        'stream_groups': max(1, len(B._groups)),
        'stream_sets_ms': [round(v, 3) for v in B.tuned.get('stream_sets_ms', [])],
    }


def _time_host(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps


def bench_radial_solvers(torch, W, maxiter=30, reps=3, cpu=True):
    """Solvers on the radial 2-D config (256^2/512^2, M 205 824), GPU and
    CPU oracle: the FP64 CG (30 iterations, rtol 0), 'dc' (3 pipe-density
    iterations), 'L1TVOLS' (10 outer iterations, rho 1) and the Toeplitz
    selfadjoint2 next to selfadjoint.  Returns (cg, others)."""
    from pynufft_amd.nufft import NUFFT
    w = W.radial2d()
    A = NUFFT()
    A.plan(w['om'], w['Nd'], w['Kd'], w['Jd'])
    x = torch.from_numpy(w['x']).to(A.device)
    y = A.forward(x)
    A.precompute_sp()
    t_cg = _time_host(torch, lambda: A.solve(y, 'cg', maxiter=maxiter, rtol=0.0), reps)
    t_cgg = _time_host(torch, lambda: A.solve_graph(y, maxiter=maxiter, rtol=0.0), reps)
    cg = {'workload': 'radial2d', 'cg_iterations': maxiter, 'gpu_solves_per_s': 1.0 / t_cg,
          'gpu_ms_per_solve': 1000.0 * t_cg, 'gpu_graph_solves_per_s': 1.0 / t_cgg}
    runs = {
        'dc3': (lambda: A.solve(y, 'dc', maxiter=3), lambda B, yc, xc: B.solve(yc, 'dc', maxiter=3)),
        'l1tvols10_rho1': (lambda: A.solve(y, 'L1TVOLS', maxiter=10, rho=1.0),
                           lambda B, yc, xc: B.solve(yc, 'L1TVOLS', maxiter=10, rho=1.0)),
        'selfadjoint2': (lambda: A.selfadjoint2(x), lambda B, yc, xc: B.selfadjoint2(xc)),
        'selfadjoint': (lambda: A.selfadjoint(x), lambda B, yc, xc: B.selfadjoint(xc)),
    }
    others = {'workload': 'radial2d'}
    for name, (fn, _) in runs.items():
        others[name] = {'gpu_ms': 1000.0 * _time_host(torch, fn, 5)}
    if cpu:
        import oracle
        B = oracle.OracleNUFFT()
        B.plan(w['om'], w['Nd'], w['Kd'], w['Jd'])
        yc, xc = y.cpu().numpy(), w['x']
        B.precompute_sp()
        t = time.perf_counter()
        B.solve(yc, 'cg', maxiter=maxiter, rtol=0.0)
        cg['cpu_oracle_solves_per_s'] = 1.0 / (time.perf_counter() - t)
        cg['cpu_cores'] = 1
        for name, (_, fc) in runs.items():
            t = time.perf_counter()
            fc(B, yc, xc)
            others[name]['cpu_oracle_ms'] = 1000.0 * (time.perf_counter() - t)
        others['cpu_cores'] = 1
    A.release()
    return cg, others


def bench_other_workloads(torch, W, steps=20):
    """forward+adjoint pairs/s of the other configurations, each planned
    with the defaults: plain 2-D 64^2 (M 4 096), radial 2-D 256^2 (M 205 824)
    and the 8-coil radial 256^2 (M 131 072) through forward_one2many +
    adjoint_many2one."""
    from pynufft_amd.nufft import NUFFT
    out = {}
    for name in ('plain2d', 'radial2d', 'multicoil2d'):
        w = getattr(W, name)()
        A = NUFFT()
        if name == 'multicoil2d':
            A.plan(w['om'], w['Nd'], w['Kd'], w['Jd'], batch=w['S'].shape[-1])
            A.set_sense(w['S'])
            x = torch.from_numpy(w['x']).to(A.device)
            fn = lambda: A.adjoint_many2one(A.forward_one2many(x))
        else:
            A.plan(w['om'], w['Nd'], w['Kd'], w['Jd'])
            x = torch.from_numpy(w['x']).to(A.device)
            fn = lambda: A.adjoint(A.forward(x))
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / steps
        out[name] = {'pairs_per_s': 1.0 / dt, 'ms_per_pair': 1000.0 * dt,
                     'Nd': list(w['Nd']), 'M': int(w['om'].shape[0]),
                     'coils': int(w['S'].shape[-1]) if name == 'multicoil2d' else 1}
        if name != 'multicoil2d':
            # the same pair replayed from a torch.cuda graph (no per-kernel
            # launch cost; the library's calls only launch, so they capture)
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, capture_error_mode='thread_local'):
                fn()
            graph.replay()
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(steps):
                graph.replay()
            torch.cuda.synchronize()
            dg = (time.perf_counter() - t) / steps
            out[name]['graph_pairs_per_s'] = 1.0 / dg
            del graph
        A.release()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--workload', default='uniform3d',
                    choices=['uniform3d', 'radial2d', 'plain2d'])
    ap.add_argument('--legacy', action='store_true',
                    help='full coefficient table (CSR) instead of factors')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-sample', type=int, default=131072)
    ap.add_argument('--no-batched', action='store_true',
                    help='skip the batched uneven-problem CG comparison')
    ap.add_argument('--only', choices=['batched', 'cg'], default=None,
                    help='run only the batched uneven-problem CG comparison or only the '
                         'radial 2-D solver timings (for profiling them alone); prints '
                         'that section as JSON')
    ap.add_argument('--copy-probe-mib', type=int, default=0,
                    help='after timing, copy this many MiB with an elementwise '
                         'kernel (known bytes, to calibrate HBM byte counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU: the operator does not shard across '
              'GPUs (see the sharding design); run with --gpus 1', file=sys.stderr)
        sys.exit(2)

    import torch
    from pynufft_amd import workloads as W
    from pynufft_amd.nufft import NUFFT

    if args.only == 'batched':
        print(json.dumps({'batched_cg': bench_batched(torch, W)}))
        return
    if args.only == 'cg':
        cg, others = bench_radial_solvers(torch, W, cpu=False)
        print(json.dumps({'cg_single': cg, 'solvers_radial2d': others}))
        return

    progress('workload %s' % args.workload)
    w = getattr(W, args.workload)()
    t0 = time.perf_counter()
    A = NUFFT(legacy=args.legacy)
    A.plan(w['om'], w['Nd'], w['Kd'], w['Jd'])
    torch.cuda.synchronize()
    t_plan = time.perf_counter() - t0
    dev = A.device
    x = torch.from_numpy(w['x']).to(dev)
    M = w['om'].shape[0]

    def pair():
        y = A.forward(x)
        return A.adjoint(y)

    progress('planned in %.1f s' % t_plan)
    for _ in range(args.warmup):
        pair()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pair()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = 1000.0 * dt / args.steps
    value = args.steps / dt

    progress('timed %d pairs: %.3f ms each' % (args.steps, ms_per_step))
    # dominant kernel: k2y vs y2k timed alone with HIP events
    k = A.xx2k(x)
    y = A.k2y(k)
    reps = max(5, min(50, args.steps))
    t_k2y = time_kernel(lambda: A.k2y(k), reps, torch)
    t_y2k = time_kernel(lambda: A.y2k(y), reps, torch)
    t_fft = time_kernel(lambda: A.xx2k(x), reps, torch)
    fmt = A._format
    packed = 'item_bound' in A._t
    if t_k2y >= t_y2k:
        kname, kms = 'k2y', t_k2y
    else:
        kname, kms = 'y2k', t_y2k
    phys = physical_bytes(fmt, M, w['Jd'], w['Kd'])
    csr_bytes = algorithmic_bytes(kname, fmt, M, w['Jd'], w['Kd'])
    achieved = phys / (kms * 1e-3) / 1e9
    traffic, traffic_note = None, 'no measurement for this workload'
    tf = os.path.join(HERE, 'profiles', 'traffic.json')
    if os.path.exists(tf):
        with open(tf) as f:
            rec = json.load(f)
        key = '%s_%s_%s' % (args.workload, fmt, kname)
        if rec.get('source_sha16') != source_hash():
            traffic_note = 'profiles/traffic.json was measured on other kernel sources'
        elif key in rec:
            traffic, traffic_note = rec[key], 'rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated'

    if args.copy_probe_mib:
        # a plain elementwise copy (reads and writes copy_probe_mib MiB each)
        src = torch.empty(args.copy_probe_mib << 18, dtype=torch.float32, device=dev)
        src.fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    batched = None
    if not args.no_batched:
        del A, k, y, x
        torch.cuda.empty_cache()
        progress('batched CG workload')
        batched = bench_batched(torch, W)
        progress('batched done')

    progress('single-problem solvers (radial 2-D)')
    cg_single, solvers = bench_radial_solvers(torch, W, cpu=not args.no_cpu_baseline) \
        if not args.no_batched else (None, None)
    others = None
    if not args.no_batched:
        progress('other workloads')
        others = bench_other_workloads(torch, W)

    cpu = None
    if not args.no_cpu_baseline:
        progress('CPU oracle baseline')
        cpu = cpu_baseline(w, min(args.cpu_sample, M))

    pair_bytes = (2 * physical_bytes(fmt, M, w['Jd'], w['Kd']) +
                  2 * fft_bytes(w['Nd'], w['Kd']))
    pair_csr_bytes = (2 * algorithmic_bytes('k2y', fmt, M, w['Jd'], w['Kd']) +
                      2 * fft_bytes(w['Nd'], w['Kd']))
    roofline = {'bound': 'hbm', 'kernel': kname, 'achieved': achieved,
                'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                'traffic': traffic, 'traffic_note': traffic_note,
                'algorithmic_bytes': phys, 'launch_ms': kms,
                'bytes_note': ('bytes the %s format must move per launch: per sample '
                               'sum(Jd) complex64 factors, ndim int32 first cells, the '
                               'int32 row index and the sample; the grid once' % fmt),
                'csr_equivalent': {
                    'bytes': csr_bytes, 'achieved': csr_bytes / (kms * 1e-3) / 1e9,
                    'note': 'M*prod(Jd)*8 coefficient bytes of a coefficient-streaming '
                            'kernel, which the factor format does not read'}}
    if kname == 'y2k':
        per_coef = 1 if packed else 2
        floor = lds_atomic_floor_ms(M, w['Jd'], per_coef=per_coef)
        roofline['lds_atomic'] = {
            'floor_ms': floor, 'frac': floor / kms,
            'note': '%d ds_add_u64 per coefficient (%s accumulators), 8 B per lane, LDS at '
                    '%.0f B/clock per CU, %d CUs at %.1f GHz, no bank conflicts; the op '
                    'also holds the exponent pass and the gather'
                    % (per_coef, 'packed' if per_coef == 1 else 'separate real/imaginary',
                       LDS_BYTES_PER_CLK, N_CU, CLOCK_HZ / 1e9)}
    kern_frac = {n: physical_bytes(fmt, M, w['Jd'], w['Kd']) / (t * 1e-3) / 1e9 / HBM_PEAK_GBS
                 for n, t in (('k2y', t_k2y), ('y2k', t_y2k))}
    kern_frac['xx2k'] = fft_bytes(w['Nd'], w['Kd']) / (t_fft * 1e-3) / 1e9 / HBM_PEAK_GBS
    out = {
        'metric': METRIC, 'value': value, 'unit': 'pairs/s', 'n_gpus': 1,
        'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': ms_per_step,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'c64', 'data': 'synthetic',
        'config': {'workload': args.workload, 'Nd': list(w['Nd']),
                   'Kd': list(w['Kd']), 'Jd': list(w['Jd']), 'M': M,
                   'format': fmt, 'step': 'forward+adjoint'},
        'roofline': roofline,
        'kernels_ms': {'k2y': t_k2y, 'y2k': t_y2k, 'xx2k': t_fft},
        'kernels_hbm_frac': kern_frac,
        'pair_hbm_frac': pair_bytes * value / 1e9 / HBM_PEAK_GBS,
        'pair_algorithmic_bytes': pair_bytes,
        'pair_csr_equivalent_bytes': pair_csr_bytes,
        'source_sha16': source_hash(),
        'plan_s': t_plan,
        'batched_cg': batched,
        'cg_single': cg_single,
        'solvers_radial2d': solvers,
        'other_workloads': others,
        'cpu_baseline': cpu,
    }
    print(json.dumps(out))


if __name__ == '__main__':
    main()
