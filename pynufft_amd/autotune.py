"""Measured choice of the gridding (y2k) work decomposition.

How fast y2k runs depends on three knobs of the plan:

* the tile shape, which sets the LDS accumulator size (tile + J - 1 per split
  axis), hence how many workgroups fit on a CU, and how many padded partial
  cells the gather pass reads back;
* the workgroup size (threads), i.e. how many samples are gridded at once
  into one accumulator;
* the rows per work item (chunk), which splits heavy tiles so that dense
  trajectory regions (the centre of radial and spiral k-space) spread over
  many CUs.

``tune`` times y2k on the device (HIP events on the launch stream) for a short
list of valid candidates and returns the fastest.  ``tune_batch`` does the same
for a ``BatchNUFFT`` by timing a few iterations of its batched FP64 CG, the
workload the batch exists for.  The FP64 host design is
computed once; each candidate only redoes the row sort.  Results are cached
per device and geometry in a JSON file (``PYNUFFT_AMD_TUNE_CACHE``, default
``~/.cache/pynufft_amd/tune.json``); ``cache=False`` skips it.
"""
import hashlib
import json
import os

import numpy as np
import torch

from . import helper

_TILES = {
    1: [(1024,), (2048,), (4096,), (512,)],
    2: [(16, 16), (32, 32), (16, 32), (32, 16), (8, 8), (64, 64)],
    3: [(16, 16, 16), (16, 16, 8), (16, 8, 8), (8, 16, 16), (8, 8, 8), (24, 16, 8)],
}
_CHUNKS = (256, 512, 1024, 2048, 4096)


def candidates(Kd, Jd):
    """(tile, threads) pairs to try: the preferred tiles moved to the nearest
    valid extent per axis, with workgroup sizes that suit the accumulator
    (large accumulators leave room for one or two workgroups per CU, so they
    get the larger workgroups)."""
    Kd, Jd = tuple(Kd), tuple(Jd)
    out = []
    seen = set()
    for pref in [helper.tile_shape(Kd, Jd)] + _TILES[len(Kd)]:
        t = tuple(helper._valid_near(p, K, J) for p, K, J in zip(pref, Kd, Jd))
        try:
            t = helper.check_tile(t, Kd, Jd)
        except ValueError:
            continue
        pad_bytes = int(np.prod(helper.tile_pad(t, Kd, Jd))) * 16
        for thr in (256, 512, 1024):
            if pad_bytes > 81920 and thr < 512:
                continue
            if pad_bytes <= 20480 and thr > 512:
                continue
            if (t, thr) not in seen:
                seen.add((t, thr))
                out.append((t, thr))
    return out


def _cache_path():
    return os.environ.get('PYNUFFT_AMD_TUNE_CACHE',
                          os.path.join(os.path.expanduser('~'), '.cache', 'pynufft_amd',
                                       'tune.json'))


def _key(device, Kd, Jd, M, nb):
    name = torch.cuda.get_device_name(device)
    mb = int(round(np.log2(max(M, 1)) * 2))      # half-octave bucket of M
    return '%s|K%s|J%s|m%d|nb%d' % (name, 'x'.join(map(str, Kd)),
                                   'x'.join(map(str, Jd)), mb, nb)


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, cfg):
    path = _cache_path()
    data = _load()
    data[key] = cfg
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + '.tmp%d' % os.getpid()
        with open(tmp, 'w') as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _time(fn, reps):
    fn()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        fn()
    stop.record()
    stop.synchronize()
    return start.elapsed_time(stop) / reps


def tune(om, Nd, Kd, Jd, batch=None, host_plan=None, device=None, reps=5,
         cache=True, chunks=_CHUNKS, verbose=False):
    """Fastest y2k configuration for this geometry on this device.

    Returns ``{'tile': tuple, 'threads': int, 'chunk': int or None, 'ms':
    float}``.  Two stages: every (tile, threads) candidate with the default
    rows per item, then every (rows per item, threads) pair for the winning
    tile.
    """
    from .nufft import NUFFT
    Nd, Kd, Jd = tuple(Nd), tuple(Kd), tuple(Jd)
    nb = 1 if batch is None else int(batch)
    st = host_plan if host_plan is not None else helper.plan(om, Nd, Kd, Jd)
    M = int(st['M'])
    dev = device if device is not None else torch.device('cuda', torch.cuda.current_device())
    key = _key(dev, Kd, Jd, M, nb)
    if cache:
        hit = _load().get(key)
        if hit is not None:
            return {'tile': tuple(hit['tile']), 'threads': int(hit['threads']),
                    'chunk': hit['chunk'], 'ms': float(hit['ms'])}
    g = torch.Generator().manual_seed(12345)
    shape = (M,) if batch is None else (M, nb)
    y = torch.complex(torch.randn(shape, generator=g), torch.randn(shape, generator=g))
    y = y.to(dev, torch.complex64)

    def measure(tile, threads, chunk):
        perm, items, ts = helper.sort_and_bin(st['base'], Kd, Jd, tile, chunk)
        s2 = dict(st)
        s2.update(perm=perm, items=items, tile_slots=ts, tile=tile)
        A = NUFFT(dev)
        A.plan(om, Nd, Kd, Jd, batch=batch, host_plan=s2, threads=threads)
        ms = _time(lambda: A.y2k(y), reps)
        A.release()
        if verbose:
            print('tune: tile %s threads %d chunk %s -> %.3f ms' % (tile, threads, chunk, ms),
                  flush=True)
        return ms

    best = None
    for tile, threads in candidates(Kd, Jd):
        ms = measure(tile, threads, None)
        if best is None or ms < best['ms']:
            best = {'tile': tile, 'threads': threads, 'chunk': None, 'ms': ms}
    # rows per item and workgroup size act together (rows per thread), so the
    # second stage tries every pair of them for the winning tile
    thr_opts = sorted({thr for t, thr in candidates(Kd, Jd) if t == best['tile']})
    for chunk in chunks:
        for thr in thr_opts:
            ms = measure(best['tile'], thr, chunk)
            if ms < best['ms']:
                best = dict(best, threads=thr, chunk=chunk, ms=ms)
    if cache:
        _store(key, {'tile': list(best['tile']), 'threads': best['threads'],
                     'chunk': best['chunk'], 'ms': best['ms']})
    return best


def resort(st, Kd, Jd, tile, chunk):
    """A host plan with its rows re-sorted for another tile / rows per item."""
    tile = helper.check_tile(tile, Kd, Jd) if tile is not None else helper.tile_shape(Kd, Jd)
    perm, items, ts = helper.sort_and_bin(st['base'], Kd, Jd, tile, chunk)
    s2 = dict(st)
    s2.update(perm=perm, items=items, tile_slots=ts, tile=tile)
    return s2


def batch_candidates(problems):
    """(tile, chunk, threads) triples for a batch: each problem's own tile
    (None) or a common smaller or larger one where valid for every problem,
    with rows per item and workgroup sizes around the defaults."""
    d = len(problems[0][2])
    tiles = [None] + {1: [(512,), (2048,)], 2: [(8, 8), (32, 32)],
                      3: [(16, 16, 8), (8, 8, 8)]}[d]
    out = []
    for t in tiles:
        if t is not None:
            try:
                for (_, _, Kd, Jd) in problems:
                    helper.check_tile(t, Kd, Jd)
            except ValueError:
                continue
        for chunk in (None, 1024):
            for thr in (0, 512):
                out.append((t, chunk, thr))
    return out


def tune_batch(batch, problems, host_plans, iters=3, reps=2, cache=True, verbose=False):
    """Fastest (tile, chunk, threads) for a BatchNUFFT: each candidate is
    planned (rows re-sorted from the given host designs) and timed on
    ``iters`` iterations of the batched CG.  Returns the winner's dict."""
    sig = hashlib.sha1(repr([(np.asarray(om).shape, float(np.asarray(om)[:64].sum()),
                              Nd, Kd, Jd)
                             for (om, Nd, Kd, Jd) in problems]).encode()).hexdigest()[:16]
    key = '%s|batch|%s' % (torch.cuda.get_device_name(batch.device), sig)
    if cache:
        hit = _load().get(key)
        if hit is not None:
            return {'tile': tuple(hit['tile']) if hit['tile'] else None,
                    'chunk': hit['chunk'], 'threads': int(hit['threads']),
                    'ms': float(hit['ms'])}
    g = torch.Generator().manual_seed(12345)
    best = None
    import concurrent.futures
    pool = concurrent.futures.ThreadPoolExecutor(8)
    for tile, chunk, thr in batch_candidates(problems):
        sts = list(pool.map(lambda a: resort(a[1], a[0][2], a[0][3], tile, chunk),
                            zip(problems, host_plans)))
        batch._build(problems, sts, thr)
        y = torch.complex(torch.randn(batch.n_samples, generator=g),
                          torch.randn(batch.n_samples, generator=g)).to(batch.device,
                                                                        torch.complex64)
        ms = _time(lambda: batch.solve(y, maxiter=iters, rtol=0.0), reps)
        if verbose:
            print('tune_batch: tile %s chunk %s threads %d -> %.3f ms' % (tile, chunk, thr, ms),
                  flush=True)
        if best is None or ms < best['ms']:
            best = {'tile': tile, 'chunk': chunk, 'threads': thr, 'ms': ms}
    pool.shutdown()
    batch.release()
    if cache:
        _store(key, {'tile': list(best['tile']) if best['tile'] else None,
                     'chunk': best['chunk'], 'threads': best['threads'], 'ms': best['ms']})
    return best


def tune_streams(batch, sets=4, iters=10, reps=2):
    """Stream sets for a BatchNUFFT's problem groups, kept by measurement.

    The runtime spreads streams over a few hardware queues when they are
    first used, and groups that share a queue run one after the other; which
    groups end up sharing is not under the caller's control, stays fixed
    for the streams' lifetime and moved the 64-problem 2-D CG between about
    3 430 and 3 600 solves/s.  So ``sets`` fresh stream sets are timed on
    ``iters`` CG iterations each and the fastest is kept (the results do not
    depend on the streams).  Returns the per-set milliseconds."""
    groups = batch._groups
    if not groups:
        return []
    g = torch.Generator().manual_seed(54321)
    y = torch.complex(torch.randn(batch.n_samples, generator=g),
                      torch.randn(batch.n_samples, generator=g)).to(batch.device,
                                                                    torch.complex64)
    best, times = None, []
    for _ in range(max(1, int(sets))):
        ss = [torch.cuda.Stream(batch.device) for _ in groups]
        batch._groups = [(a, b, sub, s) for (a, b, sub, _), s in zip(groups, ss)]
        ms = _time(lambda: batch.solve(y, maxiter=iters, rtol=0.0), reps)
        times.append(ms)
        if best is None or ms < best[0]:
            best = (ms, ss)
    batch._groups = [(a, b, sub, s) for (a, b, sub, _), s in zip(groups, best[1])]
    torch.cuda.synchronize(batch.device)
    return times
