"""Ragged column batches on the persistent gfx950 kernel, with an autotuner.

A ``RaggedBatch`` holds a CSR batch in HBM (torch tensors): column i's levels
are ``offsets[i]:offsets[i+1]`` of the concatenated p, t, td (hPa, degC), and
each column has its own parcel source (1 surface, 2 most-unstable, 3
mixed-layer) and adiabat (1-4, as ``calc_cape``).  ``RaggedPlan`` wraps the
C-ABI plan (work order, launch geometry, captured hipGraph); ``autotune``
times the plan over wavefront chunk sizes and LDS level-tile depths and keeps
the fastest, cached per device and batch signature in a JSON file.

Warm-started plans run each adiabat's columns through the grid path's warm
kernels (a setup pass, then a persistent ascent whose lanes take columns one
at a time), which have no chunk or tile to tune: their autotune times the one
plan.
"""
import ctypes
import hashlib
import json
import os
import time

import numpy as np

from . import _hip

CHUNKS = (64, 128, 256)
TILES = (0, 8, 16, 32)
# columns per workgroup, one per lane: 16 and 32 leave lanes of the
# wavefront idle, 64 / 128 / 256 are one, two and four full wavefronts
COLS_PER_BLOCK = (16, 32, 64, 128, 256)
WARM_COLS_PER_BLOCK = (64, 128, 256)
DEFAULT = dict(chunk=64, tile=16, cols=256)
CACHE = os.environ.get('XCAPE_AUTOTUNE_CACHE',
                       os.path.join(os.path.expanduser('~'), '.cache', 'xcape_amd',
                                    'autotune.json'))


class RaggedBatch:
    """Device-resident ragged batch (torch tensors on the GPU)."""

    def __init__(self, offsets, p, t, td, ps, ts, tds, source, adiabat, device='cuda'):
        import torch

        def dev(a, dt):
            return torch.as_tensor(np.ascontiguousarray(a)).to(device=device, dtype=dt)
        self.offsets_host = np.ascontiguousarray(offsets, np.int64)
        self.source_host = np.ascontiguousarray(source, np.int32)
        self.adiabat_host = np.ascontiguousarray(adiabat, np.int32)
        self.ncol = len(self.offsets_host) - 1
        self.offsets = dev(self.offsets_host, torch.int64)
        self.source = dev(self.source_host, torch.int32)
        self.adiabat = dev(self.adiabat_host, torch.int32)
        self.p, self.t, self.td = (dev(a, torch.float32) for a in (p, t, td))
        self.ps, self.ts, self.tds = (dev(a, torch.float32) for a in (ps, ts, tds))
        self.cape = torch.empty(self.ncol, dtype=torch.float32, device=device)
        self.cin = torch.empty_like(self.cape)
        self.zmulev = torch.empty_like(self.cape)
        self.mulev = torch.empty(self.ncol, dtype=torch.int32, device=device)

    @classmethod
    def from_dict(cls, d, device='cuda'):
        return cls(d['offsets'], d['p'], d['t'], d['td'], d['ps'], d['ts'], d['tds'],
                   d['source'], d['adiabat'], device=device)

    def signature(self):
        h = hashlib.sha1()
        for a in (self.offsets_host, self.source_host, self.adiabat_host):
            h.update(a.tobytes())
        return h.hexdigest()[:16]


class RaggedPlan:
    def __init__(self, batch, pinc=500., ml_depth=500., chunk=64, tile=16, cols=256,
                 fast_math=False, warm_start=False):
        self.batch = batch
        self.cfg = dict(chunk=chunk, tile=tile, cols=cols)
        b = batch
        P = _hip._ptr
        h = ctypes.c_void_p()
        flags = (_hip.XCAPE_FAST_MATH if fast_math else 0) | \
            (_hip.XCAPE_WARM_START if warm_start else 0)
        rc = _hip.lib().xcape_ragged_plan_create(
            P(b.offsets_host), P(b.source_host), P(b.adiabat_host), b.ncol,
            P(b.offsets), P(b.source), P(b.adiabat), P(b.p), P(b.t), P(b.td), P(b.ps), P(b.ts),
            P(b.tds), pinc, ml_depth, P(b.cape), P(b.cin), P(b.mulev), P(b.zmulev),
            chunk, tile, cols, flags, ctypes.byref(h))
        _hip.check(rc)
        self._h = h

    def run(self, stats=False):
        import torch
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        st = _hip.Stats()
        _hip.check(_hip.lib().xcape_ragged_plan_run(self._h, ctypes.byref(st) if stats else None,
                                                    stream))
        return st.as_dict() if stats else None

    def info(self):
        g, b = ctypes.c_int(), ctypes.c_int()
        _hip.check(_hip.lib().xcape_ragged_plan_info(self._h, ctypes.byref(g), ctypes.byref(b)))
        return dict(grid=g.value, block=b.value)

    def close(self):
        if self._h:
            _hip.lib().xcape_ragged_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _time_plan(plan, reps=3):
    import torch
    plan.run()                       # capture + first replay
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        plan.run()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _load_cache():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save_cache(c):
    os.makedirs(os.path.dirname(CACHE), exist_ok=True)
    tmp = CACHE + '.tmp'
    with open(tmp, 'w') as f:
        json.dump(c, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def autotune(batch, pinc=500., ml_depth=500., fast_math=False, chunks=CHUNKS, tiles=TILES,
             cols=None, reps=3, use_cache=True, log=None, warm_start=False):
    """Time every (chunk, tile, columns-per-workgroup) plan on `batch` and
    return (best config, {config: ms}).  Warm-started plans have no chunk or
    tile: their sweep is over the persistent ascent's columns per workgroup.
    Results are cached per device, batch signature and method settings."""
    import torch
    dev = torch.cuda.get_device_name(0)
    key = f'{dev}|{batch.signature()}|{pinc}|{ml_depth}|{int(fast_math)}|{int(warm_start)}'
    cache = _load_cache() if use_cache else {}
    if key in cache:
        ent = cache[key]
        return ent['best'], ent['times']
    times = {}
    best, best_ms = None, float('inf')
    if cols is None:
        cols = WARM_COLS_PER_BLOCK if warm_start else COLS_PER_BLOCK
    if warm_start:
        # the warm kernels take neither a chunk nor a tile
        chunks, tiles = (DEFAULT['chunk'],), (0,)
        cols = tuple(c for c in cols if c >= 64)
    for ch in chunks:
        for tl in tiles:
            for cb in cols:
                plan = RaggedPlan(batch, pinc, ml_depth, ch, tl, cb, fast_math, warm_start)
                t0 = time.perf_counter()
                ms = _time_plan(plan, reps)
                plan.close()
                name = f'chunk={ch},tile={tl},cols={cb}'
                times[name] = ms
                if log:
                    log(f'{name}: {ms:.3f} ms ({time.perf_counter() - t0:.1f} s)')
                if ms < best_ms:
                    best_ms, best = ms, dict(chunk=ch, tile=tl, cols=cb)
    if use_cache:
        cache[key] = {'best': best, 'times': times}
        _save_cache(cache)
    return best, times


def calc_cape_ragged(offsets, p, t, td, ps, ts, tds, source, adiabat, pinc=500.,
                     ml_depth=500., fast_math=False, tune=False):
    """One-shot ragged call on host arrays: returns numpy cape, cin, mulev,
    zmulev (ncol,).  Same per-column semantics as ``calc_cape``."""
    b = RaggedBatch(offsets, p, t, td, ps, ts, tds, source, adiabat)
    cfg = autotune(b, pinc, ml_depth, fast_math)[0] if tune else DEFAULT
    plan = RaggedPlan(b, pinc, ml_depth, fast_math=fast_math, **cfg)
    plan.run()
    import torch
    torch.cuda.synchronize()
    out = tuple(x.cpu().numpy() for x in (b.cape, b.cin, b.mulev, b.zmulev))
    plan.close()
    return out
