"""Throughput benchmark: the combined column workload on one MI355X.

Workload (one "step" = one pass of the hot path over one batch): 2048 x 1024
columns x 137 sigma levels, float32, most-unstable parcel, pseudo-adiabatic
ascent with ice, pinc 500 Pa, plus storm-relative helicity 0-3 km with Bunkers
motion, all in one fused kernel launch per step.  Inputs are generated in HBM
(seeded synthetic soundings and winds) before the timed region; outputs stay
in HBM.

Prints one JSON line:
  value        columns/s of the timed steps (whole job, one GPU)
  roofline     the fused kernel against the FP32 compute peak: algorithmic
               operations per launch (unit-weighted FP32 ops of the plain
               method, from the kernel's own level/sub-step/iteration counts)
               divided by the mean launch time measured with HIP events on the
               launch stream
  cpu_baseline the reference Fortran (compiled from source, oracle/_ref) on the
               host cores over a bounded sample of the same workload, or the C
               restatement of the oracle when that build is absent
  accuracy     CAPE/CIN error of a column sample against the FP32 oracle
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NCOL_X, NCOL_Y, NLEV = 2048, 1024, 137
PINC, SOURCE, ADIABAT, DEPTH = 500.0, 2, 3, 3000.0
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0

# unit-weighted FP32 operations per unit of plain-method work: environment per
# level (26), ascent bookkeeping per level (35), per sub-step (6), per
# fixed-point iteration with the ice fraction (74), per most-unstable level
# scanned (26, theta-e)
W_ENV, W_LEVEL, W_SUBSTEP, W_ITER_ICE, W_ITER_LIQ, W_MU = 26, 35, 6, 74, 69, 26


def algorithmic_ops(stats, nk, ncol, mu_levels):
    iters = stats['iterations']
    w_it = W_ITER_ICE if ADIABAT in (3, 4) else W_ITER_LIQ
    return (W_ENV * (nk + 1) * stats['columns'] + W_MU * mu_levels + W_LEVEL * stats['levels']
            + W_SUBSTEP * stats['substeps'] + w_it * iters)


def bytes_per_column(nk):
    # compulsory HBM bytes: p, t, td, u, v profiles + 5 surface values in,
    # cape, cin, mulev, zmulev (4 B) + srh_rm, srh_lm (8 B) + rm, lm, mean (8 B)
    return 4 * (5 * nk + 5) + 4 * 4 + 2 * 8 + 3 * 8


def cpu_baseline(host, threads, seconds):
    """Reference Fortran (or the C oracle) on `threads` host threads over
    chunks of the same input columns, until `seconds` of wall time have
    passed; columns/s of the chunks completed (ctypes releases the GIL)."""
    import threading
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O
    from oracle import ref_driver as R

    kind = 'reference' if R.available() else 'port'
    ncol = host['t'].shape[0]
    chunk = 1024
    nxt = [0]
    lock = threading.Lock()
    deadline = time.perf_counter() + seconds

    def run_chunk(a):
        if kind == 'reference':
            R.loopcape_ml(a['p'].T, a['t'].T, a['td'].T, a['ps'], a['ts'], a['tds'], PINC,
                          SOURCE, 500., ADIABAT)
            h, hs = R.stdheight_ml(a['p'].T, a['t'].T, a['td'].T, a['ps'], a['ts'], a['tds'], 2.)
            rm, lm, mw = R.bunkers_ml_timing(a['u'].T, a['v'].T, h.astype(np.float32), a['us'],
                                             a['vs'], hs.astype(np.float32))
            R.sreh_ml(a['u'].T, a['v'].T, h, a['us'], a['vs'], hs, rm[0], rm[1], lm[0], lm[1],
                      DEPTH)
        else:
            O.cape(a['p'], a['t'], a['td'], a['ps'], a['ts'], a['tds'], PINC, SOURCE, 500.,
                   ADIABAT)
            O.srh(a['p'], a['t'], a['td'], a['u'], a['v'], a['ps'], a['ts'], a['tds'], a['us'],
                  a['vs'], DEPTH)

    def worker(_):
        done = 0
        while time.perf_counter() < deadline:
            with lock:
                i0 = nxt[0]
                nxt[0] += chunk
            if i0 >= ncol:
                break
            idx = slice(i0, min(i0 + chunk, ncol))
            run_chunk({k: v[idx] for k, v in host.items()})
            done += idx.stop - idx.start
        return done

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        n = sum(ex.map(worker, range(threads)))
    dt = time.perf_counter() - t0
    src = 'reference Fortran compiled from source' if kind == 'reference' else 'C restatement'
    return {'value': n / dt, 'unit': 'columns/s', 'cores': threads, 'kind': kind,
            'sample': f'{n} columns of the same inputs (137 sigma levels, MU, pseudo-ice, '
                      f'pinc 500, + stdheight/Bunkers/SREH 0-3 km), {src}, {dt:.1f} s wall '
                      f'on {threads} threads'}


def accuracy(outs, inp, sample):
    """CAPE/CIN of the first `sample` columns against the FP32 oracle."""
    from oracle import oracle as O
    a = {k: v[:sample].cpu().numpy() for k, v in inp.items()}
    o = O.cape(a['p'], a['t'], a['td'], a['ps'], a['ts'], a['tds'], PINC, SOURCE, 500., ADIABAT)
    res = {}
    for name, g, r in (('cape', outs[0], o[0]), ('cin', outs[1], o[1])):
        g = g[:sample].cpu().numpy().astype(np.float64)
        r = r.astype(np.float64)
        d = np.abs(g - r)
        tol = np.maximum(0.5, 1e-3 * np.abs(r))
        res[name] = {'max_abs_err': float(d.max()), 'frac_within_tol': float(np.mean(d <= tol))}
    res['columns'] = sample
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--fast-math', action='store_true')
    ap.add_argument('--plain', action='store_true',
                    help='report the cold-start kernel as the value (default: warm start)')
    ap.add_argument('--cpu-sample', type=int, default=1 << 20,
                    help='columns copied to the host for the CPU baseline (upper bound)')
    ap.add_argument('--cpu-seconds', type=float, default=10.0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--ncol', type=int, default=NCOL_X * NCOL_Y)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU; --gpus {args.gpus} is not supported '
              f'(columns are independent: a multi-GPU run would shard them with no exchange)',
              file=sys.stderr)
        sys.exit(2)

    import torch

    from xcape_amd import _hip, synthetic
    torch.cuda.set_device(0)
    ncol = args.ncol
    print(f'[bench] generating {ncol} columns x {NLEV} levels on the GPU', file=sys.stderr,
          flush=True)
    inp = synthetic.sigma_columns_torch(ncol, NLEV, 4004, device='cuda', winds=True)
    torch.cuda.synchronize()
    keys = ('p', 't', 'td', 'u', 'v', 'ps', 'ts', 'tds', 'us', 'vs')
    a = [inp[k] for k in keys]

    def step(warm, stats=False):
        return _hip.cape_srh(*a, pinc=PINC, source=SOURCE, adiabat=ADIABAT, depth=DEPTH,
                             fast_math=args.fast_math, warm_start=warm, stats=stats)

    def timed(warm):
        for _ in range(args.warmup):
            step(warm)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            outs = step(warm)
        e1.record(stream)
        torch.cuda.synchronize()
        return outs, e0.elapsed_time(e1) / args.steps, time.perf_counter() - t0

    # counting passes: level / sub-step / iteration totals of both methods
    st_plain = step(False, stats=True)[-1]
    st_warm = step(True, stats=True)[-1]
    torch.cuda.synchronize()
    mu_levels = int((inp['p'] >= 500.0).sum().item()) + ncol   # levels scanned at p >= 500 hPa
    # the plain method's work is credited to both kernels
    ops = algorithmic_ops(st_plain, NLEV, ncol, mu_levels)

    outs_p, ms_plain, _ = timed(False)
    warm = not args.plain
    outs, ms, wall = timed(warm) if warm else (outs_p, ms_plain, None)
    value = ncol / (ms / 1e3)
    achieved = ops / (ms / 1e3) / 1e12
    acc = accuracy(outs, inp, 2048)
    acc_plain = accuracy(outs_p, inp, 2048)
    cpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        nhost = min(ncol, args.cpu_sample)
        host = {k: v[:nhost].cpu().numpy() for k, v in inp.items()}
        print(f'[bench] CPU baseline on {threads} threads, {args.cpu_seconds} s', file=sys.stderr,
              flush=True)
        cpu = cpu_baseline(host, threads, args.cpu_seconds)
    del a, inp

    line = {
        'metric': 'columns/sec CAPE+CIN, surrogate-warm-start vs plain GPU, at <=0.5 J/kg '
                  'CAPE error',
        'value': value,
        'unit': 'columns/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms,
        'wall_s_timed': wall,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic (seeded soundings and winds generated in HBM)',
        'config': {'workload': 'combined: 2048x1024 columns x 137 sigma levels, most-unstable, '
                               'pseudo-ice, pinc 500 Pa, + calc_srh depth 3000 m with Bunkers '
                               'motion, fused kernel',
                   'ncol': ncol, 'nlev': NLEV, 'warm_start': warm,
                   'fast_math': bool(args.fast_math), 'parallelism': 'single GPU'},
        'roofline': {'bound': 'mfma', 'pipe': 'VALU FP32 (no matrix work in the plain method)',
                     'achieved': achieved, 'peak': FP32_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': achieved / FP32_PEAK_TFLOPS, 'traffic': None,
                     'ops_per_launch': ops,
                     'hbm_gbs_algorithmic': bytes_per_column(NLEV) * value / 1e9},
        'plain_kernel': {'value': ncol / (ms_plain / 1e3), 'ms_per_step': ms_plain,
                         'counts': st_plain, 'accuracy_vs_oracle': acc_plain},
        'speedup_vs_plain': ms_plain / ms,
        'kernel_counts': st_warm if warm else st_plain,
        'accuracy_vs_oracle': acc,
        'cpu_baseline': cpu,
    }
    print(json.dumps(line))


if __name__ == '__main__':
    main()
