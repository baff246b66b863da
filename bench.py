"""Throughput benchmark: the combined column workload on one MI355X.

Workload (one "step" = one pass of the hot path over one batch): 2048 x 1024
columns x 137 sigma levels, float32, most-unstable parcel, pseudo-adiabatic
ascent with ice, pinc 500 Pa, plus storm-relative helicity 0-3 km with Bunkers
motion on the same columns (xcape_cape_srh_f32).  Inputs are generated in HBM
(seeded synthetic soundings and winds) before the timed region; outputs stay
in HBM.

Prints one JSON line:
  value         columns/s of the warm-started path (surrogate + extrapolation)
  plain_kernel  the same step with the reference's cold-started method
  roofline      the dominant kernel (the warm-started ascent, CAPE only) against
                the FP32 peak: algorithmic operations per launch (unit-weighted
                FP32 ops of the reference method, from the kernel's own level /
                sub-step / iteration counters) over its mean launch time,
                measured with HIP events on the launch stream
  attribution   the CAPE-only times of the reference loops, the warm start's
                loop with cold starts, extrapolation only, and with the surrogate
  accuracy      CAPE/CIN/MUlev/zMUlev of a column sample against the FP32 oracle
                and against the FP64 re-run of the GPU kernel; SRH and Bunkers
                against the oracle
  cpu_baseline  the reference Fortran (compiled from source, oracle/_ref) on the
                host cores over a bounded sample of the same workload, or the C
                restatement of the oracle when that build is absent

--profile-pass runs each kernel variant once (plus a 1 GiB device copy for
counter calibration) and prints nothing else, for rocprofv3 counter runs.
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
SEED = 4004

# unit-weighted FP32 operations per unit of reference-method work:
# environment per level (26), ascent bookkeeping per level (35), per
# sub-step (6), per fixed-point iteration with the ice fraction (74) or
# without (69), per most-unstable level scanned (26, theta-e)
W_ENV, W_LEVEL, W_SUBSTEP, W_ITER_ICE, W_ITER_LIQ, W_MU = 26, 35, 6, 74, 69, 26
PMC_FILE = os.path.join(ROOT, 'profiles', 'cape_warm_combined_2048x1024x137_pmc.json')


def algorithmic_ops(stats, nk, mu_levels):
    w_it = W_ITER_ICE if ADIABAT in (3, 4) else W_ITER_LIQ
    return (W_ENV * (nk + 1) * stats['columns'] + W_MU * mu_levels + W_LEVEL * stats['levels']
            + W_SUBSTEP * stats['substeps'] + w_it * stats['iterations'])


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


def full_grid_agreement(outs, ref):
    """CAPE, CIN, MUlev of every column against a reference run on the GPU,
    with the tolerance max(0.5 J/kg, 0.1 %)."""
    res = {'columns': int(ref[0].numel())}
    for i, name in enumerate(('cape', 'cin')):
        g, r = outs[i].double(), ref[i].double()
        d = (g - r).abs()
        tol = (1e-3 * r.abs()).clamp(min=0.5)
        res[name] = {'max_abs_diff': float(d.max()),
                     'frac_within_tol': float((d <= tol).double().mean())}
    res['mulev_frac_equal'] = float((outs[2] == ref[2]).double().mean())
    return res


def accuracy(outs, inp, sample, fp64_outs):
    """Outputs of the first `sample` columns against the FP32 oracle (CAPE,
    CIN, MUlev, zMUlev, SRH, Bunkers) and against the FP64 re-run of the GPU
    kernel (CAPE, CIN, MUlev, zMUlev)."""
    from oracle import oracle as O
    a = {k: v[:sample].cpu().numpy() for k, v in inp.items()}
    o = O.cape(a['p'], a['t'], a['td'], a['ps'], a['ts'], a['tds'], PINC, SOURCE, 500., ADIABAT)
    g = [x[:sample].cpu().numpy() for x in outs]
    f64 = [x[:sample].cpu().numpy() for x in fp64_outs]
    res = {}
    for i, name in enumerate(('cape', 'cin')):
        d = np.abs(g[i].astype(np.float64) - o[i])
        tol = np.maximum(0.5, 1e-3 * np.abs(o[i].astype(np.float64)))
        res[name] = {'max_abs_err': float(d.max()), 'frac_within_tol': float(np.mean(d <= tol)),
                     'max_abs_err_vs_fp64_gpu': float(np.abs(g[i] - f64[i]).max())}
    res['mulev'] = {'frac_equal': float(np.mean(g[2] == o[2])),
                    'frac_equal_fp64_gpu': float(np.mean(g[2] == f64[2]))}
    res['zmulev'] = {'max_abs_err': float(np.abs(g[3].astype(np.float64) - o[3]).max()),
                     'max_abs_err_vs_fp64_gpu': float(np.abs(g[3] - f64[3]).max())}
    srh = O.srh(a['p'], a['t'], a['td'], a['u'], a['v'], a['ps'], a['ts'], a['tds'], a['us'],
                a['vs'], DEPTH)
    res['srh'] = {'max_abs_err': float(max(np.abs(g[4] - srh[0]).max(),
                                           np.abs(g[5] - srh[1]).max()))}
    res['bunkers'] = {'max_abs_err': float(max(np.abs(g[6] - srh[2]).max(),
                                               np.abs(g[7] - srh[3]).max(),
                                               np.abs(g[8] - srh[4]).max()))}
    res['columns'] = sample
    return res


def pmc_summary(fetch_dir, write_dir, out, ncol, fast_math):
    """HBM bytes per launch of the ascent kernels from two rocprofv3 counter
    runs of --profile-pass (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch),
    each scaled by what the same counter reads for the pass's 1 GiB
    device-to-device copy (HIP's copyBuffer kernel).  Writes the JSON summary that the timed runs report as
    roofline.traffic."""
    import csv
    import glob

    def per_dispatch(d, counter):
        rows = []
        for f in glob.glob(os.path.join(d, '**', '*counter_collection.csv'), recursive=True):
            with open(f) as fh:
                rows += [r for r in csv.DictReader(fh) if r['Counter_Name'] == counter]
        acc = {}
        for r in rows:
            k = int(r['Dispatch_Id'])
            name, v = r['Kernel_Name'], float(r['Counter_Value'])
            acc[k] = (name, acc.get(k, (name, 0.0))[1] + v)
        return acc

    def first(disp, pred):
        """Counter value of the first dispatch whose kernel name matches."""
        for k in sorted(disp):
            if pred(disp[k][0]):
                return disp[k][1]
        return None

    # the pass runs the plain CAPE-only call, then the warm one (setup pass +
    # persistent ascent), then a combined step: the first dispatch of each
    kernels = {
        'plain': lambda n: 'cape_grid_kernel<float' in n,
        'warm_setup': lambda n: 'warm_setup_kernel<true, true, true, false>' in n,
        'warm_ascent': lambda n: 'warm_ascent_kernel<true, true, true, false, false>' in n,
    }
    res = {'ncol': ncol, 'fast_math': fast_math,
           'kernels': {'plain': 'cape_grid_kernel<float, unsigned int, true, true, false, false>',
                       'warm': 'warm_setup_kernel<true, true, true, false> + '
                               'warm_ascent_kernel<true, true, true, false, false>'}}
    gib = float(1 << 30)
    for label, d, counter in (('fetch', fetch_dir, 'FETCH_SIZE'), ('write', write_dir, 'WRITE_SIZE')):
        disp = per_dispatch(d, counter)
        copies = [v for n, v in disp.values() if 'copyBuffer' in n]
        cal = max(copies) * 1024.0 if copies else None
        factor = gib / cal if cal else 1.0
        res[f'{label}_calibration'] = {'copy_bytes': gib, 'counter_bytes': cal, 'factor': factor}
        for mode, pred in kernels.items():
            v = first(disp, pred)
            if v is not None:
                res[f'{mode}_{label}_bytes'] = v * 1024.0 * factor
        parts = [f'warm_{x}_{label}_bytes' for x in ('setup', 'ascent')]
        if all(x in res for x in parts):
            res[f'warm_{label}_bytes'] = sum(res[x] for x in parts)
    for mode in ('plain', 'warm'):
        if f'{mode}_fetch_bytes' in res and f'{mode}_write_bytes' in res:
            res[f'{mode}_hbm_bytes_per_launch'] = (res[f'{mode}_fetch_bytes'] +
                                                    res[f'{mode}_write_bytes'])
    with open(out, 'w') as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res))


def ragged_section(fast_math, log, ncol=200_000):
    """The ragged-batch workload (200 000 columns, level counts 30/40/60/91/137,
    SB/ML/MU sources, pseudo-liquid/pseudo-ice): autotuned persistent plans,
    graph-replayed, plain and warm-started; accuracy on a column sample."""
    import torch

    from oracle import oracle as O
    from xcape_amd import ragged, synthetic
    t0 = time.perf_counter()
    c = synthetic.ragged_columns(ncol, 1002)
    b = ragged.RaggedBatch.from_dict(c)
    res = {'columns': ncol, 'generate_s': time.perf_counter() - t0}
    for name, warm in (('plain', False), ('warm', True)):
        best, times = ragged.autotune(b, fast_math=fast_math, use_cache=False, warm_start=warm)
        plan = ragged.RaggedPlan(b, fast_math=fast_math, warm_start=warm, **best)
        ms = ragged._time_plan(plan, reps=10)
        st = plan.run(stats=True)
        plan.close()
        torch.cuda.synchronize()
        res[name] = {'best': best, 'ms': ms, 'columns_per_s': ncol / (ms / 1e3),
                     'configs_swept': len(times), 'iterations': st['iterations'],
                     'fallbacks': st['fallbacks']}
        log(f'[bench] ragged {name}: {best} {ms:.3f} ms')
    # accuracy of the warm plan on the first 2000 columns (grouped oracle)
    sel = np.arange(2000)
    got = b.cape.cpu().numpy()[sel]
    off = c['offsets']
    err = 0.0
    for i in sel:
        s0, s1 = off[i], off[i + 1]
        r = O.cape(c['p'][None, s0:s1], c['t'][None, s0:s1], c['td'][None, s0:s1],
                   c['ps'][i:i + 1], c['ts'][i:i + 1], c['tds'][i:i + 1], 500.,
                   int(c['source'][i]), 500., int(c['adiabat'][i]))
        err = max(err, abs(float(got[i]) - float(r[0][0])))
    res['warm_cape_max_abs_err_2000'] = err
    res['speedup_warm_vs_plain'] = res['plain']['ms'] / res['warm']['ms']
    return res


def host_array_section(a, ncol, args, log):
    """The PCIe-inclusive rate: the combined call on host (numpy) copies of
    the device inputs `a`, as the reference's callers hold them, over the
    whole grid.  The library streams them through pinned staging; uploads and
    downloads are inside the timed region.  Its roofline is the host-to-device
    copy rate from pinned memory, measured here on a 512 MiB copy."""
    import torch

    from xcape_amd import _hip

    log('[bench] host-array path')
    host = [x.cpu().numpy() for x in a]
    torch.cuda.synchronize()
    kw_h = dict(pinc=PINC, source=SOURCE, adiabat=ADIABAT, fast_math=args.fast_math,
                depth=DEPTH, warm_start=not args.plain)
    _hip.cape_srh(*host, **kw_h)
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        _hip.cape_srh(*host, **kw_h)
    host_ms = (time.perf_counter() - t0) / reps * 1e3
    in_bytes = int(sum(x.nbytes for x in host))
    del host
    pin = torch.empty(1 << 27, dtype=torch.float32, pin_memory=True)
    dst = torch.empty(1 << 27, dtype=torch.float32, device='cuda')
    dst.copy_(pin, non_blocking=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        dst.copy_(pin, non_blocking=True)
    torch.cuda.synchronize()
    h2d_gbs = 5 * pin.numel() * 4 / (time.perf_counter() - t0) / 1e9
    del pin, dst
    host_gbs = in_bytes / (host_ms / 1e3) / 1e9
    host_path = {'columns': ncol, 'ms_per_call': host_ms, 'columns_per_s': ncol / (host_ms / 1e3),
                 'input_bytes': in_bytes, 'input_gbs': host_gbs,
                 'pinned_h2d_gbs': h2d_gbs, 'frac_of_pinned_h2d': host_gbs / h2d_gbs}
    return host_path


def load_surrogate(path, log):
    """Activate the committed surrogate, or train one on the GPU when the
    file is absent.  Returns a description for the JSON line."""
    from xcape_amd import surrogate
    if path and os.path.exists(path):
        s = surrogate.Surrogate.load(path)
        src = os.path.relpath(path, ROOT)
    else:
        log('[bench] no surrogate file: training one on the GPU (2 M columns)')
        s = surrogate.train(2_000_000, 200_000, epochs=6, log=None)
        src = 'trained in this run'
    s.activate()
    return {'source': src, 'val_guess_error_K': s.meta.get('val_guess_error_K'),
            'train_samples': s.meta.get('train_samples')}


def timed(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        out = fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return out, e0.elapsed_time(e1) / steps, time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--fast-math', action='store_true')
    ap.add_argument('--plain', action='store_true',
                    help='report the cold-started reference method as the value')
    ap.add_argument('--surrogate', default=os.path.join(ROOT, 'xcape_amd', 'data',
                                                        'surrogate_v1.npz'))
    ap.add_argument('--cpu-sample', type=int, default=1 << 20,
                    help='columns copied to the host for the CPU baseline (upper bound)')
    ap.add_argument('--cpu-seconds', type=float, default=10.0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-ragged', action='store_true',
                    help='skip the ragged-batch (autotuned persistent kernel) section')
    ap.add_argument('--ncol', type=int, default=NCOL_X * NCOL_Y)
    ap.add_argument('--profile-pass', action='store_true',
                    help='one launch of each kernel variant, for counter runs')
    ap.add_argument('--pmc-summary', nargs=2, metavar=('FETCH_DIR', 'WRITE_DIR'),
                    help='summarise two rocprofv3 counter runs of --profile-pass')
    ap.add_argument('--pmc-out', default=PMC_FILE)
    ap.add_argument('--host-only', action='store_true',
                    help='only the host-array (PCIe-inclusive) section, for traces')
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary, args.pmc_out, args.ncol, bool(args.fast_math))
        return
    if args.gpus != 1:
        print(f'bench.py runs on one GPU; --gpus {args.gpus} is not supported '
              f'(columns are independent: a multi-GPU run would shard them with no exchange)',
              file=sys.stderr)
        sys.exit(2)

    import torch

    from xcape_amd import _hip, synthetic

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    torch.cuda.set_device(0)
    ncol = args.ncol
    if args.profile_pass:
        # the counters' calibration: a 1 GiB device-to-device copy (HIP's
        # copyBuffer blit kernel; the summary takes the largest such dispatch)
        x = torch.ones(1 << 28, dtype=torch.float32, device='cuda')
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y
    log(f'[bench] generating {ncol} columns x {NLEV} levels on the GPU')
    inp = synthetic.sigma_columns_torch(ncol, NLEV, SEED, device='cuda', winds=True)
    torch.cuda.synchronize()
    keys = ('p', 't', 'td', 'u', 'v', 'ps', 'ts', 'tds', 'us', 'vs')
    a = [inp[k] for k in keys]
    cape_in = (a[0], a[1], a[2], a[5], a[6], a[7])
    kw = dict(pinc=PINC, source=SOURCE, adiabat=ADIABAT, fast_math=args.fast_math)
    stream = torch.cuda.current_stream()
    sur = load_surrogate(args.surrogate, log)
    if args.host_only:
        print(json.dumps({'host_arrays_pcie_inclusive': host_array_section(a, ncol, args, log),
                          'surrogate': sur}))
        return

    def step(warm, stats=False):
        return _hip.cape_srh(*a, depth=DEPTH, warm_start=warm, stats=stats, **kw)

    # the CAPE-only launches (the roofline's dominant kernel) read level-major
    # copies made here, outside every timed region: the HIP events around them
    # then bracket the ascent kernel alone, without the staging transpose that
    # column-contiguous inputs go through
    cape_lm = tuple(x.t().contiguous() if x.dim() == 2 else x for x in cape_in)

    def cape_only(warm, stats=False):
        return _hip.cape(*cape_lm, warm_start=warm, stats=stats, level_major=True, **kw)

    if args.profile_pass:
        for warm in (False, True):
            cape_only(warm)
            torch.cuda.synchronize()
        step(True)
        torch.cuda.synchronize()
        log('[bench] profile pass done')
        return

    # counting passes: level / sub-step / iteration totals of each method
    log('[bench] counting passes')
    counts = {m: cape_only(w, stats=True)[-1]
              for m, w in (('plain', False), ('cold_loop', 'cold'),
                           ('extrapolation', 'extrapolation'), ('warm', True))}
    mu_levels = int((inp['p'] >= 500.0).sum().item()) + ncol   # levels scanned at p >= 500 hPa
    ops = algorithmic_ops(counts['plain'], NLEV, mu_levels)

    log('[bench] timing the CAPE-only launches')
    ms_cape = {}
    for m, w in (('plain', False), ('cold_loop', 'cold'), ('extrapolation', 'extrapolation'),
                 ('warm', True)):
        _, ms_cape[m], _ = timed(lambda: cape_only(w), args.steps, args.warmup, stream)
    # the helicity chain alone, on column-contiguous input as in the combined
    # step (its staging transpose included)
    _, ms_srh, _ = timed(lambda: _hip.srh(*a, depth=DEPTH), args.steps, args.warmup, stream)

    log('[bench] timing the combined step')
    outs_p, ms_plain, _ = timed(lambda: step(False), args.steps, args.warmup, stream)
    warm = not args.plain
    if warm:
        outs, ms, wall = timed(lambda: step(True), args.steps, args.warmup, stream)
    else:
        outs, ms, wall = outs_p, ms_plain, None
    value = ncol / (ms / 1e3)

    log('[bench] accuracy')
    sample = 2048
    sl = [x[:sample] for x in cape_in]
    fp64 = _hip.cape(*sl, pinc=PINC, source=SOURCE, adiabat=ADIABAT, fp64=True)
    acc = accuracy(outs, inp, sample, fp64)
    acc_plain = accuracy(outs_p, inp, sample, fp64)
    # every column of the reported method against the reference method in the
    # accurate math mode, on the GPU (the oracle is too slow for the full grid)
    ref_full = _hip.cape(*cape_lm, level_major=True, pinc=PINC, source=SOURCE, adiabat=ADIABAT)
    acc['full_grid_vs_plain_accurate_gpu'] = full_grid_agreement(outs, ref_full)
    del ref_full

    # the dominant kernel: the ascent (CAPE-only launch of the reported method)
    dom = 'warm' if warm else 'plain'
    achieved = ops / (ms_cape[dom] / 1e3) / 1e12
    ops_exec = algorithmic_ops(counts[dom], NLEV, mu_levels)
    traffic = None
    if os.path.exists(PMC_FILE):
        with open(PMC_FILE) as f:
            pmc = json.load(f)
        if pmc.get('ncol') == ncol and pmc.get('fast_math') == bool(args.fast_math):
            traffic = pmc.get(f'{dom}_hbm_bytes_per_launch')

    rag = None
    if not args.no_ragged:
        log('[bench] ragged batches')
        rag = ragged_section(bool(args.fast_math), log)

    host_path = host_array_section(a, ncol, args, log)

    cpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        nhost = min(ncol, args.cpu_sample)
        host = {k: v[:nhost].cpu().numpy() for k, v in inp.items()}
        log(f'[bench] CPU baseline on {threads} threads, {args.cpu_seconds} s')
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
                               'motion (xcape_cape_srh_f32)',
                   'ncol': ncol, 'nlev': NLEV, 'warm_start': warm,
                   'fast_math': bool(args.fast_math), 'parallelism': 'single GPU'},
        'roofline': {'bound': 'valu', 'pipe': 'VALU FP32 (the ascent has no matrix work)',
                     'kernel': ('warm_setup_kernel + warm_ascent_kernel (the warm-started '
                                'CAPE-only call)' if warm else
                                'cape_grid_kernel (the plain CAPE-only call)'),
                     'achieved': achieved, 'peak': FP32_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': achieved / FP32_PEAK_TFLOPS, 'traffic': traffic,
                     'ops_per_launch': ops, 'ms_per_launch': ms_cape[dom],
                     # the same unit weights over the work the kernel executed
                     # (its own iteration counts rather than the plain method's)
                     'ops_executed_per_launch': ops_exec,
                     'achieved_executed': ops_exec / (ms_cape[dom] / 1e3) / 1e12,
                     'hbm_gbs_algorithmic': bytes_per_column(NLEV) * value / 1e9},
        'plain_kernel': {'value': ncol / (ms_plain / 1e3), 'ms_per_step': ms_plain,
                         'accuracy': acc_plain},
        'speedup_vs_plain': ms_plain / ms,
        'attribution_cape_only_ms': ms_cape,
        'srh_only_ms': ms_srh,
        'kernel_counts': counts,
        'surrogate': sur,
        'accuracy': acc,
        'ragged_200k': rag,
        'host_arrays_pcie_inclusive': host_path,
        'cpu_baseline': cpu,
    }
    print(json.dumps(line))


if __name__ == '__main__':
    main()
