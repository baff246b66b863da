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
  accuracy      CAPE/CIN/MUlev/zMUlev of a column sample against the FP32 oracle;
                every output (CAPE, CIN, MUlev, zMUlev, LFC, EL, SRH, Bunkers)
                against the FP64 re-run of the GPU kernels (fp64_rerun); the
                CAPE outputs of every column the CPU baseline finished against
                the reference's (vs_reference_cpu_sample)
  configs       the 64x64x40 sigma grid (surface, pseudo-liquid) and the
                1024x512x37 pressure-level grid (surface, mixed-layer,
                most-unstable): plain and warm-started calls, their rooflines,
                the reference on the host cores and the outputs against it
  ragged_200k   the ragged batch: both plans autotuned (every config timed is
                listed), rooflines, the reference grouped by (level count,
                source, adiabat) and the outputs against it
  cpu_baseline  the reference Fortran (compiled from source, oracle/_ref) on the
                host cores over a bounded sample of the same workload, or the C
                restatement of the oracle when that build is absent

--profile-pass runs each kernel variant once (plus a 1 GiB device copy for
counter calibration) and prints nothing else, for rocprofv3 counter runs;
--profile-reps N repeats each call N times back to back, for kernel traces.
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


def algorithmic_ops(stats, nk, mu_levels, w_it=None, env_levels=None):
    """Unit-weighted FP32 operations of the reference method's work counted
    in `stats`: the environment of env_levels levels (default (nk + 1) per
    column), the most-unstable scan over mu_levels, and the ascent's levels,
    sub-steps and iterations (w_it per iteration; default that of ADIABAT)."""
    if w_it is None:
        w_it = W_ITER_ICE if ADIABAT in (3, 4) else W_ITER_LIQ
    if env_levels is None:
        env_levels = (nk + 1) * stats['columns']
    return (W_ENV * env_levels + W_MU * mu_levels + W_LEVEL * stats['levels']
            + W_SUBSTEP * stats['substeps'] + w_it * stats['iterations'])


def bytes_per_column(nk):
    # compulsory HBM bytes: p, t, td, u, v profiles + 5 surface values in,
    # cape, cin, mulev, zmulev (4 B) + srh_rm, srh_lm (8 B) + rm, lm, mean (8 B)
    return 4 * (5 * nk + 5) + 4 * 4 + 2 * 8 + 3 * 8


def reference_kind():
    from oracle import ref_driver as R
    return 'reference' if R.available() else 'port'


def run_tasks(tasks, threads, seconds, fn):
    """fn(columns) on `threads` host threads over `tasks` (arrays of column
    indices), in order, until `seconds` of wall time have passed (ctypes
    releases the GIL).  Returns (columns done, wall seconds, {task number:
    (columns, fn's result)})."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    nxt = [0]
    lock = threading.Lock()
    res = {}
    deadline = time.perf_counter() + seconds

    def worker(_):
        done = 0
        while time.perf_counter() < deadline:
            with lock:
                i = nxt[0]
                nxt[0] += 1
            if i >= len(tasks):
                break
            res[i] = (tasks[i], fn(tasks[i]))
            done += len(tasks[i])
        return done

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        n = sum(ex.map(worker, range(threads)))
    return n, time.perf_counter() - t0, res


def ranges(ncol, chunk):
    return [np.arange(i0, min(i0 + chunk, ncol)) for i0 in range(0, ncol, chunk)]


def gathered(res):
    """The columns and CAPE outputs of run_tasks' results, by column."""
    keys = sorted(res)
    if not keys:
        return np.zeros(0, np.int64), [np.zeros(0)] * 4
    idx = np.concatenate([res[i][0] for i in keys])
    outs = [np.concatenate([np.asarray(res[i][1][j]) for i in keys]) for j in range(4)]
    order = np.argsort(idx, kind='stable')
    return idx[order], [o[order] for o in outs]


def cpu_baseline(host, threads, seconds):
    """Reference Fortran (or the C oracle) on `threads` host threads over
    chunks of the same input columns, until `seconds` of wall time have
    passed; columns/s of the chunks completed.  Also returns the CAPE
    outputs of the columns done (the reference side of
    accuracy.vs_reference_cpu_sample; Bunkers runs in its timing build,
    whose outputs are not kept)."""
    from oracle import oracle as O
    from oracle import ref_driver as R

    kind = reference_kind()

    def run_chunk(cols):
        a = {k: v[cols[0]:cols[-1] + 1] for k, v in host.items()}
        if kind == 'reference':
            out = R.loopcape_ml(a['p'].T, a['t'].T, a['td'].T, a['ps'], a['ts'], a['tds'], PINC,
                                SOURCE, 500., ADIABAT)
            h, hs = R.stdheight_ml(a['p'].T, a['t'].T, a['td'].T, a['ps'], a['ts'], a['tds'], 2.)
            rm, lm, mw = R.bunkers_ml_timing(a['u'].T, a['v'].T, h.astype(np.float32), a['us'],
                                             a['vs'], hs.astype(np.float32))
            R.sreh_ml(a['u'].T, a['v'].T, h, a['us'], a['vs'], hs, rm[0], rm[1], lm[0], lm[1],
                      DEPTH)
        else:
            out = O.cape(a['p'], a['t'], a['td'], a['ps'], a['ts'], a['tds'], PINC, SOURCE, 500.,
                         ADIABAT)
            O.srh(a['p'], a['t'], a['td'], a['u'], a['v'], a['ps'], a['ts'], a['tds'], a['us'],
                  a['vs'], DEPTH)
        return out

    n, dt, res = run_tasks(ranges(host['t'].shape[0], 1024), threads, seconds, run_chunk)
    src = 'reference Fortran compiled from source' if kind == 'reference' else 'C restatement'
    return {'value': n / dt, 'unit': 'columns/s', 'cores': threads, 'kind': kind,
            'sample': f'{n} columns of the same inputs (137 sigma levels, MU, pseudo-ice, '
                      f'pinc 500, + stdheight/Bunkers/SREH 0-3 km), {src}, {dt:.1f} s wall '
                      f'on {threads} threads'}, res


def _within(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return (np.abs(a - b) <= np.maximum(0.5, 1e-3 * np.abs(b))) | (np.isnan(a) & np.isnan(b))


PERTURB_SEEDS = range(1, 17)
ITERATE_AMP = 1e-4     # K: a converged sub-step's other admissible iterates


def vs_reference(got, ref, oracle_fn=None, level_fn=None):
    """A GPU run's (cape, cin, mulev, zmulev) against the reference's on the
    same columns (host arrays, position j of each is column cols[j] of the
    configuration): columns outside max(0.5 J/kg, 0.1 %) for CAPE and for
    CIN; of those, the ones whose reference CAPE is below 0.05 J/kg and the
    ones that one run reports as CAPE = CIN = 0 and the other not; MUlev
    equality and the zMUlev error.

    With `oracle_fn(j, fp64=False, perturb=None, perturb_iterate=0)` (the
    oracle on positions j: its FP64 run, its FP32 run with every exp, log
    and pow moved by up to an ulp, and that run also accepting another point
    of each sub-step's convergence band), each column outside the tolerance
    is re-examined: it is ill-conditioned when the FP64 run or one of 32
    perturbed runs lands outside the tolerance of the reference's own FP32
    value (the reference's result then turns on the last bit of its C
    library, or on which iterate of its tolerance band it accepts), and the
    GPU's value is accounted for when it is within the tolerance of one of
    those runs.  With `level_fn(j, k)` ((p [hPa], t, td [degC]) of column j's
    1-based level k, 1 the surface), an unequal MUlev is a reported tie when
    the two levels' theta-e (the oracle's getthe) agree to 4 ulp."""
    cape, cin, mul, z = (np.asarray(x) for x in got)
    rc, rci, rm, rz = (np.asarray(x) for x in ref)
    res = {'columns': int(len(rc))}
    sliver = rc < 0.05
    stop_g = (cape == 0) & (cin == 0)
    stop_r = (rc == 0) & (rci == 0)
    outs = {}
    for name, g, r in (('cape', cape, rc), ('cin', cin, rci)):
        d = np.abs(g.astype(np.float64) - r)
        out = ~_within(g, r)
        outs[name] = out
        res[name] = {'outside_tol': int(out.sum()),
                     'outside_tol_sliver': int((out & sliver).sum()),
                     'outside_tol_stop_flip': int((out & (stop_g != stop_r)).sum()),
                     'max_abs_err': float(np.nanmax(np.where(np.isnan(d), 0.0, d)))
                     if len(d) else 0.0}
    bad = np.nonzero(outs['cape'] | outs['cin'])[0]
    if oracle_fn is not None and len(bad):
        alts = [oracle_fn(bad, fp64=True)] + [oracle_fn(bad, perturb=s) for s in PERTURB_SEEDS]
        alts += [oracle_fn(bad, perturb=s, perturb_iterate=ITERATE_AMP) for s in PERTURB_SEEDS]
        ill = np.zeros(len(bad), bool)
        seen = np.zeros(len(bad), bool)
        for o in alts:
            ill |= ~(_within(o[0], rc[bad]) & _within(o[1], rci[bad]))
            seen |= _within(cape[bad], o[0]) & _within(cin[bad], o[1])
        res['outside_tol_columns'] = {
            'count': int(len(bad)),
            'ill_conditioned': int(ill.sum()),
            'gpu_within_tol_of_a_perturbed_or_fp64_oracle_run': int(seen.sum()),
            'neither': int((~ill & ~seen).sum()),
            'note': 'ill-conditioned: the oracle in FP64, or with exp/log/pow moved by up to '
                    'an ulp (16 seeds), or also accepting another iterate of each sub-step\'s '
                    'convergence band (16 seeds), lands outside the tolerance of the '
                    'reference itself'}
    res['zero_cape_and_cin'] = {'gpu': int(stop_g.sum()), 'reference': int(stop_r.sum()),
                                'flips': int((stop_g != stop_r).sum())}
    neq = np.nonzero(mul != rm)[0]
    res['mulev'] = {'frac_equal': float(np.mean(mul == rm)) if len(rm) else 1.0,
                    'unequal': int(len(neq))}
    if level_fn is not None and len(neq):
        res['mulev']['unequal_theta_e_ties'] = int(sum(
            _theta_e_tie(level_fn(j, int(mul[j])), level_fn(j, int(rm[j]))) for j in neq))
    dz = np.abs(z.astype(np.float64) - rz)
    res['zmulev_max_abs_err'] = float(np.nanmax(dz)) if len(dz) else 0.0
    return res


def _theta_e_tie(la, lb):
    """The oracle's theta-e (getthe with q = getqvs(p, td), as the
    most-unstable search evaluates it) of two levels, equal to 4 ulp."""
    import ctypes

    from oracle import oracle as O
    lib = O.lib()
    q, f = lib.xo_getqvs_f32, lib.xo_getthe_f32
    q.restype, q.argtypes = ctypes.c_float, [ctypes.c_float] * 2
    f.restype, f.argtypes = ctypes.c_float, [ctypes.c_float] * 4

    def the(lv):
        p = np.float32(np.float32(100.0) * np.float32(lv[0]))
        t = np.float32(np.float32(273.15) + np.float32(lv[1]))
        td = np.float32(np.float32(273.15) + np.float32(lv[2]))
        return np.float32(f(p, t, td, q(p, td)))
    a, b = the(la), the(lb)
    return bool(abs(a - b) <= 4 * np.spacing(max(a, b)))


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
    CIN, MUlev, zMUlev, SRH, Bunkers), against the oracle's FP64 run (CAPE,
    CIN) and against the FP64 re-run of the GPU kernel (CAPE, CIN, MUlev,
    zMUlev)."""
    from oracle import oracle as O
    a = {k: v[:sample].cpu().numpy() for k, v in inp.items()}
    o = O.cape(a['p'], a['t'], a['td'], a['ps'], a['ts'], a['tds'], PINC, SOURCE, 500., ADIABAT)
    o64 = O.cape(*(a[k].astype(np.float64) for k in ('p', 't', 'td', 'ps', 'ts', 'tds')), PINC,
                 SOURCE, 500., ADIABAT, fp64=True)
    g = [x[:sample].cpu().numpy() for x in outs]
    f64 = [x[:sample].cpu().numpy() for x in fp64_outs]
    res = {}
    for i, name in enumerate(('cape', 'cin')):
        d = np.abs(g[i].astype(np.float64) - o[i])
        tol = np.maximum(0.5, 1e-3 * np.abs(o[i].astype(np.float64)))
        d64 = np.abs(g[i].astype(np.float64) - o64[i])
        res[name] = {'max_abs_err': float(d.max()), 'frac_within_tol': float(np.mean(d <= tol)),
                     'max_abs_err_vs_fp64_gpu': float(np.abs(g[i] - f64[i]).max()),
                     'max_abs_err_vs_fp64_oracle': float(d64.max()),
                     'frac_within_tol_vs_fp64_oracle':
                         float(np.mean(d64 <= np.maximum(0.5, 1e-3 * np.abs(o64[i]))))}
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


def fp64_rerun(outs, inp, sample, warm, fast_math):
    """Every output of the reported method on the first `sample` columns
    against the FP64 re-run of the GPU kernels on the same FP32 inputs:
    CAPE, CIN, MUlev, zMUlev (xcape_cape_f64), LFC and EL
    (xcape_cape_levels_f64 against xcape_cape_levels_f32 with the same
    method), SRH and Bunkers motion (xcape_srh_f64)."""
    from xcape_amd import _hip
    s = {k: v[:sample] for k, v in inp.items()}
    ca = [s[k] for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
    kw = dict(pinc=PINC, source=SOURCE, adiabat=ADIABAT)
    f64 = [x.cpu().numpy() for x in _hip.cape(*ca, fp64=True, levels=True, **kw)]
    f32 = [x.cpu().numpy() for x in _hip.cape(*ca, warm_start=warm, fast_math=fast_math,
                                                levels=True, **kw)]
    srh64 = [x.cpu().numpy() for x in _hip.srh(*(s[k] for k in ('p', 't', 'td', 'u', 'v', 'ps',
                                                                 'ts', 'tds', 'us', 'vs')),
                                                 depth=DEPTH, fp64=True)]
    g = [x[:sample].cpu().numpy() for x in outs]
    res = {'columns': sample}
    for i, name in enumerate(('cape', 'cin')):
        d = np.abs(g[i].astype(np.float64) - f64[i])
        tol = np.maximum(0.5, 1e-3 * np.abs(f64[i]))
        res[name] = {'max_abs_err': float(np.nanmax(d)), 'frac_within_tol': float(np.mean(d <= tol))}
    res['mulev'] = {'frac_equal': float(np.mean(g[2] == f64[2]))}
    res['zmulev'] = {'max_abs_err': float(np.nanmax(np.abs(g[3] - f64[3])))}
    for i, name in ((4, 'lfc_hpa'), (5, 'el_hpa')):
        fin = np.isfinite(f32[i]) & np.isfinite(f64[i])
        d = np.abs(f32[i][fin].astype(np.float64) - f64[i][fin])
        res[name] = {'max_abs_err': float(d.max()) if d.size else 0.0,
                     'frac_within_1hpa': float(np.mean(d <= 1.0)) if d.size else 1.0,
                     'presence_mismatch': int((np.isfinite(f32[i]) != np.isfinite(f64[i])).sum())}
    res['srh'] = {'max_abs_err': float(max(np.abs(g[4] - srh64[0]).max(),
                                           np.abs(g[5] - srh64[1]).max()))}
    res['bunkers'] = {'max_abs_err': float(max(np.abs(g[j] - srh64[j - 4]).max()
                                               for j in (6, 7, 8)))}
    return res


def roofline_entry(ops, ms, bytes_col, ncol):
    """VALU (unit-weighted FP32 ops) and HBM (compulsory bytes) fractions of
    a call taking `ms`."""
    achieved = ops / (ms / 1e3) / 1e12
    gbs = bytes_col * ncol / (ms / 1e3) / 1e9
    return {'bound': 'valu', 'achieved': achieved, 'peak': FP32_PEAK_TFLOPS, 'unit': 'TFLOP/s',
            'frac': achieved / FP32_PEAK_TFLOPS, 'ops_per_launch': ops,
            'hbm_gbs_algorithmic': gbs, 'hbm_frac': gbs / HBM_PEAK_GBS,
            'bytes_per_column': bytes_col}


def sigma_oracle(host, source, adiabat):
    """The oracle and the level lookup on columns of a sigma grid (host
    arrays p, t, td (ncol, nk) and ps, ts, tds), for vs_reference."""
    from oracle import oracle as O

    def ofn(cols, **kw):
        a = [host[k][cols] for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
        return O.cape(*a, PINC, source, 500., adiabat, **kw)

    def lfn(col, k):
        if k == 1:
            return host['ps'][col], host['ts'][col], host['tds'][col]
        return host['p'][col, k - 2], host['t'][col, k - 2], host['td'][col, k - 2]
    return ofn, lfn


def at(fn, idx):
    """fn on configuration columns, called with positions into idx."""
    return lambda j, *a, **k: fn(idx[j], *a, **k)


def cape_config(name, dev, kw, p_shared, env_levels, mu_levels, w_it, bytes_col, ref_fn,
                threads, seconds, steps, warmup, log, ofn=None, lfn=None):
    """One grid configuration: the plain and warm-started CAPE calls on
    column-contiguous device arrays `dev` (the whole call timed with HIP
    events: staging transpose and kernels), their work counters and
    roofline, the reference on the host cores over chunks of the same
    columns (`ref_fn(columns)`, bounded by `seconds`), and each method's
    outputs against the reference's on the columns it finished."""
    import torch

    from xcape_amd import _hip
    stream = torch.cuda.current_stream()
    ncol = dev[1].shape[0]
    res = {'workload': name, 'columns': ncol}
    outs = {}
    for m, warm in (('plain', False), ('warm', True)):
        st = _hip.cape(*dev, warm_start=warm, stats=True, p_shared=p_shared, **kw)[-1]
        out, ms, _ = timed(lambda: _hip.cape(*dev, warm_start=warm, p_shared=p_shared, **kw),
                           steps, warmup, stream)
        outs[m] = [x.cpu().numpy() for x in out]
        res[m] = {'ms': ms, 'columns_per_s': ncol / (ms / 1e3), 'iterations': st['iterations'],
                  'substeps': st['substeps'], 'fallbacks': st['fallbacks'],
                  'nonconverged': st['nonconverged']}
        if m == 'plain':
            ops = algorithmic_ops(st, None, mu_levels, w_it, env_levels)
        res[m]['roofline'] = roofline_entry(ops, ms, bytes_col, ncol)
        log(f'[bench] {name.split(":")[0]} {m}: {ms:.3f} ms')
    kind = reference_kind()
    n, dt, rres = run_tasks(ranges(ncol, 1024), threads, seconds, ref_fn)
    idx, ref = gathered(rres)
    res['cpu_baseline'] = {'value': n / dt, 'unit': 'columns/s', 'cores': threads, 'kind': kind,
                           'sample': f'{n} of the {ncol} columns, {dt:.2f} s wall on '
                                     f'{threads} threads'}
    for m in ('plain', 'warm'):
        res[m]['vs_reference'] = vs_reference(
            [o[idx] for o in outs[m]], ref, ofn and at(ofn, idx),
            lfn and (lambda j, k: lfn(int(idx[j]), k)))
        res[m]['speedup_vs_cpu'] = res[m]['columns_per_s'] / res['cpu_baseline']['value']
    return res


def c0_section(args, threads, log):
    """C0: 64 x 64 columns x 40 sigma levels, surface parcel, pseudo-liquid,
    pinc 500 Pa (the reference's CPU-runnable case), whole grid."""
    import torch

    from oracle import oracle as O
    from oracle import ref_driver as R
    from xcape_amd import synthetic
    ncol, nk = 64 * 64, 40
    c = synthetic.sigma_columns(ncol, nk, 1000)
    keys = ('p', 't', 'td', 'ps', 'ts', 'tds')
    dev = [torch.from_numpy(c[k]).cuda() for k in keys]
    kw = dict(pinc=PINC, source=1, ml_depth=500., adiabat=1, fast_math=args.fast_math)
    ref = R.available()

    def ref_fn(cols):
        a = [c[k][cols] for k in keys]
        if ref:
            return R.loopcape_ml(a[0].T, a[1].T, a[2].T, a[3], a[4], a[5], PINC, 1, 500., 1)
        return O.cape(*a, PINC, 1, 500., 1)
    ofn, lfn = sigma_oracle(c, 1, 1)
    return cape_config('C0: 64x64 columns x 40 sigma levels, surface, pseudo-liquid, pinc 500 Pa',
                       dev, kw, False, ncol * (nk + 1), 0, W_ITER_LIQ, 3 * nk * 4 + 12 + 16,
                       ref_fn, threads, 5.0, 50, 5, log, ofn, lfn)


def c1_sections(args, threads, log):
    """C1: 1024 x 512 columns x 37 shared pressure levels (1000-100 hPa),
    levels below ground masked by each column's start level, pseudo-liquid,
    pinc 500 Pa; surface, mixed-layer (500 m) and most-unstable runs, the
    reference's loopcape_pl1d (start levels from the reference's rule,
    core.py:286-289, inside its timing)."""
    import torch

    from oracle import oracle as O
    from oracle import ref_driver as R
    from xcape_amd import core, synthetic
    ncol = 1024 * 512
    t0 = time.perf_counter()
    c = synthetic.pressure_columns(ncol, 1001)
    log(f'[bench] C1 inputs {time.perf_counter() - t0:.1f} s')
    p1d = c['p'].astype(np.float32)
    nk = len(p1d)
    dev = [torch.from_numpy(p1d).cuda()] + [torch.from_numpy(c[k]).cuda()
                                             for k in ('t', 'td', 'ps', 'ts', 'tds')]
    start = core.pressure_start_levels(p1d.astype(np.float64), c['ps'].astype(np.float64))
    env_levels = int((nk - start + 2).sum())
    n500 = int((p1d >= 500.0).sum())
    mu_levels = int(np.maximum(0, n500 - (start - 1)).sum()) + ncol
    ref = R.available()

    def pressure_levels(col, k):
        # 1-based level k of column col counted from the surface (1), the
        # first level above it being the column's start level
        if k == 1:
            return c['ps'][col], c['ts'][col], c['tds'][col]
        g = int(start[col]) - 1 + (k - 2)
        return p1d[g], c['t'][col, g], c['td'][col, g]
    out = {}
    for name, source in (('surface', 1), ('mixed_layer_500m', 3), ('most_unstable', 2)):
        kw = dict(pinc=PINC, source=source, ml_depth=500., adiabat=1, fast_math=args.fast_math)

        def ref_fn(cols, source=source):
            ps = c['ps'][cols]
            st = core.pressure_start_levels(p1d.astype(np.float64), ps.astype(np.float64))
            if ref:
                return R.loopcape_pl1d(c['t'][cols].T, c['td'][cols].T, p1d, ps, c['ts'][cols],
                                       c['tds'][cols], PINC, source, 500., 1, st)
            return O.cape(None, c['t'][cols], c['td'][cols], ps, c['ts'][cols], c['tds'][cols],
                          PINC, source, 500., 1, p1d=p1d, start_lev=st)

        def ofn(cols, source=source, **kw):
            return O.cape(None, c['t'][cols], c['td'][cols], c['ps'][cols], c['ts'][cols],
                          c['tds'][cols], PINC, source, 500., 1, p1d=p1d, start_lev=start[cols],
                          **kw)
        out[name] = cape_config(f'C1: 1024x512 columns x 37 pressure levels, {name}, '
                                f'pseudo-liquid, pinc 500 Pa', dev, kw, True, env_levels,
                                mu_levels if source == 2 else 0, W_ITER_LIQ,
                                2 * nk * 4 + 12 + 16, ref_fn, threads, 4.0, 10, 2, log, ofn,
                                pressure_levels)
    return out


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
    return res


def ragged_section(fast_math, threads, log, ncol=200_000):
    """C2, the ragged-batch workload: 200 000 columns, level counts
    30/40/60/91/137, SB/ML/MU sources, pseudo-liquid/pseudo-ice.  Both plans
    autotuned (plain: chunk x LDS tile depth x columns per workgroup; warm:
    the persistent ascent's columns per workgroup; every config timed is
    reported), graph-replayed, with the roofline of the reference method's
    work and the reference on the host cores, one loopcape_ml call per (level
    count, source, adiabat) group of 1024-column chunks, on a bounded sample,
    whose outputs each plan's are compared with."""
    import torch

    from oracle import oracle as O
    from oracle import ref_driver as R
    from xcape_amd import ragged, synthetic
    t0 = time.perf_counter()
    c = synthetic.ragged_columns(ncol, 1002)
    b = ragged.RaggedBatch.from_dict(c)
    res = {'workload': 'C2: 200000 ragged columns (30/40/60/91/137 levels), SB/ML/MU, '
                       'pseudo-liquid/pseudo-ice, pinc 500 Pa',
           'columns': ncol, 'generate_s': time.perf_counter() - t0}
    off = c['offsets']
    nlev = c['nlev'].astype(np.int64)
    live = c['ts'] > 0
    env_levels = int(((nlev + 1) * live).sum())
    p500 = np.add.reduceat((c['p'] >= 500.0).astype(np.int64), off[:-1])
    mu_levels = int(((p500 + 1) * live * (c['source'] == 2)).sum())
    ice = float(np.mean(c['adiabat'][live] >= 3))
    w_it = ice * W_ITER_ICE + (1 - ice) * W_ITER_LIQ
    bytes_col = float(np.mean(3 * 4 * nlev)) + 12 + 8 + 8 + 16
    outs = {}
    ops = None
    for name, warm in (('plain', False), ('warm', True)):
        best, times = ragged.autotune(b, fast_math=fast_math, use_cache=False, warm_start=warm)
        plan = ragged.RaggedPlan(b, fast_math=fast_math, warm_start=warm, **best)
        ms = ragged._time_plan(plan, reps=10)
        st = plan.run(stats=True)
        plan.run()
        torch.cuda.synchronize()
        outs[name] = [x.cpu().numpy() for x in (b.cape, b.cin, b.mulev, b.zmulev)]
        plan.close()
        if ops is None:
            ops = algorithmic_ops(st, None, mu_levels, w_it, env_levels)
        res[name] = {'best': best, 'ms': ms, 'columns_per_s': ncol / (ms / 1e3),
                     'configs_swept': len(times), 'sweep_ms': times,
                     'iterations': st['iterations'], 'fallbacks': st['fallbacks'],
                     'roofline': roofline_entry(ops, ms, bytes_col, ncol)}
        log(f'[bench] ragged {name}: {best} {ms:.3f} ms ({len(times)} configs)')
    res['roofline_note'] = ('ops: the plain plan\'s counted work with the iteration weight '
                            'averaged over the batch\'s adiabat mix')
    # the reference, grouped as its callers must group a mixed batch
    kind = reference_kind()
    keys = np.stack([nlev, c['source'], c['adiabat']], 1)
    tasks = []
    for key in np.unique(keys, axis=0):
        cols = np.nonzero((keys == key).all(1))[0]
        for j in range(0, len(cols), 1024):
            tasks.append(cols[j:j + 1024])
    # interleave the groups so that a bounded run samples every one
    rng = np.random.default_rng(0)
    tasks = [tasks[i] for i in rng.permutation(len(tasks))]

    def ref_fn(cols):
        nk = int(nlev[cols[0]])
        rows = off[cols][:, None] + np.arange(nk)[None, :]
        src, adi = int(c['source'][cols[0]]), int(c['adiabat'][cols[0]])
        a = (c['p'][rows], c['t'][rows], c['td'][rows], c['ps'][cols], c['ts'][cols],
             c['tds'][cols])
        if kind == 'reference':
            return R.loopcape_ml(a[0].T, a[1].T, a[2].T, a[3], a[4], a[5], PINC, src, 500., adi)
        return O.cape(*a, PINC, src, 500., adi)
    n, dt, rres = run_tasks(tasks, threads, 5.0, ref_fn)
    idx, ref = gathered(rres)

    def ofn(cols, **kw):
        outs1 = []
        for i in cols:
            rows = np.arange(off[i], off[i + 1])
            outs1.append(O.cape(c['p'][None, rows], c['t'][None, rows], c['td'][None, rows],
                                c['ps'][i:i + 1], c['ts'][i:i + 1], c['tds'][i:i + 1], PINC,
                                int(c['source'][i]), 500., int(c['adiabat'][i]), **kw))
        return [np.concatenate([o[j] for o in outs1]) for j in range(4)]

    def lfn(j, k):
        i = int(idx[j])
        if k == 1:
            return c['ps'][i], c['ts'][i], c['tds'][i]
        r = off[i] + k - 2
        return c['p'][r], c['t'][r], c['td'][r]
    res['cpu_baseline'] = {'value': n / dt, 'unit': 'columns/s', 'cores': threads, 'kind': kind,
                           'sample': f'{n} of the {ncol} columns in 1024-column chunks of '
                                     f'(level count, source, adiabat) groups, {dt:.2f} s wall on '
                                     f'{threads} threads'}
    for name in ('plain', 'warm'):
        res[name]['vs_reference'] = vs_reference([o[idx] for o in outs[name]], ref,
                                                 at(ofn, idx), lfn)
        res[name]['speedup_vs_cpu'] = res[name]['columns_per_s'] / res['cpu_baseline']['value']
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


def timed_interleaved(fns, steps, warmup, stream):
    """Mean ms per call of each of `fns` (name -> callable), measured with
    HIP events on `stream` around every call, the calls interleaved one of
    each per round and the order rotated from round to round: the device's
    clock moves over a run (it ramps up, and drops after long heavy
    launches), and interleaving lets that drift fall on every method
    alike."""
    import torch
    names = list(fns)
    for _ in range(warmup):
        for f in fns.values():
            f()
    torch.cuda.synchronize()
    ev = {k: [] for k in fns}
    for r in range(steps):
        for i in range(len(names)):
            k = names[(i + r) % len(names)]
            f = fns[k]
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            f()
            e1.record(stream)
            ev[k].append((e0, e1))
    torch.cuda.synchronize()
    return {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in ev.items()}


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
    ap.add_argument('--no-configs', action='store_true',
                    help='skip the 64x64x40 sigma and 1024x512x37 pressure-level sections')
    ap.add_argument('--ncol', type=int, default=NCOL_X * NCOL_Y)
    ap.add_argument('--profile-pass', action='store_true',
                    help='one launch of each kernel variant, for counter runs')
    ap.add_argument('--profile-reps', type=int, default=1,
                    help='with --profile-pass: back-to-back repeats of each call, '
                         'for kernel traces')
    ap.add_argument('--pmc-summary', nargs=2, metavar=('FETCH_DIR', 'WRITE_DIR'),
                    help='summarise two rocprofv3 counter runs of --profile-pass')
    ap.add_argument('--pmc-out', default=PMC_FILE)
    ap.add_argument('--host-only', action='store_true',
                    help='only the host-array (PCIe-inclusive) section, for traces')
    args = ap.parse_args()
    # the JSON line is the only output on stdout: whatever native code
    # prints (the reference Fortran's convergence messages, which its
    # runtime flushes at exit) goes to stderr with the progress lines
    sys.stdout.flush()
    json_fd = os.dup(1)
    os.dup2(2, 1)

    def emit(obj):
        os.write(json_fd, (json.dumps(obj) + '\n').encode())

    if args.pmc_summary:
        emit(pmc_summary(*args.pmc_summary, args.pmc_out, args.ncol, bool(args.fast_math)))
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
        emit({'host_arrays_pcie_inclusive': host_array_section(a, ncol, args, log),
              'surrogate': sur})
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
            for _ in range(args.profile_reps):
                cape_only(warm)
            torch.cuda.synchronize()
        for _ in range(args.profile_reps):
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
    # the dominant kernel's launch time for the roofline: back-to-back calls
    # of the reported method (as the kernel trace in profiles/ shows them)
    dom = 'warm' if not args.plain else 'plain'
    _, ms_dom, _ = timed(lambda: cape_only(dom == 'warm'), args.steps, args.warmup, stream)
    # the methods against each other: interleaved
    methods = (('plain', False), ('cold_loop', 'cold'), ('extrapolation', 'extrapolation'),
               ('warm', True))
    ms_cape = timed_interleaved({m: (lambda w=w: cape_only(w)) for m, w in methods},
                                args.steps, args.warmup, stream)
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
    acc['fp64_rerun'] = fp64_rerun(outs, inp, sample, warm, bool(args.fast_math))
    acc_plain['fp64_rerun'] = fp64_rerun(outs_p, inp, sample, False, bool(args.fast_math))
    # every column of the reported method against the reference method in the
    # accurate math mode, on the GPU (the oracle is too slow for the full grid)
    ref_full = _hip.cape(*cape_lm, level_major=True, pinc=PINC, source=SOURCE, adiabat=ADIABAT)
    acc['full_grid_vs_plain_accurate_gpu'] = full_grid_agreement(outs, ref_full)
    del ref_full

    # the dominant kernel: the ascent (CAPE-only launch of the reported method)
    achieved = ops / (ms_dom / 1e3) / 1e12
    ops_exec = algorithmic_ops(counts[dom], NLEV, mu_levels)
    traffic = None
    if os.path.exists(PMC_FILE):
        with open(PMC_FILE) as f:
            pmc = json.load(f)
        if pmc.get('ncol') == ncol and pmc.get('fast_math') == bool(args.fast_math):
            traffic = pmc.get(f'{dom}_hbm_bytes_per_launch')

    threads = len(os.sched_getaffinity(0))
    configs = None
    if not args.no_configs:
        log('[bench] 64x64x40 sigma grid')
        configs = {'C0': c0_section(args, threads, log)}
        log('[bench] 1024x512x37 pressure-level grid')
        configs['C1'] = c1_sections(args, threads, log)
    rag = None
    if not args.no_ragged:
        log('[bench] ragged batches')
        rag = ragged_section(bool(args.fast_math), threads, log)

    host_path = host_array_section(a, ncol, args, log)

    cpu = None
    if not args.no_cpu:
        nhost = min(ncol, args.cpu_sample)
        host = {k: v[:nhost].cpu().numpy() for k, v in inp.items()}
        log(f'[bench] CPU baseline on {threads} threads, {args.cpu_seconds} s')
        cpu, ref_res = cpu_baseline(host, threads, args.cpu_seconds)
        # the combined step's CAPE outputs against the reference's on every
        # column the baseline finished
        idx, ref = gathered(ref_res)
        sel = torch.from_numpy(idx).to(outs[0].device)
        ofn, lfn = sigma_oracle(host, SOURCE, ADIABAT)
        acc['vs_reference_cpu_sample'] = vs_reference(
            [x[sel].cpu().numpy() for x in outs[:4]], ref, at(ofn, idx),
            lambda j, k: lfn(int(idx[j]), k))
        acc_plain['vs_reference_cpu_sample'] = vs_reference(
            [x[sel].cpu().numpy() for x in outs_p[:4]], ref, at(ofn, idx),
            lambda j, k: lfn(int(idx[j]), k))
        del host
        log(f"[bench] vs reference: {acc['vs_reference_cpu_sample']}")
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
                     'ops_per_launch': ops, 'ms_per_launch': ms_dom,
                     # the same unit weights over the work the kernel executed
                     # (its own iteration counts rather than the plain method's)
                     'ops_executed_per_launch': ops_exec,
                     'achieved_executed': ops_exec / (ms_dom / 1e3) / 1e12,
                     'hbm_gbs_algorithmic': bytes_per_column(NLEV) * value / 1e9},
        'plain_kernel': {'value': ncol / (ms_plain / 1e3), 'ms_per_step': ms_plain,
                         'accuracy': acc_plain},
        'speedup_vs_plain': ms_plain / ms,
        'attribution_cape_only_ms': ms_cape,
        'srh_only_ms': ms_srh,
        'kernel_counts': counts,
        'surrogate': sur,
        'accuracy': acc,
        'configs': configs,
        'ragged_200k': rag,
        'host_arrays_pcie_inclusive': host_path,
        'cpu_baseline': cpu,
    }
    emit(line)


if __name__ == '__main__':
    main()
