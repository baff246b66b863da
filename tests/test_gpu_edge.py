"""getcape's edge branches on the device, against the FP32 oracle.

The soundings of tests/edge_cases.py drive every kernel path into the
branches that only the oracle (pinned bit for bit to the compiled reference)
had seen: most-unstable with the surface above 500 hPa
(CAPE_CODE_model_lev.f90:264-267), mixed-layer with the second level above
ml_depth and with the whole column inside it (:287-301), the 100-iteration
limit that zeroes a column (:464-474), NaN levels, and the ts <= 0 gate.

Paths: the plain kernel (accurate and fast math), the warm start
(extrapolation, with the committed surrogate, fast math) and ragged plans
(plain and warm).  Bar: MUlev equal, NaN exactly where the oracle's are,
and for every column whose fate the reference decides robustly, the
oracle's: stopped by the iteration limit (CAPE = CIN = 0) where it stops,
CAPE/CIN within the tolerance elsewhere.

Robustly means that the oracle decides it the same way when its exp, log
and pow results are moved by up to an ulp (16 seeds of the oracle's
perturbed runs: the device's functions differ from the host C library's by
that much) and in its FP64 run.  Most of the edge grid's 450 hPa columns are
not so: whether the reference's relaxed iteration (x + 0.3 (f(x) - x))
converges there turns on rounding, since near their fixed points the
saturated map's slope is about -10, so the relaxed step expands on one side
of the saturation kink and contracts on the other.  There the bar is:
CAPE = CIN = 0 (stopped), or within the tolerance of an oracle run that
converged.
"""
import numpy as np
import pytest

from tests import edge_cases
from tests.tolerances import cape_ok

pytestmark = pytest.mark.gpu

ADIABATS = (1, 2, 3, 4)


@pytest.fixture(scope='module')
def hip():
    from xcape_amd import _hip
    assert _hip.device_count() >= 1, 'no HIP device visible'
    return _hip


SEEDS = range(1, 17)


@pytest.fixture(scope='module')
def ref(oracle_lib):
    """Oracle outputs for every (source, ml_depth) case and adiabat: the FP32
    run with its stats, which columns it stops, which columns it decides
    robustly, and the runs (FP64, perturbed) that converged where it
    stopped."""
    arr = edge_cases.grid()
    a64 = [a.astype(np.float64) for a in arr]
    live = arr[4] > 0
    out = {}
    for src, ml in edge_cases.CASES:
        for adi in ADIABATS:
            o32 = oracle_lib.cape(*arr, 500., src, ml, adi, stats=True)
            s32 = (o32[0] == 0) & (o32[1] == 0) & live
            robust = np.ones(len(live), bool)
            alt = [oracle_lib.cape(*a64, 500., src, ml, adi, fp64=True)]
            alt += [oracle_lib.cape(*arr, 500., src, ml, adi, perturb=s) for s in SEEDS]
            for o in alt:
                robust &= ((o[0] == 0) & (o[1] == 0) & live) == s32
            out[(src, ml, adi)] = dict(o32=o32, s32=s32, robust=robust, alt=alt)
    return out


def _check(name, got, nonconv, r):
    cape, cin, mul = (np.asarray(x) for x in got[:3])
    o = r['o32']
    robust = r['robust']
    print(f'{name}: nonconverged {nonconv} (oracle {o[-1]["nonconverged"]}), '
          f'ill-conditioned columns {np.nonzero(~robust)[0].tolist()}, '
          f'max|dCAPE| robust {np.nanmax(np.abs(cape - o[0])[robust]):.3g}', flush=True)
    nan = np.isnan(o[0])
    np.testing.assert_array_equal(np.isnan(cape), nan, err_msg=name)
    np.testing.assert_array_equal(mul, o[2], err_msg=name)
    # robust columns: the oracle's outcome
    zero = (o[0] == 0) & (o[1] == 0) & robust
    assert np.all(cape[zero] == 0) and np.all(cin[zero] == 0), name
    ok = ~nan & robust
    assert cape_ok(cape[ok], o[0][ok]).all(), name
    assert cape_ok(cin[ok], o[1][ok]).all(), name
    # ill-conditioned ones: stopped, or the values of an oracle run that
    # converged there
    for i in np.nonzero(~robust)[0]:
        if cape[i] == 0 and cin[i] == 0:
            continue
        runs = [x for x in [o] + r['alt'] if not (x[0][i] == 0 and x[1][i] == 0)]
        assert any(cape_ok(cape[i], x[0][i]) and cape_ok(cin[i], x[1][i]) for x in runs), \
            (name, i, cape[i], [float(x[0][i]) for x in runs])
    if nonconv is not None:
        # the columns each run stops: the robust ones exactly, plus any of
        # the ill-conditioned ones
        n_ill = int(np.count_nonzero(~robust))
        n_stop32 = o[-1]['nonconverged']
        assert n_stop32 - n_ill <= nonconv <= n_stop32 + n_ill, name


@pytest.mark.parametrize('adiabat', ADIABATS)
@pytest.mark.parametrize('variant', ['plain', 'fast', 'extrapolation', 'fast_extrapolation'])
def test_grid_kernels(hip, ref, variant, adiabat):
    arr = edge_cases.grid()
    warm = 'extrapolation' if 'extrapolation' in variant else False
    fast = variant.startswith('fast')
    for src, ml in edge_cases.CASES:
        print(f'[edge] {variant} source={src} ml_depth={ml} adiabat={adiabat}', flush=True)
        g = hip.cape(*arr, pinc=500., source=src, ml_depth=ml, adiabat=adiabat,
                     warm_start=warm, fast_math=fast, stats=True)
        _check(f'{variant} s{src} ml{ml} a{adiabat}', g[:4], g[4]['nonconverged'],
               ref[(src, ml, adiabat)])


@pytest.mark.parametrize('adiabat', ADIABATS)
def test_warm_with_surrogate(hip, ref, adiabat):
    from xcape_amd import surrogate
    surrogate.Surrogate.load().activate()
    arr = edge_cases.grid()
    for src, ml in edge_cases.CASES:
        print(f'[edge] surrogate source={src} ml_depth={ml} adiabat={adiabat}', flush=True)
        g = hip.cape(*arr, pinc=500., source=src, ml_depth=ml, adiabat=adiabat,
                     warm_start=True, stats=True)
        _check(f'surrogate s{src} ml{ml} a{adiabat}', g[:4], g[4]['nonconverged'],
               ref[(src, ml, adiabat)])


@pytest.mark.parametrize('adiabat', ADIABATS)
@pytest.mark.parametrize('warm', [False, True])
def test_ragged_plans(hip, ref, adiabat, warm):
    """The edge grid as a ragged batch (every column 30 levels, one source
    and adiabat per plan): the persistent ragged kernel and the warm plan."""
    import torch

    from xcape_amd import ragged
    p, t, td, ps, ts, tds = edge_cases.grid()
    ncol, nk = t.shape
    offsets = np.arange(ncol + 1, dtype=np.int64) * nk
    for src, ml in edge_cases.CASES:
        print(f'[edge] ragged warm={warm} source={src} ml_depth={ml} adiabat={adiabat}',
              flush=True)
        b = ragged.RaggedBatch(offsets, p.reshape(-1), t.reshape(-1), td.reshape(-1), ps, ts,
                               tds, np.full(ncol, src, np.int32),
                               np.full(ncol, adiabat, np.int32))
        plan = ragged.RaggedPlan(b, 500., ml, warm_start=warm)
        st = plan.run(stats=True)
        plan.run()          # the captured graph's replay
        torch.cuda.synchronize()
        got = [x.cpu().numpy() for x in (b.cape, b.cin, b.mulev)]
        plan.close()
        _check(f'ragged warm={warm} s{src} ml{ml} a{adiabat}', got, st['nonconverged'],
               ref[(src, ml, adiabat)])
