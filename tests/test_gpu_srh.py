"""GPU parity of the helicity chain (heights, Bunkers motion, SREH) and of the
fused CAPE+SRH kernel.  Bar: Bunkers motion within 0.05 m/s, SRH within
1 m^2/s^2 of the oracle; the reference's goldens within their own 1.5e-5."""
import numpy as np
import pytest

from tests.tolerances import BUNKERS_ABS, SRH_ABS, cape_ok
from xcape_amd import synthetic

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def hip():
    from xcape_amd import _hip
    assert _hip.device_count() >= 1, 'no HIP device visible'
    return _hip


def _cmp_srh(g, o, name):
    srm, slm, rm, lm, mw = g
    osrm, oslm, orm, olm, omw = o
    d_srh = max(np.abs(srm - osrm).max(), np.abs(slm - oslm).max())
    d_bk = max(np.abs(rm - orm).max(), np.abs(lm - olm).max(), np.abs(mw - omw).max())
    print(f'{name}: max|dSRH|={d_srh:.3g} max|dBunkers|={d_bk:.3g} '
          f'exactSRH={np.mean(srm == osrm):.3f} exactBk={np.mean(rm == orm):.3f}', flush=True)
    assert d_srh <= SRH_ABS and d_bk <= BUNKERS_ABS


def test_soundings_srh_goldens(hip, soundings, oracle_lib):
    """Reference test_calc_srh_sigma (test/test_core.py:173-229)."""
    s = soundings
    P, T, TD, U, V = (s[k] for k in ('pressure', 'temperature', 'dewpoint', 'u_wind_ms',
                                      'v_wind_ms'))
    args = (P[:, 1:], T[:, 1:], TD[:, 1:], U[:, 1:], V[:, 1:], P[:, 0], T[:, 0], TD[:, 0],
            U[:, 0], V[:, 0])
    g = hip.srh(*args, depth=3000.)
    # The fixture holds float64 profiles, which the reference feeds unrounded
    # to its FP64 height and helicity routines; this entry point takes float32
    # profiles.  Against the goldens the bar is therefore the SRH tolerance;
    # against the oracle run on the same float32 inputs it is the reference's
    # own decimal=5.
    assert np.all(np.abs(g[0] - s['SRH03_model_lev_rm']) <= SRH_ABS)
    assert np.all(np.abs(g[1] - s['SRH03_model_lev_lm']) <= SRH_ABS)
    a32 = [np.asarray(x, np.float32) for x in args]
    o = oracle_lib.srh(*a32, 3000.)
    np.testing.assert_almost_equal(g[0], o[0], 5)
    np.testing.assert_almost_equal(g[1], o[1], 5)
    _cmp_srh(g, o, 'soundings')


def test_era5_srh_goldens(hip, era5, oracle_lib):
    """Reference test_calc_srh_pressure (test/test_core.py:231-286)."""
    e = era5
    st = oracle_lib.start_levels(e['level'], e['sfc_p'])
    args = (e['level'], e['t'], e['td'], e['u'], e['v'], e['sfc_p'], e['sfc_t'], e['sfc_td'],
            e['sfc_u'], e['sfc_v'])
    for start in (st, None):
        g = hip.srh(*args, depth=3000., p_shared=True, start_lev=start)
        assert np.all(np.abs(g[0] - e['sfc_srh_rm']) <= SRH_ABS)
        assert np.all(np.abs(g[1] - e['sfc_srh_lm']) <= SRH_ABS)
        a32 = [np.asarray(x, np.float32) for x in args]
        o = oracle_lib.srh(None, *a32[1:], 3000., p1d=a32[0], start_lev=st)
        np.testing.assert_almost_equal(g[0], o[0], 5)
        np.testing.assert_almost_equal(g[1], o[1], 5)
        _cmp_srh(g, o, 'era5')


@pytest.mark.parametrize('depth', [1000., 3000.])
def test_sigma_synthetic_srh(hip, oracle_lib, depth):
    c = synthetic.sigma_columns(2048, 137, 1004, winds=True)
    args = [c[k] for k in ('p', 't', 'td', 'u', 'v', 'ps', 'ts', 'tds', 'us', 'vs')]
    g = hip.srh(*args, depth=depth)
    o = oracle_lib.srh(*args, depth)
    _cmp_srh(g, o, f'sigma depth {depth}')


def test_levels_above_the_outputs(hip, oracle_lib):
    """Past the helicity's top and the last Bunkers level (6 km) the kernel
    only checks that heights keep rising, and walks a column again in full
    where they do not.  Columns that make it do so, against the oracle's
    full walk: an upper level's pressure set near the surface's (the heights
    drop below 6 km and climb through the Bunkers levels again, whose last
    brackets then win), a NaN temperature at the top, and the upper levels'
    pressures above the surface's (the last level below the surface: the
    reference's decreasing orientation).  Below 6 km, where the walk keeps
    the Bunkers sums as the heights rise: a level's pressure above the one
    before it (the heights fall there and cross some Bunkers levels again),
    and a NaN dewpoint (NaN heights from there up).  The untouched columns
    too."""
    c = synthetic.sigma_columns(1024, 137, 1005, winds=True)
    p, t, td = c['p'].copy(), c['t'].copy(), c['td'].copy()
    p[0:100, 120] = c['ps'][0:100] * 0.9
    t[100:200, 136] = np.nan
    p[200:300, 100:] = c['ps'][200:300, None] * 20.0
    p[300:400, 30] = c['p'][300:400, 24]
    td[400:500, 40] = np.nan
    args = [p, t, td] + [c[k] for k in ('u', 'v', 'ps', 'ts', 'tds', 'us', 'vs')]
    g = hip.srh(*args, depth=3000.)
    o = oracle_lib.srh(*args, 3000.)
    for x, y in zip(g, o):
        np.testing.assert_array_equal(np.isnan(x), np.isnan(y))
    np.testing.assert_almost_equal(g[0], o[0], 5)
    np.testing.assert_almost_equal(g[1], o[1], 5)
    for x, y in zip(g[2:], o[2:]):
        fin = np.isfinite(y)
        np.testing.assert_array_equal(x[fin], y[fin])
    # the first group's Bunkers levels were crossed again above the dropped
    # level, and the third group's column turned over: both change the motion
    ref = oracle_lib.srh(c['p'], c['t'], c['td'], *args[3:], 3000.)
    assert np.any(o[2][0:100] != ref[2][0:100])
    assert np.any(o[2][200:300] != ref[2][200:300])
    assert np.any(o[2][300:400] != ref[2][300:400])
    _cmp_srh(g, o, 'levels above the outputs')
    # the FP64 re-run takes the same paths (its storm motion in double): the
    # untouched columns within the usual bar; on the constructed ones the
    # helicity and the NaN pattern, as their Bunkers motion differs from the
    # reference's FP32 one by that one's rounding at large magnitudes (NaN
    # heights leave targets unbracketed, and the mean wind then sums their
    # -999999 fill values: |mean| ~ 5e5 m/s, float spacing ~0.06 m/s)
    g64 = hip.srh(*args, depth=3000., fp64=True)
    for x, y in zip(g64, o):
        np.testing.assert_array_equal(np.isnan(x), np.isnan(y))
    _cmp_srh([x[500:] for x in g64], [y[500:] for y in o],
             'levels above the outputs, FP64 chain, untouched columns')
    d_bk = max(np.nanmax(np.abs(g64[j][:500] - o[j][:500])) for j in (2, 3, 4))
    print(f'levels above the outputs, FP64 chain, constructed columns: '
          f'max|dBunkers|={d_bk:.3g}', flush=True)
    for j in (0, 1):
        assert np.nanmax(np.abs(g64[j][:500] - o[j][:500])) <= SRH_ABS


def test_pressure_synthetic_srh(hip, oracle_lib):
    c = synthetic.pressure_columns(2048, 1002, winds=True)
    st = oracle_lib.start_levels(c['p'], c['ps'])
    args = [c[k] for k in ('p', 't', 'td', 'u', 'v', 'ps', 'ts', 'tds', 'us', 'vs')]
    g = hip.srh(*args, depth=3000., p_shared=True)
    o = oracle_lib.srh(None, *args[1:], 3000., p1d=c['p'], start_lev=st)
    assert np.all(np.isfinite(g[0]))
    _cmp_srh(g, o, 'pressure')


def test_shallow_column_srh_is_zero(hip, oracle_lib):
    """No level above `depth`: the reference's sum is empty."""
    c = synthetic.sigma_columns(64, 20, 9, winds=True)
    args = [c[k] for k in ('p', 't', 'td', 'u', 'v', 'ps', 'ts', 'tds', 'us', 'vs')]
    g = hip.srh(*args, depth=1.0e6)
    assert np.all(g[0] == 0) and np.all(g[1] == 0)
    _cmp_srh(g, oracle_lib.srh(*args, 1.0e6), 'shallow')


@pytest.mark.parametrize('p_shared', [False, True])
def test_fp64_storm_motion_rerun(hip, oracle_lib, p_shared):
    """xcape_srh_f64 (Bunkers motion in FP64) against the FP32 chain and the
    oracle: the FP32 motion within 0.05 m/s of the FP64 one and SRH within
    1 m^2/s^2; the FP64 motion also within 0.05 m/s of the oracle's FP32."""
    if p_shared:
        c = synthetic.pressure_columns(2048, 1002, winds=True)
        st = oracle_lib.start_levels(c['p'], c['ps'])
        o_kw = dict(p1d=c['p'], start_lev=st)
    else:
        c = synthetic.sigma_columns(2048, 137, 1004, winds=True)
        o_kw = {}
    args = [c[k] for k in ('p', 't', 'td', 'u', 'v', 'ps', 'ts', 'tds', 'us', 'vs')]
    g64 = hip.srh(*args, depth=3000., p_shared=p_shared, fp64=True)
    g32 = hip.srh(*args, depth=3000., p_shared=p_shared)
    assert all(x.dtype == np.float64 for x in g64)
    _cmp_srh(g32, g64, f'fp32 vs fp64 chain (p_shared={p_shared})')
    o = (oracle_lib.srh(None, *args[1:], 3000., **o_kw) if p_shared
         else oracle_lib.srh(*args, 3000.))
    _cmp_srh(g64, o, f'fp64 chain vs oracle (p_shared={p_shared})')


def test_srh_in_parts_matches_one_launch(hip, monkeypatch):
    """A helicity-only call on 2^18 column-contiguous columns or more is
    staged in four parts, each part's helicity launch on a second stream
    beside the next part's copies; forced here down to 128 columns (3000
    columns: parts of 768, the last of 696).  The outputs equal the one-launch
    call's: sigma and shared pressure levels, FP32 and the FP64 re-run."""
    c = synthetic.sigma_columns(3000, 137, 31, winds=True)
    pc = synthetic.pressure_columns(3000, 32, winds=True)
    keys = ('p', 't', 'td', 'u', 'v', 'ps', 'ts', 'tds', 'us', 'vs')

    def run():
        out = [hip.srh(*(c[k] for k in keys), depth=d, fp64=f)
               for d in (1000., 3000.) for f in (False, True)]
        out.append(hip.srh(*(pc[k] for k in keys), depth=3000., p_shared=True))
        return out

    monkeypatch.delenv('XCAPE_STAGE_COLUMNS', raising=False)
    monkeypatch.delenv('XCAPE_SRH_PIPE_COLUMNS', raising=False)
    ref = run()
    monkeypatch.setenv('XCAPE_SRH_PIPE_COLUMNS', '128')
    for x, r in zip(run(), ref):
        for a, b in zip(x, r):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize('fast', [False, True])
def test_fused_cape_srh_matches_separate_calls(hip, fast):
    c = synthetic.sigma_columns(2048, 137, 1004, winds=True)
    a = [c[k] for k in ('p', 't', 'td', 'u', 'v', 'ps', 'ts', 'tds', 'us', 'vs')]
    f = hip.cape_srh(*a, pinc=500., source=2, adiabat=3, depth=3000., fast_math=fast)
    capes = hip.cape(a[0], a[1], a[2], a[5], a[6], a[7], pinc=500., source=2, adiabat=3,
                     fast_math=fast)
    srh = hip.srh(*a, depth=3000.)
    for x, y in zip(f[:4], capes):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(f[4:], srh):
        np.testing.assert_array_equal(x, y)


def test_fused_c4_shape_against_oracle(hip, oracle_lib):
    """The combined workload's recipe (137 sigma levels, most unstable,
    pseudo-ice, SRH 0-3 km) on a slice of columns, against the oracle."""
    c = synthetic.sigma_columns(1024, 137, 1004, winds=True)
    a = [c[k] for k in ('p', 't', 'td', 'u', 'v', 'ps', 'ts', 'tds', 'us', 'vs')]
    f = hip.cape_srh(*a, pinc=500., source=2, adiabat=3, depth=3000.)
    o = oracle_lib.cape(a[0], a[1], a[2], a[5], a[6], a[7], pinc=500., source=2, adiabat=3)
    assert np.all(cape_ok(f[0], o[0])) and np.all(cape_ok(f[1], o[1]))
    _cmp_srh(f[4:], oracle_lib.srh(*a, 3000.), 'fused C4')
