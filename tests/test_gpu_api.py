"""The reference's own golden tests through the drop-in API on the GPU.

Ports of test/test_core.py:68-116 (CAPE, sigma), :120-170 (CAPE, pressure),
:173-229 and :231-286 (SRH, sigma and pressure, output_var 'srh' and 'all'):
``xcape_amd.core.calc_cape`` / ``calc_srh`` with ``method='gpu'`` (and the
reference's own ``'fortran'``, which selects the same kernels), the float64
fixture arrays passed as given.  So the API's own code paths meet the goldens
and the oracle on the device: the host start levels of float64 pressure
input, the float64 helicity inputs, the (2, ...) storm-motion reshapes.

Bars: the reference's decimal=0 (CAPE, CIN, MUlev, zMUlev) and decimal=5
(SRH) against its goldens; against the FP32 oracle on the same inputs (f2py
rounds the CAPE inputs to float32), the CAPE/CIN tolerance, MUlev exact, and
Bunkers motion within 0.05 m/s.
"""
import numpy as np
import pytest

from tests.tolerances import BUNKERS_ABS, cape_ok
from xcape_amd import core

pytestmark = pytest.mark.gpu

DEC_CAPE = 0


@pytest.fixture(scope='module')
def gpu():
    from xcape_amd import _hip
    assert _hip.device_count() >= 1, 'no HIP device visible'
    return _hip


def _snd(s):
    P, T, TD = s['pressure'], s['temperature'], s['dewpoint']
    return P[:, 1:], T[:, 1:], TD[:, 1:], P[:, 0], T[:, 0], TD[:, 0]


@pytest.mark.parametrize('method', ['gpu', 'fortran'])
@pytest.mark.parametrize('sourcein,pinc', [('surface', 100), ('mixed-layer', 1000),
                                           ('most-unstable', 100)])
def test_calc_cape_sigma(gpu, soundings, oracle_lib, sourcein, pinc, method):
    """test_core.py:68-116 through calc_cape(method='gpu')."""
    s = soundings
    args = _snd(s)
    assert all(a.dtype == np.float64 for a in args)
    out = core.calc_cape(*args, source=sourcein, ml_depth=500., adiabat='pseudo-liquid',
                         pinc=pinc, method=method, vertical_lev='sigma')
    src = core._SOURCE[sourcein]
    o = oracle_lib.cape(*[np.asarray(a, np.float32) for a in args], pinc=pinc, source=src,
                        ml_depth=500., adiabat=1)
    cape, cin = out[0], out[1]
    assert cape.shape == (13,) and cin.shape == (13,)
    assert cape_ok(cape, o[0]).all() and cape_ok(cin, o[1]).all()
    if sourcein == 'most-unstable':
        assert len(out) == 4
        mulv, zmulv = out[2], out[3]
        np.testing.assert_almost_equal(cape, s['MU_CAPE_pinc100'], DEC_CAPE)
        np.testing.assert_almost_equal(cin, s['MU_CIN_pinc100'], DEC_CAPE)
        np.testing.assert_almost_equal(mulv, s['MU_lv_pinc100'].astype('int32'), DEC_CAPE)
        np.testing.assert_almost_equal(zmulv, s['MU_z_pinc100'], DEC_CAPE)
        np.testing.assert_array_equal(mulv, o[2])
    else:
        assert len(out) == 2
        key = 'SB' if sourcein == 'surface' else 'ML'
        kc = {'SB': 'SB_CAPE_pinc100', 'ML': 'ML_CAPE_pinc1000_mldepth500'}[key]
        ki = {'SB': 'SB_CIN_pinc100', 'ML': 'ML_CIN_pinc1000_mldepth500'}[key]
        np.testing.assert_almost_equal(cape, s[kc], DEC_CAPE)
        np.testing.assert_almost_equal(cin, s[ki], DEC_CAPE)


@pytest.mark.parametrize('sourcein,key', [('surface', 'sp500'), ('mixed-layer', 'ml300p500'),
                                          ('most-unstable', 'mup500')])
def test_calc_cape_pressure(gpu, era5, oracle_lib, sourcein, key):
    """test_core.py:120-170 through calc_cape(method='gpu'): float64 levels,
    so the start levels come from the API's host rule (core.py:286-289)."""
    e = era5
    out = core.calc_cape(e['level'], e['t'], e['td'], e['sfc_p'], e['sfc_t'], e['sfc_td'],
                         source=sourcein, ml_depth=300., adiabat='pseudo-liquid', pinc=500,
                         method='gpu', vertical_lev='pressure')
    np.testing.assert_almost_equal(out[0], e['sfc_cape' + key], DEC_CAPE)
    np.testing.assert_almost_equal(out[1], e['sfc_cin' + key], DEC_CAPE)
    st = oracle_lib.start_levels(e['level'], e['sfc_p'])
    f32 = [np.asarray(e[k], np.float32) for k in ('t', 'td', 'sfc_p', 'sfc_t', 'sfc_td')]
    o = oracle_lib.cape(None, *f32, 500., core._SOURCE[sourcein], 300., 1,
                        p1d=np.asarray(e['level'], np.float32), start_lev=st)
    assert cape_ok(out[0], o[0]).all() and cape_ok(out[1], o[1]).all()
    if sourcein == 'most-unstable':
        np.testing.assert_array_equal(out[2], o[2])


def _srh_oracle(oracle_lib, args, **kw):
    """The oracle on the inputs as the reference's wrappers see them: float64
    for the heights and the helicity, float32 roundings for Bunkers."""
    return oracle_lib.srh(*args, 3000., **kw)


@pytest.mark.parametrize('output_var,n_returns', [('all', 8), ('srh', 2)])
def test_calc_srh_sigma(gpu, soundings, oracle_lib, output_var, n_returns):
    """test_core.py:173-229 through calc_srh(method='gpu')."""
    s = soundings
    P, T, TD, U, V = (s[k] for k in ('pressure', 'temperature', 'dewpoint', 'u_wind_ms',
                                      'v_wind_ms'))
    args = (P[:, 1:], T[:, 1:], TD[:, 1:], U[:, 1:], V[:, 1:], P[:, 0], T[:, 0], TD[:, 0],
            U[:, 0], V[:, 0])
    out = core.calc_srh(*args, depth=3000, vertical_lev='sigma', output_var=output_var)
    assert len(out) == n_returns
    np.testing.assert_almost_equal(out[0], s['SRH03_model_lev_rm'], 5)
    np.testing.assert_almost_equal(out[1], s['SRH03_model_lev_lm'], 5)
    o = _srh_oracle(oracle_lib, args)
    np.testing.assert_almost_equal(out[0], o[0], 8)
    np.testing.assert_almost_equal(out[1], o[1], 8)
    if output_var == 'all':
        rm0, rm1, lm0, lm1, mw0, mw1 = out[2:]
        for g, r in ((rm0, o[2][:, 0]), (rm1, o[2][:, 1]), (lm0, o[3][:, 0]),
                     (lm1, o[3][:, 1]), (mw0, o[4][:, 0]), (mw1, o[4][:, 1])):
            assert g.shape == (13,)
            assert np.abs(g - r).max() <= BUNKERS_ABS


@pytest.mark.parametrize('output_var,n_returns', [('all', 8), ('srh', 2)])
def test_calc_srh_pressure(gpu, era5, oracle_lib, output_var, n_returns):
    """test_core.py:231-286 through calc_srh(method='gpu')."""
    e = era5
    args = (e['level'], e['t'], e['td'], e['u'], e['v'], e['sfc_p'], e['sfc_t'], e['sfc_td'],
            e['sfc_u'], e['sfc_v'])
    out = core.calc_srh(*args, depth=3000, vertical_lev='pressure', output_var=output_var)
    assert len(out) == n_returns
    np.testing.assert_almost_equal(out[0], e['sfc_srh_rm'], 5)
    np.testing.assert_almost_equal(out[1], e['sfc_srh_lm'], 5)
    st = oracle_lib.start_levels(e['level'], e['sfc_p'])
    o = oracle_lib.srh(None, *args[1:], 3000., p1d=args[0], start_lev=st)
    np.testing.assert_almost_equal(out[0], o[0], 8)
    np.testing.assert_almost_equal(out[1], o[1], 8)
    if output_var == 'all':
        for g, r in zip(out[2:], (o[2][:, 0], o[2][:, 1], o[3][:, 0], o[3][:, 1],
                                  o[4][:, 0], o[4][:, 1])):
            assert g.shape == (17,)
            assert np.abs(g - r).max() <= BUNKERS_ABS


def test_calc_srh_float32_inputs_match_the_f32_entry(gpu, soundings, oracle_lib):
    """float32 inputs take the float32 entry point; their float64 promotion
    through the float64 one gives the same helicity (FP64 arithmetic on the
    same values)."""
    s = soundings
    P, T, TD, U, V = (np.asarray(s[k], np.float32) for k in ('pressure', 'temperature',
                                                             'dewpoint', 'u_wind_ms',
                                                             'v_wind_ms'))
    args = (P[:, 1:], T[:, 1:], TD[:, 1:], U[:, 1:], V[:, 1:], P[:, 0], T[:, 0], TD[:, 0],
            U[:, 0], V[:, 0])
    a32 = core.calc_srh(*args, depth=3000, vertical_lev='sigma', output_var='all')
    a64 = core.calc_srh(*[np.asarray(a, np.float64) for a in args], depth=3000,
                        vertical_lev='sigma', output_var='all')
    for x, y in zip(a32, a64):
        np.testing.assert_array_equal(x, y)


def test_nd_shapes_through_the_api(gpu):
    """(4, 6, nlev) grids: the API's reshapes against the flat call, for
    CAPE (all four returns) and SRH with output_var='all' ((2, ...) motions
    split into components)."""
    from xcape_amd import synthetic
    c = synthetic.sigma_columns(24, 30, 17, winds=True)
    sh3, sh2 = (4, 6, 30), (4, 6)
    p, t, td, u, v = (c[k].reshape(sh3) for k in ('p', 't', 'td', 'u', 'v'))
    ps, ts, tds, us, vs = (c[k].reshape(sh2) for k in ('ps', 'ts', 'tds', 'us', 'vs'))
    out = core.calc_cape(p, t, td, ps, ts, tds, source='most-unstable', adiabat='pseudo-ice',
                         method='gpu', vertical_lev='sigma')
    flat = gpu.cape(c['p'], c['t'], c['td'], c['ps'], c['ts'], c['tds'], source=2, adiabat=3)
    for g, r in zip(out, flat):
        assert g.shape == sh2
        np.testing.assert_array_equal(g.reshape(-1), r)
    out = core.calc_srh(p, t, td, u, v, ps, ts, tds, us, vs, depth=3000, vertical_lev='sigma',
                        output_var='all')
    srm, slm, rm, lm, mw = gpu.srh(c['p'], c['t'], c['td'], c['u'], c['v'], c['ps'], c['ts'],
                                   c['tds'], c['us'], c['vs'], depth=3000.)
    assert all(x.shape == sh2 for x in out)
    np.testing.assert_array_equal(out[0].reshape(-1), srm)
    np.testing.assert_array_equal(out[1].reshape(-1), slm)
    for i, m in enumerate((rm, lm, mw)):
        np.testing.assert_array_equal(out[2 + 2 * i].reshape(-1), m[:, 0])
        np.testing.assert_array_equal(out[3 + 2 * i].reshape(-1), m[:, 1])
