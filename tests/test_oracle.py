"""Pin the CPU oracle before trusting it as the GPU's checker.

1. The reference's own goldens (test/test_core.py tolerances).
2. Bit-exact agreement with the reference Fortran compiled from source into
   oracle/_ref (skipped where that build is absent), over every source,
   adiabat and the edge branches no fixture covers.
3. Closed forms and invariants the mathematics fixes.
"""
import ctypes

import numpy as np
import pytest

from xcape_amd import synthetic


def _ref():
    from oracle import ref_driver
    if not ref_driver.available():
        pytest.skip('reference build (oracle/_ref) not present')
    return ref_driver


def _snd_args(s):
    P, T, TD = s['pressure'], s['temperature'], s['dewpoint']
    return (P[:, 1:], T[:, 1:], TD[:, 1:], P[:, 0], T[:, 0], TD[:, 0])


# ---------------------------------------------------------------- goldens --
def test_soundings_goldens(oracle_lib, soundings):
    s = soundings
    a = _snd_args(s)
    c, ci, _, _ = oracle_lib.cape(*a, pinc=100., source=1)
    np.testing.assert_almost_equal(c, s['SB_CAPE_pinc100'], 0)
    np.testing.assert_almost_equal(ci, s['SB_CIN_pinc100'], 0)
    c, ci, _, _ = oracle_lib.cape(*a, pinc=1000., source=3, ml_depth=500.)
    np.testing.assert_almost_equal(c, s['ML_CAPE_pinc1000_mldepth500'], 0)
    np.testing.assert_almost_equal(ci, s['ML_CIN_pinc1000_mldepth500'], 0)
    c, ci, mu, z = oracle_lib.cape(*a, pinc=100., source=2)
    np.testing.assert_almost_equal(c, s['MU_CAPE_pinc100'], 0)
    np.testing.assert_almost_equal(ci, s['MU_CIN_pinc100'], 0)
    np.testing.assert_array_equal(mu, s['MU_lv_pinc100'].astype(np.int32))
    np.testing.assert_almost_equal(z, s['MU_z_pinc100'], 0)


def test_era5_goldens(oracle_lib, era5):
    e = era5
    lev = e['level']
    st = oracle_lib.start_levels(lev, e['sfc_p'])
    # the fixture stores the 0-based masked argmin (core.py:286-288)
    np.testing.assert_array_equal(st - 1, e['sfc_start3d'].astype(np.int32))
    for src, key in ((1, 'sp500'), (3, 'ml300p500'), (2, 'mup500')):
        c, ci, _, _ = oracle_lib.cape(None, e['t'], e['td'], e['sfc_p'], e['sfc_t'], e['sfc_td'],
                                      500., src, 300., 1, p1d=lev, start_lev=st)
        np.testing.assert_almost_equal(c, e['sfc_cape' + key], 0)
        np.testing.assert_almost_equal(ci, e['sfc_cin' + key], 0)


def test_srh_goldens(oracle_lib, soundings, era5):
    s = soundings
    P, T, TD, U, V = (s[k] for k in ('pressure', 'temperature', 'dewpoint', 'u_wind_ms',
                                      'v_wind_ms'))
    srm, slm, *_ = oracle_lib.srh(P[:, 1:], T[:, 1:], TD[:, 1:], U[:, 1:], V[:, 1:], P[:, 0],
                                  T[:, 0], TD[:, 0], U[:, 0], V[:, 0], 3000.)
    np.testing.assert_almost_equal(srm, s['SRH03_model_lev_rm'], 5)
    np.testing.assert_almost_equal(slm, s['SRH03_model_lev_lm'], 5)
    e = era5
    st = oracle_lib.start_levels(e['level'], e['sfc_p'])
    srm, slm, *_ = oracle_lib.srh(None, e['t'], e['td'], e['u'], e['v'], e['sfc_p'],
                                  e['sfc_t'], e['sfc_td'], e['sfc_u'], e['sfc_v'], 3000.,
                                  p1d=e['level'], start_lev=st)
    np.testing.assert_almost_equal(srm, e['sfc_srh_rm'], 5)
    np.testing.assert_almost_equal(slm, e['sfc_srh_lm'], 5)


# ------------------------------------------------- compiled reference ------
def _bitexact(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize('adiabat', [1, 2, 3, 4])
@pytest.mark.parametrize('source', [1, 2, 3])
def test_cape_bitexact_vs_reference_sigma(oracle_lib, source, adiabat):
    R = _ref()
    c = synthetic.sigma_columns(384, 40, 1000 + adiabat)
    r = R.loopcape_ml(c['p'].T, c['t'].T, c['td'].T, c['ps'], c['ts'], c['tds'], 500., source,
                      500., adiabat)
    o = oracle_lib.cape(c['p'], c['t'], c['td'], c['ps'], c['ts'], c['tds'], 500., source,
                        500., adiabat)
    for x, y in zip(o, r):
        assert _bitexact(x, y)


@pytest.mark.parametrize('adiabat', [1, 3])
@pytest.mark.parametrize('source', [1, 2, 3])
def test_cape_bitexact_vs_reference_pressure(oracle_lib, source, adiabat):
    R = _ref()
    c = synthetic.pressure_columns(384, 1001)
    st = oracle_lib.start_levels(c['p'], c['ps'])
    r = R.loopcape_pl1d(c['t'].T, c['td'].T, c['p'], c['ps'], c['ts'], c['tds'], 500., source,
                        500., adiabat, st)
    o = oracle_lib.cape(None, c['t'], c['td'], c['ps'], c['ts'], c['tds'], 500., source, 500.,
                        adiabat, p1d=c['p'], start_lev=st)
    for x, y in zip(o, r):
        assert _bitexact(x, y)


@pytest.mark.parametrize('adiabat', [1, 3])
@pytest.mark.parametrize('source', [1, 2, 3])
def test_cape_bitexact_vs_reference_pressure_2d(oracle_lib, source, adiabat):
    """loopcape_pl (CAPE_CODE_pressure_lev.f90:4-84), unreachable from the
    reference's Python: pressure per column, each column from its own start
    level.  The columns' pressures differ (the shared levels scaled by up to
    0.1 % per column)."""
    R = _ref()
    c = synthetic.pressure_columns(384, 1002)
    st = oracle_lib.start_levels(c['p'], c['ps'])
    rng = np.random.default_rng(7)
    p2 = (c['p'][None, :] * (1.0 + 0.001 * rng.uniform(-1.0, 1.0, (384, 1)))).astype(np.float32)
    r = R.loopcape_pl(p2.T, c['t'].T, c['td'].T, c['ps'], c['ts'], c['tds'], 500., source,
                      500., adiabat, st)
    o = oracle_lib.cape(p2, c['t'], c['td'], c['ps'], c['ts'], c['tds'], 500., source, 500.,
                        adiabat, start_lev=st)
    for x, y in zip(o, r):
        assert _bitexact(x, y)
    # the shared levels repeated per column give loopcape_pl1d's results
    o1 = oracle_lib.cape(None, c['t'], c['td'], c['ps'], c['ts'], c['tds'], 500., source,
                         500., adiabat, p1d=c['p'], start_lev=st)
    ot = oracle_lib.cape(np.tile(np.asarray(c['p'], np.float32), (384, 1)), c['t'], c['td'],
                         c['ps'], c['ts'], c['tds'], 500., source, 500., adiabat, start_lev=st)
    for x, y in zip(o1, ot):
        assert _bitexact(x, y)


def test_cape_edge_branches_vs_reference(oracle_lib):
    """Branches no fixture reaches (tests/edge_cases.py): ML with the second
    level above ml_depth and with the whole column inside it, MU with a
    surface above 500 hPa, the cold-surface gate, NaN levels and the
    100-iteration limit; every adiabat."""
    import edge_cases
    R = _ref()
    p, t, td, ps, ts, tds = edge_cases.grid()
    for src, ml in edge_cases.CASES:
        for adi in (1, 2, 3, 4):
            r = R.loopcape_ml(p.T, t.T, td.T, ps, ts, tds, 500., src, ml, adi)
            o = oracle_lib.cape(p, t, td, ps, ts, tds, 500., src, ml, adi)
            for x, y in zip(o, r):
                assert _bitexact(x, y), (src, ml, adi)
    # the limit is reached: the surface source stops columns 47 and 102
    # (pseudo-liquid) and zeroes them
    o = oracle_lib.cape(p, t, td, ps, ts, tds, 500., 1, 500., 1, stats=True)
    assert o[-1]['nonconverged'] == 2
    assert o[0][47] == 0 and o[1][47] == 0 and o[0][102] == 0


def test_perturbed_library_runs(oracle_lib):
    """The oracle's perturbed runs (exp/log/pow results moved by up to an
    ulp, the stand-in for another C library) are active and mild where the
    reference's solve is well conditioned: on the 64x64x40 configuration
    nearly every column's CAPE changes in its low bits, within the CAPE/CIN
    tolerance wherever both runs converge.  They also show where it is not:
    in warm, moist sub-steps the map's slope is near -5, the relaxed step
    contracts by only ~0.8, and ulp-level noise in the map keeps |f(x) - x|
    near the 2e-4 K test, so that whether the reference converges within its
    100 iterations -- or zeroes the column -- turns on its library's
    rounding.  A few columns per thousand at most."""
    from tolerances import cape_ok
    c = synthetic.sigma_columns(4096, 40, 1000)
    a = [c[k] for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
    o = oracle_lib.cape(*a, 500., 1, 500., 1, stats=True)
    assert o[-1]['nonconverged'] == 0
    runs = [oracle_lib.cape(*a, 500., 1, 500., 1, perturb=s, stats=True) for s in (1, 2)]
    pos = o[0] > 0
    assert np.mean(runs[0][0][pos] != o[0][pos]) > 0.5
    assert np.any(runs[0][0] != runs[1][0])
    for q in runs:
        stopped = (q[0] == 0) & (q[1] == 0) & pos
        assert np.count_nonzero(stopped) == q[-1]['nonconverged']
        assert np.count_nonzero(stopped) <= 4
        ok = ~stopped
        assert cape_ok(q[0][ok], o[0][ok]).all() and cape_ok(q[1][ok], o[1][ok]).all()
        np.testing.assert_array_equal(q[2], o[2])


def test_iterate_perturbed_runs(oracle_lib):
    """The runs that also accept another point of each converged sub-step's
    tolerance band: with amplitude 0 they are the library-perturbed runs bit
    for bit; with 1e-4 K they move CAPE in many columns, by far less than the
    tolerance; with 1 K, far beyond the band, the alternatives that fail the
    convergence test are dropped, so the results stay as close (the band is
    enforced by the test, not assumed)."""
    from tolerances import cape_ok
    c = synthetic.sigma_columns(2048, 40, 1000)
    a = [c[k] for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
    for src, adi in ((1, 1), (2, 3)):
        lib = oracle_lib.cape(*a, 500., src, 500., adi, perturb=5)
        z = oracle_lib.cape(*a, 500., src, 500., adi, perturb=5, perturb_iterate=0.0)
        for x, y in zip(lib, z):
            np.testing.assert_array_equal(x, y)
        it = oracle_lib.cape(*a, 500., src, 500., adi, perturb=5, perturb_iterate=1e-4)
        pos = lib[0] > 0
        assert np.mean(it[0][pos] != lib[0][pos]) > 0.1
        ok = ~((it[0] == 0) & (it[1] == 0) & pos)
        assert np.mean(cape_ok(it[0][ok], lib[0][ok]) & cape_ok(it[1][ok], lib[1][ok])) > 0.999
        np.testing.assert_array_equal(it[2], lib[2])
        far = oracle_lib.cape(*a, 500., src, 500., adi, perturb=5, perturb_iterate=1.0)
        ok = ~((far[0] == 0) & (far[1] == 0) & pos)
        assert np.mean(cape_ok(far[0][ok], lib[0][ok]) & cape_ok(far[1][ok], lib[1][ok])) > 0.999


@pytest.mark.parametrize('pinc', [100., 1000., 2500.])
def test_cape_bitexact_vs_reference_soundings(oracle_lib, soundings, pinc):
    R = _ref()
    a = _snd_args(soundings)
    for src in (1, 2, 3):
        r = R.loopcape_ml(a[0].T, a[1].T, a[2].T, a[3], a[4], a[5], pinc, src, 500., 3)
        o = oracle_lib.cape(*a, pinc=pinc, source=src, ml_depth=500., adiabat=3)
        for x, y in zip(o, r):
            assert _bitexact(x, y)


def test_srh_bitexact_vs_reference(oracle_lib):
    R = _ref()
    c = synthetic.sigma_columns(64, 137, 1004, winds=True)
    h_r, hs_r = R.stdheight_ml(c['p'].T, c['t'].T, c['td'].T, c['ps'], c['ts'], c['tds'], 2.)
    srm, slm, rm, lm, mw = oracle_lib.srh(c['p'], c['t'], c['td'], c['u'], c['v'], c['ps'],
                                          c['ts'], c['tds'], c['us'], c['vs'], 3000.)
    h_o, hs_o = oracle_lib.stdheight(c['p'], c['t'], c['td'], c['ps'], c['ts'], c['tds'], 2.)
    assert _bitexact(h_o, h_r.T) and _bitexact(hs_o, hs_r)
    rm_r, lm_r, mw_r = R.bunkers_ml(c['u'].T, c['v'].T, h_r.astype(np.float32), c['us'],
                                    c['vs'], hs_r.astype(np.float32))
    assert _bitexact(rm, rm_r.T) and _bitexact(lm, lm_r.T) and _bitexact(mw, mw_r.T)
    s_r = R.sreh_ml(c['u'].T, c['v'].T, h_r, c['us'], c['vs'], hs_r, rm_r[0], rm_r[1],
                    lm_r[0], lm_r[1], 3000.)
    assert _bitexact(srm, s_r[0]) and _bitexact(slm, s_r[1])


def test_srh_bitexact_vs_reference_pressure(oracle_lib):
    R = _ref()
    c = synthetic.pressure_columns(48, 1002, winds=True)
    st = oracle_lib.start_levels(c['p'], c['ps'])
    h_r, hs_r = R.stdheight_pl1d(c['t'].T, c['td'].T, c['p'], c['ps'], c['ts'], c['tds'], 2., st)
    h_o, hs_o = oracle_lib.stdheight(None, c['t'], c['td'], c['ps'], c['ts'], c['tds'], 2.,
                                     p1d=c['p'], start_lev=st)
    assert _bitexact(h_o, h_r.T) and _bitexact(hs_o, hs_r)
    srm, slm, rm, lm, mw = oracle_lib.srh(None, c['t'], c['td'], c['u'], c['v'], c['ps'],
                                          c['ts'], c['tds'], c['us'], c['vs'], 3000.,
                                          p1d=c['p'], start_lev=st)
    rm_r, lm_r, mw_r = R.bunkers_pl(c['u'].T, c['v'].T, h_r.astype(np.float32), c['us'],
                                    c['vs'], hs_r.astype(np.float32), st)
    assert _bitexact(rm, rm_r.T) and _bitexact(lm, lm_r.T) and _bitexact(mw, mw_r.T)
    s_r = R.sreh_pl(c['u'].T, c['v'].T, h_r, c['us'], c['vs'], hs_r, rm_r[0], rm_r[1],
                    lm_r[0], lm_r[1], 3000., st)
    assert _bitexact(srm, s_r[0]) and _bitexact(slm, s_r[1])


# ------------------------------------------ closed forms and invariants ----
def _f32fn(lib, name, nargs):
    f = getattr(lib, name)
    f.restype = ctypes.c_float
    f.argtypes = [ctypes.c_float] * nargs
    return f


def test_bolton_saturation_closed_form(oracle_lib):
    """getqvs/getqvi against Bolton (1980) eq. 10 and r = eps e/(p - e) in
    float64 (a swapped constant or a dropped term shows up at 1e-3)."""
    qvs = _f32fn(oracle_lib.lib(), 'xo_getqvs_f32', 2)
    qvi = _f32fn(oracle_lib.lib(), 'xo_getqvi_f32', 2)
    eps = 287.04 / 461.5
    for p in (100000., 85000., 50000., 20000.):
        for t in (240., 260., 273.15, 290., 305.):
            e = 611.2 * np.exp(17.67 * (t - 273.15) / (t - 29.65))
            np.testing.assert_allclose(qvs(p, t), eps * e / (p - e), rtol=2e-6)
            ei = 611.2 * np.exp(21.8745584 * (t - 273.15) / (t - 7.66))
            np.testing.assert_allclose(qvi(p, t), eps * ei / (p - ei), rtol=2e-6)
    # at 0 degC liquid and ice saturation agree (both equal 611.2 Pa)
    np.testing.assert_allclose(qvs(80000., 273.15), qvi(80000., 273.15), rtol=1e-6)


def test_thetae_closed_form(oracle_lib):
    """getthe: saturated air (td == t) uses T_LCL = T; then theta-e equals
    T (p0/p)^(0.2854(1-0.28 r)) exp((3376/T - 2.54) r (1 + 0.81 r))."""
    the = _f32fn(oracle_lib.lib(), 'xo_getthe_f32', 4)
    for p, t, r in ((100000., 300., 0.02), (70000., 280., 0.008), (50000., 260., 0.001)):
        ref = t * (1e5 / p) ** (0.2854 * (1 - 0.28 * r)) * np.exp((3376. / t - 2.54) * r *
                                                                  (1 + 0.81 * r))
        np.testing.assert_allclose(the(p, t, t, r), ref, rtol=3e-6)
    # dry air: theta-e reduces to the potential temperature T (p0/p)^0.2854
    np.testing.assert_allclose(the(50000., 260., 200., 0.0), 260. * 2 ** 0.2854, rtol=3e-6)


def test_dry_ascent_is_poisson(oracle_lib):
    """An unsaturated sub-step has a closed form independent of the iterate:
    theta2 = theta1 (p2/p1)^(R_m/c_pm - R_d/c_p) (CAPE_CODE_model_lev.f90:460-462
    with ql = qi = 0).  Check every logged sub-step of a bone-dry column."""
    c = synthetic.sigma_columns(8, 40, 3)
    td = c['t'] - 60.0          # dewpoint 60 K below: never saturates below 100 hPa
    rec = oracle_lib.cape_log(c['p'], c['t'], td, c['ps'], c['ts'], c['tds'] - 40.0,
                              pinc=500., source=1, adiabat=1)
    assert len(rec) > 100
    # rec: p2, theta-e, qt, th1, th2, iters.  Record i runs from p1 = p2[i-1]
    # to p2[i]; it was unsaturated when the pseudo-adiabatic update left qt
    # unchanged (qt[i+1] == qt[i]).
    rd, rv, cp, cpv = 287.04, 461.5, 1005.7, 1875.0
    p1 = rec[:-2, 0].astype(np.float64)
    p2 = rec[1:-1, 0].astype(np.float64)
    qt = rec[1:-1, 2].astype(np.float64)
    dry = (p2 < p1) & (rec[2:, 0] < rec[1:-1, 0]) & (rec[2:, 2] == rec[1:-1, 2])
    th1 = rec[1:-1, 3].astype(np.float64)
    th2 = rec[1:-1, 4].astype(np.float64)
    expo = (rd + rv * qt) / (cp + cpv * qt) - rd / cp
    pred = th1 * (p2 / p1) ** expo
    assert dry.sum() > 100
    np.testing.assert_allclose(th2[dry], pred[dry], rtol=3e-6)


def test_invariants(oracle_lib):
    c = synthetic.sigma_columns(600, 40, 21)
    a = [c[k] for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
    for src in (1, 2, 3):
        cape, cin, mu, z = oracle_lib.cape(*a, pinc=500., source=src, adiabat=1)
        assert np.all(cape >= 0) and np.all(cin >= 0)
        # permutation equivariance over columns
        perm = np.random.default_rng(0).permutation(600)
        pc = oracle_lib.cape(*[x[perm] for x in a], pinc=500., source=src, adiabat=1)
        np.testing.assert_array_equal(pc[0], cape[perm])
        np.testing.assert_array_equal(pc[1], cin[perm])
    # MU whose most unstable level is the surface equals the surface parcel
    sb = oracle_lib.cape(*a, pinc=500., source=1, adiabat=1)
    mu = oracle_lib.cape(*a, pinc=500., source=2, adiabat=1)
    at_sfc = mu[2] == 1
    assert at_sfc.sum() > 100
    np.testing.assert_array_equal(mu[0][at_sfc], sb[0][at_sfc])
    np.testing.assert_array_equal(mu[1][at_sfc], sb[1][at_sfc])


def test_pressure_below_ground_is_never_read(oracle_lib):
    c = synthetic.pressure_columns(300, 77)
    st = oracle_lib.start_levels(c['p'], c['ps'])
    t2, td2 = c['t'].copy(), c['td'].copy()
    below = np.isnan(t2)
    assert below.any()
    t2[below] = 1.0e4
    td2[below] = -1.0e4
    for src in (1, 2, 3):
        a = oracle_lib.cape(None, c['t'], c['td'], c['ps'], c['ts'], c['tds'], 500., src, 500.,
                            1, p1d=c['p'], start_lev=st)
        b = oracle_lib.cape(None, t2, td2, c['ps'], c['ts'], c['tds'], 500., src, 500., 1,
                            p1d=c['p'], start_lev=st)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


def test_fp32_close_to_fp64(oracle_lib):
    """The FP32 algorithm stays near its FP64 evaluation: CAPE within
    max(2 J/kg, 0.5 %) (bounds the reference's own rounding noise)."""
    c = synthetic.sigma_columns(400, 40, 31)
    a = [c[k] for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
    for src in (1, 2, 3):
        f = oracle_lib.cape(*a, pinc=500., source=src, adiabat=3)
        d = oracle_lib.cape(*[x.astype(np.float64) for x in a], pinc=500., source=src,
                            adiabat=3, fp64=True)
        err = np.abs(f[0] - d[0])
        assert np.all(err <= np.maximum(2.0, 5e-3 * np.abs(d[0]))), err.max()


def test_pinc_refinement_converges(oracle_lib):
    """Halving pinc changes CAPE less and less (the sub-stepping converges)."""
    c = synthetic.sigma_columns(200, 40, 41)
    a = [c[k].astype(np.float64) for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
    vals = [oracle_lib.cape(*a, pinc=pi, source=1, adiabat=1, fp64=True)[0]
            for pi in (2000., 1000., 500., 250.)]
    d1 = np.abs(vals[1] - vals[0]).mean()
    d3 = np.abs(vals[3] - vals[2]).mean()
    assert d3 < d1


def test_srh_unidirectional_shear_closed_form(oracle_lib):
    """With v == 0 everywhere and u increasing with height, SREH reduces to
    -c_v (u(top) - u(0)) for a storm motion (c_u, c_v) (sum telescopes)."""
    n, nk = 5, 30
    z = np.tile(np.linspace(100., 9000., nk), (n, 1))
    u = 0.004 * z
    v = np.zeros_like(u)
    us, vs, zs = np.zeros(n), np.zeros(n), np.full(n, 2.0)
    cu = np.full(n, 3.0)
    cv = np.linspace(-5, 5, n)
    srm, slm = oracle_lib.sreh(u, v, z, us, vs, zs, cu, cv, cu, -cv, 3000.)
    utop = 0.004 * 3000.0
    np.testing.assert_allclose(srm, -cv * (utop - 0.0), rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(slm, cv * (utop - 0.0), rtol=1e-12, atol=1e-9)


def test_bunkers_closed_form(oracle_lib):
    """Linear hodograph u = a z, v = 0 sampled exactly at the 500 m nodes:
    mean wind = mean of the 13 node values and RM = mean + 7.5 (0, -1)."""
    n = 3
    z = np.tile(np.arange(250., 12001., 250.), (n, 1)).astype(np.float32)
    a = np.array([0.002, 0.004, 0.008])[:, None]
    u = (a * z).astype(np.float32)
    v = np.zeros_like(u)
    zs = np.zeros(n, np.float32)
    rm, lm, mw = oracle_lib.bunkers(u, v, z, np.zeros(n, np.float32), np.zeros(n, np.float32), zs)
    nodes = np.concatenate([[0.0], np.arange(500., 6001., 500.)])
    mean_u = (a[:, 0:1] * nodes).mean(axis=1)
    np.testing.assert_allclose(mw[:, 0], mean_u, rtol=1e-6)
    np.testing.assert_allclose(rm[:, 1], -7.5, rtol=1e-6)
    np.testing.assert_allclose(lm[:, 1], 7.5, rtol=1e-6)
    np.testing.assert_allclose(rm[:, 0], mean_u, rtol=1e-6)


# ------------------------------------------------------------- LFC / EL --
def _dry_column_levels(nk=80, ncol=40, seed=5):
    """Bone-dry soundings whose buoyancy changes sign several times: the
    environment's theta oscillates around the surface parcel's.  Returns the
    inputs and, computed independently in float64 from the closed-form dry
    ascent (the sub-steps telescope to theta_p = theta_s (p/p_s)^e with
    e = R_m/c_pm - R_d/c_p), the expected LFC and EL [hPa]."""
    rng = np.random.default_rng(seed)
    rd, rv, cp, cpv, g = 287.04, 461.5, 1005.7, 1875.0, 9.81
    kap, eps = rd / cp, rd / rv
    ps = rng.uniform(980., 1020., ncol)
    # top at 0.3 ps: with a dewpoint 100 K below, the parcel never saturates
    p = ps[:, None] * np.linspace(0.99, 0.3, nk)[None, :]
    ts = rng.uniform(295., 305., ncol)
    amp = rng.uniform(1.0, 3.0, ncol)
    per = rng.uniform(15000., 30000., ncol)
    th_s = ts * (1000. / ps) ** kap
    th_env = th_s[:, None] + amp[:, None] * np.cos(2 * np.pi * (ps[:, None] - p) * 100. / per[:, None])
    t = th_env * (p / 1000.) ** kap
    td = t - 100.
    tds = ts - 100.
    # float32 inputs, as the oracle will see them
    P, T, TD = (np.float32(x) for x in (p, t - 273.15, td - 273.15))
    PS, TS, TDS = (np.float32(x) for x in (ps, ts - 273.15, tds - 273.15))

    def qsat(pp, tt):   # Bolton over liquid, as getqvs
        es = 611.2 * np.exp(17.67 * (tt - 273.15) / (tt - 29.65))
        return eps * es / (pp - es)
    lfc = np.full(ncol, np.nan)
    el = np.full(ncol, np.nan)
    for i in range(ncol):
        pa = 100. * np.concatenate([[PS[i]], P[i]]).astype(np.float64)
        ta = 273.15 + np.concatenate([[TS[i]], T[i]]).astype(np.float64)
        tda = 273.15 + np.concatenate([[TDS[i]], TD[i]]).astype(np.float64)
        q = qsat(pa, tda)
        th = ta * (1e5 / pa) ** kap
        thv = th * (1 + q / eps) / (1 + q)
        e = (rd + rv * q[0]) / (cp + cpv * q[0]) - rd / cp
        thp = th[0] * (pa / pa[0]) ** e
        b = g * (thp * (1 + q[0] / eps) / (1 + q[0]) - thv) / thv
        b[0] = 0.0
        for k in range(1, len(pa)):
            b1, b2 = b[k - 1], b[k]
            enter, leave = b2 >= 0 > b1, b2 < 0 < b1
            parea = b2 if enter else (b1 if leave else (0.0 if b2 < 0 else b1 + b2))
            if parea > 0:
                ps_ = pa[k - 1] + (pa[k] - pa[k - 1]) * (0 - b1) / (b2 - b1)
                if np.isnan(lfc[i]):
                    lfc[i] = ps_ if enter else pa[k - 1]
                el[i] = ps_ if leave else pa[k]
            if pa[k] <= 10000. and b2 < 0:
                break
    return (P, T, TD, PS, TS, TDS), lfc / 100., el / 100.


def test_lfc_el_dry_closed_form(oracle_lib):
    """LFC and EL (no reference output) against the same crossings computed
    independently from the closed-form dry ascent in float64."""
    args, lfc, el = _dry_column_levels()
    r = oracle_lib.cape(*args, pinc=500., source=1, adiabat=1, levels=True)
    got_lfc, got_el = r[4], r[5]
    assert np.all(np.isfinite(lfc)) and np.all(np.isfinite(el))
    np.testing.assert_allclose(got_lfc, lfc, atol=0.01)
    np.testing.assert_allclose(got_el, el, atol=0.01)


@pytest.mark.parametrize('source', [1, 2, 3])
@pytest.mark.parametrize('adiabat', [1, 3])
def test_lfc_el_invariants(oracle_lib, source, adiabat):
    """LFC and EL exist exactly when CAPE is positive, the LFC lies at or
    below the EL (higher pressure) and at or above the surface, and asking for
    them changes no other output."""
    c = synthetic.sigma_columns(1500, 60, 33)
    a = [c[k] for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
    r = oracle_lib.cape(*a, pinc=500., source=source, adiabat=adiabat, levels=True)
    cape, lfc, el = r[0], r[4], r[5]
    pos = cape > 0
    assert pos.sum() > 100
    np.testing.assert_array_equal(np.isfinite(lfc), pos)
    np.testing.assert_array_equal(np.isfinite(el), pos)
    assert np.all(lfc[pos] >= el[pos])
    assert np.all(lfc[pos] <= c['ps'][pos] + 1e-3)
    for x, y in zip(r[:4], oracle_lib.cape(*a, pinc=500., source=source, adiabat=adiabat)):
        np.testing.assert_array_equal(x, y)


def test_dewpoint_from_specific_humidity_round_trip(oracle_lib):
    """The humidity adaptor (xcape_amd.core.dewpoint_from_specific_humidity)
    inverts the reference's own saturation mixing ratio: getqvs(p, Td(q, p))
    gives back r = q / (1 - q), over the range of tropospheric humidities."""
    from xcape_amd.core import dewpoint_from_specific_humidity as td_of
    qvs = _f32fn(oracle_lib.lib(), 'xo_getqvs_f32', 2)
    rng = np.random.default_rng(7)
    p = rng.uniform(100., 1050., 400)
    q = 10 ** rng.uniform(-5.5, -1.6, 400)
    td = td_of(q, p)
    assert td.dtype == np.float64 or td.dtype == np.float32
    r = q / (1 - q)
    got = np.array([qvs(100. * pp, 273.15 + tt) for pp, tt in zip(p, td)])
    np.testing.assert_allclose(got, r, rtol=2e-5)
    # saturated air at 0 degC: e = 6.112 hPa gives Td = 0
    eps = 287.04 / 461.5
    r0 = eps * 6.112 / (850. - 6.112)
    assert abs(float(td_of(r0 / (1 + r0), 850.))) < 1e-9
