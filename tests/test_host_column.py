"""The device column routines compiled for the host (tests/host_column.hip)
against the oracle, without a GPU.

The plain method must be bit-exact (the host build uses the same libm as the
oracle); the warm-started one must stay within the CAPE/CIN tolerance with
far fewer iterations.  Both layouts (column-contiguous and level-major) and
both grids (sigma, shared pressure levels with start levels) are covered.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..'))
from oracle import oracle as O  # noqa: E402
from tolerances import cape_ok  # noqa: E402
from xcape_amd import synthetic as S  # noqa: E402

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')


@pytest.fixture(scope='module')
def host_lib(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip('hipcc not available')
    out = str(tmp_path_factory.mktemp('host') / 'libhost_column.so')
    cmd = [HIPCC, '-O2', '-std=c++17', '-fPIC', '-shared', '--cuda-host-only',
           '-ffp-contract=off', '-w', os.path.join(HERE, 'host_column.hip'), '-o', out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return ctypes.CDLL(out)


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def host_cape(L, p, t, td, ps, ts, tds, source, adiabat, warm, p1d=None, start=None,
              level_major=False, pinc=500., levels=False, ml_depth=500.):
    ncol, nk = t.shape
    if level_major:
        t, td = np.ascontiguousarray(t.T), np.ascontiguousarray(td.T)
        if p is not None:
            p = np.ascontiguousarray(p.T)
    out = [np.zeros(ncol, np.float32), np.zeros(ncol, np.float32), np.zeros(ncol, np.int32),
           np.zeros(ncol, np.float32)]
    lv = [np.zeros(ncol, np.float32), np.zeros(ncol, np.float32)] if levels else [None, None]
    st = np.zeros(6, np.int64)
    L.host_cape(_p(p), _p(p1d), _p(start), _p(t), _p(td), _p(ps), _p(ts), _p(tds),
                ctypes.c_int(nk), ctypes.c_int(ncol), ctypes.c_int(int(level_major)),
                ctypes.c_float(pinc), ctypes.c_int(source), ctypes.c_float(ml_depth),
                ctypes.c_int(adiabat), ctypes.c_int(int(warm)), *[_p(o) for o in out + lv],
                _p(st))
    if levels:
        out = out + lv
    return out, dict(zip(('levels', 'substeps', 'iterations', 'nonconverged', 'fallbacks',
                          'first'), st.tolist()))


@pytest.mark.parametrize('nlev,seed', [(40, 1000), (137, 1004)])
@pytest.mark.parametrize('source', [1, 2, 3])
@pytest.mark.parametrize('adiabat', [1, 2, 3, 4])
def test_sigma_plain_bitexact_warm_within_tolerance(host_lib, nlev, seed, source, adiabat):
    c = S.sigma_columns(300, nlev, seed)
    arr = [c[k] for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
    ref = O.cape(*arr, 500., source, 500., adiabat)
    plain, sp = host_cape(host_lib, *arr, source, adiabat, False, level_major=(seed % 2 == 0))
    for g, r in zip(plain, ref):
        np.testing.assert_array_equal(g, r)
    warm, sw = host_cape(host_lib, *arr, source, adiabat, True)
    assert cape_ok(warm[0], ref[0]).all() and cape_ok(warm[1], ref[1]).all()
    np.testing.assert_array_equal(warm[2], ref[2])
    assert sw['substeps'] == sp['substeps'] and sw['levels'] == sp['levels']
    assert sw['iterations'] < 0.35 * sp['iterations']
    assert sw['fallbacks'] < 0.1 * sw['substeps']
    # no history for the first sub-step: without a surrogate it starts cold
    assert sw['first'] == sp['first']


def test_warm_start_iteration_budget(host_lib):
    """What the warm start buys on the combined workload's recipe (137 sigma
    levels, most unstable, pseudo-ice), without a surrogate: the reference
    method takes about 9.4 iterations per sub-step; the predicted first
    iterate, the secant second iterate and the liquid and ice shapes bring it
    to about 1.45, with under 2 % of the sub-steps falling back to the cold
    start.  Bounds with some margin, so that a lost piece fails here."""
    c = S.sigma_columns(400, 137, 4004)
    arr = [c[k] for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
    _, sp = host_cape(host_lib, *arr, 2, 3, False)
    warm, sw = host_cape(host_lib, *arr, 2, 3, True)
    ref = O.cape(*arr, 500., 2, 500., 3)
    assert cape_ok(warm[0], ref[0]).all() and cape_ok(warm[1], ref[1]).all()
    per_sub = sw['iterations'] / sw['substeps']
    print('iterations per sub-step: plain', sp['iterations'] / sp['substeps'], 'warm', per_sub,
          'fallbacks', sw['fallbacks'] / sw['substeps'], flush=True)
    assert per_sub < 1.55
    assert sw['iterations'] < 0.17 * sp['iterations']
    assert sw['fallbacks'] < 0.02 * sw['substeps']


@pytest.mark.parametrize('source', [1, 2, 3])
def test_pressure_levels_plain_bitexact(host_lib, source):
    c = S.pressure_columns(400, 77)
    p1d = np.asarray(c['p'], np.float32)
    start = O.start_levels(p1d, c['ps'])
    arr = [c[k] for k in ('t', 'td', 'ps', 'ts', 'tds')]
    ref = O.cape(None, *arr, 500., source, 500., 1, p1d=p1d, start_lev=start)
    for lm in (False, True):
        got, _ = host_cape(host_lib, None, *arr, source, 1, False, p1d=p1d, start=start,
                           level_major=lm)
        for g, r in zip(got, ref):
            np.testing.assert_array_equal(g, r)


@pytest.mark.parametrize('source', [1, 2, 3])
@pytest.mark.parametrize('adiabat', [1, 2, 3, 4])
def test_lfc_el_plain_bitexact_warm_close(host_lib, source, adiabat):
    """The device column code's LFC/EL (host build): the plain method equals
    the oracle's bit for bit; the warm start, whose fixed points differ within
    the solver tolerance, moves them by little (crossings interpolate between
    levels, so a buoyancy change of ~1e-5 m/s^2 shifts them by a fraction of
    the layer where the layer's buoyancy change is small)."""
    c = S.sigma_columns(400, 60, 51)
    arr = [c[k] for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
    ref = O.cape(*arr, 500., source, 500., adiabat, levels=True)
    plain, _ = host_cape(host_lib, *arr, source, adiabat, False, levels=True)
    for g, r in zip(plain, ref):
        np.testing.assert_array_equal(g, r)
    warm, _ = host_cape(host_lib, *arr, source, adiabat, True, levels=True)
    pos = ref[0] > 0
    np.testing.assert_array_equal(np.isfinite(warm[4]), np.isfinite(ref[4]))
    for w, r in ((warm[4], ref[4]), (warm[5], ref[5])):
        d = np.abs(w[pos] - r[pos])
        assert np.mean(d <= 1.0) >= 0.99, np.sort(d)[-5:]


@pytest.mark.parametrize('adiabat', [1, 2, 3, 4])
def test_edge_branches(host_lib, adiabat):
    """The branches only the oracle had seen (tests/edge_cases.py): ML with
    ml_depth 1 m and 1e6 m, MU with a surface above 500 hPa, the cold-surface
    gate, NaN levels and the 100-iteration limit.  The plain method equals
    the oracle bit for bit (NaN included) with the same count of columns
    stopped by the limit.  The warm start stops the same columns (after its
    fallback the sub-step is the reference's cold solve), zeroes them, and
    agrees elsewhere within the CAPE/CIN tolerance."""
    import edge_cases
    arr = edge_cases.grid()
    for src, ml in edge_cases.CASES:
        ref = O.cape(*arr, 500., src, ml, adiabat, stats=True)
        st_ref = ref[-1]
        plain, sp = host_cape(host_lib, *arr, src, adiabat, False, ml_depth=ml)
        for g, r in zip(plain, ref[:4]):
            np.testing.assert_array_equal(g, r)
        assert sp['nonconverged'] == st_ref['nonconverged'], (src, ml)
        warm, sw = host_cape(host_lib, *arr, src, adiabat, True, ml_depth=ml)
        assert sw['nonconverged'] == st_ref['nonconverged'], (src, ml)
        np.testing.assert_array_equal(warm[2], ref[2])
        nan = np.isnan(ref[0])
        np.testing.assert_array_equal(np.isnan(warm[0]), nan)
        ok = ~nan
        assert cape_ok(warm[0][ok], ref[0][ok]).all() and cape_ok(warm[1][ok], ref[1][ok]).all()
        stopped = (ref[0] == 0) & (ref[1] == 0)
        assert np.all(warm[0][stopped] == 0) and np.all(warm[1][stopped] == 0)
