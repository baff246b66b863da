"""Host-side behaviour of the calc_cape/calc_srh front-end and the C-ABI
library, without a GPU: shapes and return counts (the reference's fake
backend), argument errors, pressure start levels, and that libxcape_hip.so
loads and exports every symbol include/xcape_hip.h declares."""
import ctypes
import os
import re

import numpy as np
import pytest

from xcape_amd import _hip, core

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope='module')
def p_t_td_3d():
    rng = np.random.default_rng(0)
    return tuple(rng.random((5, 10, 20)) for _ in range(3))


@pytest.fixture(scope='module')
def p_t_td_surface():
    rng = np.random.default_rng(1)
    return tuple(rng.random((5, 10)) for _ in range(3))


# reference test_calc_cape_shape_3d (test/test_core.py:39-58), numpy inputs
@pytest.mark.parametrize('sourcein,n_returns', [('surface', 2), ('most-unstable', 4),
                                                ('mixed-layer', 2)])
@pytest.mark.parametrize('vertical_levin', ['sigma', 'pressure'])
def test_calc_cape_shape_3d(p_t_td_3d, p_t_td_surface, sourcein, n_returns, vertical_levin):
    p, t, td = p_t_td_3d
    ps, ts, tds = p_t_td_surface
    if vertical_levin == 'sigma':
        args = (p, t, td, ps, ts, tds)
    else:
        args = (np.ones(t.shape[-1]), t, td, ps, ts, tds)
    result = core.calc_cape(*args, source=sourcein, vertical_lev=vertical_levin, method='dummy')
    assert len(result) == n_returns
    for data in result:
        assert data.shape == p.shape[:-1]


def test_calc_cape_single_profile_shape():
    p = np.linspace(1000, 100, 20)
    out = core.calc_cape(p, p, p, np.float64(1000.), np.float64(20.), np.float64(10.),
                         vertical_lev='sigma', method='dummy')
    assert all(o.shape == (1,) for o in out)


def test_dask_shapes(p_t_td_3d, p_t_td_surface):
    da = pytest.importorskip('dask.array')
    p, t, td = p_t_td_3d
    ps, ts, tds = p_t_td_surface
    args = [da.from_array(a) for a in (p, t, td, ps, ts, tds)]
    res = core.calc_cape(*args, source='most-unstable', vertical_lev='sigma', method='dummy')
    assert len(res) == 4
    for r in res:
        assert isinstance(r, da.Array)
        assert r.compute().shape == p.shape[:-1]


def test_argument_errors(p_t_td_3d, p_t_td_surface):
    p, t, td = p_t_td_3d
    ps, ts, tds = p_t_td_surface
    with pytest.raises(ValueError, match='Too few'):
        core.calc_cape(p, t, td, ps, ts, vertical_lev='sigma')
    with pytest.raises(ValueError, match='Too many'):
        core.calc_cape(p, t, td, ps, ts, tds, tds, vertical_lev='sigma')
    with pytest.raises(KeyError):
        core.calc_cape(p, t, td, ps, ts, tds)
    with pytest.raises(ValueError, match='vertical_lev'):
        core.calc_cape(p, t, td, ps, ts, tds, vertical_lev='hybrid')
    with pytest.raises(ValueError, match='P should be 1d'):
        core.calc_cape(p, t, td, ps, ts, tds, vertical_lev='pressure', method='dummy')
    with pytest.raises(ValueError, match='invalid method'):
        core.calc_cape(p, t, td, ps, ts, tds, vertical_lev='sigma', method='numba')
    with pytest.raises(KeyError):
        core.calc_cape(p, t, td, ps, ts, tds, vertical_lev='sigma', source='lifted',
                       method='dummy')
    with pytest.raises(ValueError, match='same shape'):
        core.calc_cape(p, t[..., :-1], td, ps, ts, tds, vertical_lev='sigma', method='dummy')


def test_pressure_start_levels_match_fixture(era5):
    st = core.pressure_start_levels(era5['level'], era5['sfc_p'])
    np.testing.assert_array_equal(st - 1, era5['sfc_start3d'].astype(np.int32))


def test_calc_srh_shapes_dummy():
    rng = np.random.default_rng(2)
    a3 = [rng.random((4, 6, 30)) for _ in range(5)]
    a2 = [rng.random((4, 6)) for _ in range(5)]
    r = core.calc_srh(*a3, *a2, vertical_lev='sigma', output_var='all', method='dummy')
    assert len(r) == 8 and all(x.shape == (4, 6) for x in r)
    r = core.calc_srh(*a3, *a2, vertical_lev='sigma', output_var='srh', method='dummy')
    assert len(r) == 2 and all(x.shape == (4, 6) for x in r)


# ------------------------------------------------------------- C-ABI ------
def _declared_symbols():
    txt = open(os.path.join(ROOT, 'include', 'xcape_hip.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(xcape_[a-z0-9_]+)\s*\(', txt)))


def test_library_exports_every_declared_symbol():
    L = _hip.lib()
    names = _declared_symbols()
    assert len(names) >= 5
    for n in names:
        assert hasattr(L, n), n
    # the binding declares exactly the header's entry points
    assert sorted(_hip.SIGNATURES) == names


def test_invalid_arguments_rejected_without_device():
    """Argument checks run before any device work and report EINVAL."""
    L = _hip.lib()
    x = np.zeros((2, 3), np.float32)
    s = np.zeros(2, np.float32)
    o = np.zeros(2, np.float32)
    m = np.zeros(2, np.int32)
    P = _hip._ptr

    def call(nk=3, pinc=500., source=1, adiabat=1):
        return L.xcape_cape_f32(P(x), 0, None, P(x), P(x), P(s), P(s), P(s), nk, 2, pinc, source,
                                500., adiabat, P(o), P(o), P(m), P(o), 0, None, None)
    assert call(source=7) == -1 and b'source' in L.xcape_last_error()
    assert call(adiabat=0) == -1 and b'adiabat' in L.xcape_last_error()
    assert call(pinc=0.0) == -1 and b'pinc' in L.xcape_last_error()
    assert call(pinc=float('nan')) == -1
    assert call(nk=0) == -1
    # per-column p with start levels (loopcape_pl): start levels are checked
    # against 1..nk, and the helicity chain does not take them
    st = np.array([1, 4], np.int32)
    assert L.xcape_cape_f32(P(x), 0, P(st), P(x), P(x), P(s), P(s), P(s), 3, 2, 500., 1,
                            500., 1, P(o), P(o), P(m), P(o), 0, None, None) == -1
    assert b'start_lev' in L.xcape_last_error()
    d = np.zeros(2, np.float64)
    w = np.zeros(4, np.float32)
    assert L.xcape_srh_f32(P(x), 0, P(np.ones(2, np.int32)), P(x), P(x), P(x), P(x), P(s),
                           P(s), P(s), P(s), P(s), 3, 2, 3000., P(d), P(d), P(w), P(w), P(w),
                           0, None) == -1
    assert b'CAPE calls only' in L.xcape_last_error()
    # ragged plans: columns per workgroup from the tuner's set, 64 or more
    # for warm-started plans
    off = np.array([0, 3, 5], np.int64)
    one = np.ones(2, np.int32)
    p5 = np.zeros(5, np.float32)
    h = ctypes.c_void_p()

    def plan(cols, flags=0):
        return L.xcape_ragged_plan_create(P(off), P(one), P(one), 2, P(off), P(one), P(one),
                                          P(p5), P(p5), P(p5), P(s), P(s), P(s), 500., 500.,
                                          P(o), P(o), P(m), P(o), 64, 0, cols, flags,
                                          ctypes.byref(h))
    assert plan(48) == -1 and b'cols_per_block' in L.xcape_last_error()
    assert plan(16, _hip.XCAPE_WARM_START) == -1 and b'warm' in L.xcape_last_error()
    assert L.xcape_version().startswith(b'xcape-hip')


def test_levels_keyword_appends_two_outputs(p_t_td_3d, p_t_td_surface):
    """levels=True (not in the reference) appends p_lfc, p_el shaped like the
    other outputs; the fake backend gives NaN levels."""
    p, t, td = p_t_td_3d
    ps, ts, tds = p_t_td_surface
    res = core.calc_cape(p, t, td, ps, ts, tds, source='most-unstable', vertical_lev='sigma',
                         method='dummy', levels=True)
    assert len(res) == 6
    assert all(r.shape == p.shape[:-1] for r in res)
    assert np.all(np.isnan(res[4])) and np.all(np.isnan(res[5]))
    res = core.calc_cape(p, t, td, ps, ts, tds, source='surface', vertical_lev='sigma',
                         method='dummy', levels=True)
    assert len(res) == 4


def test_memory_mapped_profiles_reach_the_library_uncopied(tmp_path):
    """Out-of-core grids: a memory-mapped float32 (X.., Z) array reaches the
    C-ABI as a view of the mapping (reshaped, never copied into memory); the
    library's host threads read it chunk by chunk as they stage it."""
    m = np.lib.format.open_memmap(tmp_path / 't.npy', mode='w+', dtype=np.float32,
                                  shape=(30, 40, 50))
    m[:] = 1.0
    m.flush()
    r = np.load(tmp_path / 't.npy', mmap_mode='r')
    (flat,) = core._reshape_inputs(r)
    passed = _hip._Args([flat]).f32(flat)
    assert passed.shape == (1200, 50) and passed.flags['C_CONTIGUOUS']
    assert np.shares_memory(passed, r)
