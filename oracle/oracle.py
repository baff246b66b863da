"""ctypes front-end of the C oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker; the product
package never imports it.

All functions take the reference's Python-level layout: profiles ``(ncol,
nk)`` C-contiguous (levels last, i.e. the transpose of the reference's
``(nk, ncol)`` Fortran-order view), surface arrays ``(ncol,)``, hPa and degC.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        'columns', 'levels', 'substeps', 'iterations', 'iterations_ice',
        'nonconverged', 'mu_levels', 'env_levels')]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


def build():
    subprocess.run(['make', '-s', '-C', _HERE, 'liboracle.so'], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'liboracle.so')
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


def start_levels(p1d, ps):
    """1-based first pressure level at or above the ground, per column.

    Same rule as the reference's ``core.py:286-289``:
    ``masked_less(ps - p, 0).argmin(axis=0) + 1``.
    """
    p1d = np.asarray(p1d)
    ps = np.asarray(ps).reshape(-1)
    tmp = ps[None, :] - p1d[:, None]
    return (np.ma.masked_less(tmp, 0).argmin(axis=0) + 1).astype(np.int32)


def cape(p, t, td, ps, ts, tds, pinc=500., source=1, ml_depth=500., adiabat=1,
         p1d=None, start_lev=None, fp64=False, stats=False, levels=False, perturb=None,
         perturb_iterate=0.0):
    """Column CAPE/CIN/MUlvl/z_out.  ``p`` is (ncol, nk) (ignored when ``p1d``
    gives shared pressure levels, in which case ``start_lev`` is required).
    ``p`` with ``start_lev``: each column from its start level of its own p
    (loopcape_pl).  With ``levels`` the LFC and EL pressures [hPa] follow
    z_out (NaN without positive area).  ``perturb``: a seed; every exp, log
    and pow result is then kept or moved by one ulp (xo_cape_pert32), as a
    C library an ulp away from this one would give them (FP32, no levels);
    with ``perturb_iterate`` > 0 every converged sub-step also accepts
    another point of its tolerance band (the last iterate moved by up to that
    many K, kept when it passes the same convergence test)."""
    dt = np.float64 if fp64 else np.float32
    t = _c(t, dt)
    ncol, nk = t.shape
    td = _c(td, dt)
    ps, ts, tds = (_c(np.reshape(a, -1), dt) for a in (ps, ts, tds))
    if p1d is not None:
        p1d = _c(p1d, dt)
        start_lev = _c(start_lev, np.int32)
        p = None
    else:
        p = _c(p, dt)
        if start_lev is not None:
            start_lev = _c(start_lev, np.int32)
    cape_o = np.zeros(ncol, dt)
    cin_o = np.zeros(ncol, dt)
    mul = np.zeros(ncol, np.int32)
    zo = np.zeros(ncol, dt)
    st = Stats()
    lfc = np.zeros(ncol, dt)
    el = np.zeros(ncol, dt)
    R = ctypes.c_double if fp64 else ctypes.c_float
    if perturb is not None:
        if fp64 or levels:
            raise ValueError('perturbed runs are FP32 without levels')
        lib().xo_set_perturbation(ctypes.c_uint32(perturb))
        lib().xo_set_iterate_perturbation(ctypes.c_float(perturb_iterate))
        fn = lib().xo_cape_pert32
        fn.restype = None
        fn(_p(p), _p(p1d), _p(t), _p(td), _p(ps), _p(ts), _p(tds), _p(start_lev),
           ctypes.c_int64(nk), ctypes.c_int64(ncol), R(pinc), ctypes.c_int(source),
           R(ml_depth), ctypes.c_int(adiabat), _p(cape_o), _p(cin_o), _p(mul), _p(zo),
           ctypes.byref(st))
    else:
        fn = lib().xo_cape_levels_f64 if fp64 else lib().xo_cape_levels_f32
        fn.restype = None
        fn(_p(p), _p(p1d), _p(t), _p(td), _p(ps), _p(ts), _p(tds), _p(start_lev),
           ctypes.c_int64(nk), ctypes.c_int64(ncol), R(pinc), ctypes.c_int(source),
           R(ml_depth), ctypes.c_int(adiabat), _p(cape_o), _p(cin_o), _p(mul), _p(zo),
           _p(lfc) if levels else None, _p(el) if levels else None, ctypes.byref(st))
    out = (cape_o, cin_o, mul, zo) + ((lfc, el) if levels else ())
    if stats:
        return out + (st.as_dict(),)
    return out


def cape_log(p, t, td, ps, ts, tds, pinc=500., source=1, ml_depth=500., adiabat=1,
             cap=1 << 20):
    """Solver samples, one row per sub-step: p2 [Pa], theta-e(start), qt,
    th1, converged th2, iterations."""
    t = _c(t, np.float32)
    ncol, nk = t.shape
    p, td = _c(p, np.float32), _c(td, np.float32)
    ps, ts, tds = (_c(np.reshape(a, -1), np.float32) for a in (ps, ts, tds))
    rec = np.zeros((cap, 6), np.float32)
    fn = lib().xo_cape_log_f32
    fn.restype = ctypes.c_int64
    n = fn(_p(p), _p(t), _p(td), _p(ps), _p(ts), _p(tds), ctypes.c_int64(nk),
           ctypes.c_int64(ncol), ctypes.c_float(pinc), ctypes.c_int(source),
           ctypes.c_float(ml_depth), ctypes.c_int(adiabat), _p(rec), ctypes.c_int64(cap))
    return rec[:n]


def stdheight(p, t, td, ps, ts, tds, aglh0=2., p1d=None, start_lev=None):
    t = _c(t, np.float64)
    ncol, nk = t.shape
    td = _c(td, np.float64)
    ps, ts, tds = (_c(np.reshape(a, -1), np.float64) for a in (ps, ts, tds))
    if p1d is not None:
        p1d, start_lev, p = _c(p1d, np.float64), _c(start_lev, np.int32), None
    else:
        p = _c(p, np.float64)
    h = np.zeros((ncol, nk), np.float64)
    hs = np.zeros(ncol, np.float64)
    fn = lib().xo_stdheight
    fn.restype = None
    fn(_p(p), _p(p1d), _p(t), _p(td), _p(ps), _p(ts), _p(tds), _p(start_lev),
       ctypes.c_int64(nk), ctypes.c_int64(ncol), ctypes.c_double(aglh0), _p(h), _p(hs))
    return h, hs


def bunkers(u, v, aglh, us, vs, aglhs, start_lev=None):
    """rm, lm, mean6: (ncol, 2) float32."""
    u, v, aglh = (_c(a, np.float32) for a in (u, v, aglh))
    ncol, nk = u.shape
    us, vs, aglhs = (_c(np.reshape(a, -1), np.float32) for a in (us, vs, aglhs))
    st = _c(start_lev, np.int32)
    rm = np.zeros((ncol, 2), np.float32)
    lm = np.zeros((ncol, 2), np.float32)
    mw = np.zeros((ncol, 2), np.float32)
    fn = lib().xo_bunkers
    fn.restype = None
    fn(_p(u), _p(v), _p(aglh), _p(us), _p(vs), _p(aglhs), _p(st),
       ctypes.c_int64(nk), ctypes.c_int64(ncol), _p(rm), _p(lm), _p(mw))
    return rm, lm, mw


def sreh(u, v, aglh, us, vs, aglhs, cu_rm, cv_rm, cu_lm, cv_lm, top, start_lev=None):
    u, v, aglh = (_c(a, np.float64) for a in (u, v, aglh))
    ncol, nk = u.shape
    us, vs, aglhs = (_c(np.reshape(a, -1), np.float64) for a in (us, vs, aglhs))
    cs = [_c(np.reshape(a, -1), np.float64) for a in (cu_rm, cv_rm, cu_lm, cv_lm)]
    st = _c(start_lev, np.int32)
    srm = np.zeros(ncol, np.float64)
    slm = np.zeros(ncol, np.float64)
    fn = lib().xo_sreh
    fn.restype = None
    fn(_p(u), _p(v), _p(aglh), _p(us), _p(vs), _p(aglhs), *[_p(c) for c in cs],
       ctypes.c_double(top), _p(st), ctypes.c_int64(nk), ctypes.c_int64(ncol), _p(srm), _p(slm))
    return srm, slm


def srh(p, t, td, u, v, ps, ts, tds, us, vs, depth=3000., p1d=None, start_lev=None):
    """The reference's calc_srh chain (core.py:516-541, srh.py:41-61): FP64
    stdheight with a 2 m surface height, FP32 Bunkers on the heights cast to
    f32, FP64 SREH on the FP64 heights.  Returns srh_rm, srh_lm (f64) and rm,
    lm, mean6 ((ncol, 2) f32)."""
    h, hs = stdheight(p, t, td, ps, ts, tds, 2., p1d=p1d, start_lev=start_lev)
    rm, lm, mw = bunkers(u, v, h.astype(np.float32), us, vs, hs.astype(np.float32),
                         start_lev=start_lev)
    srm, slm = sreh(u, v, h, us, vs, hs, rm[:, 0], rm[:, 1], lm[:, 0], lm[:, 1], depth,
                    start_lev=start_lev)
    return srm, slm, rm, lm, mw
