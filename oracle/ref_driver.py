"""ctypes driver for the reference Fortran compiled by ``make -C oracle ref``.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path imports this module;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may.  It calls the reference routines the way the reference's
f2py wrappers do (``src/xcape/fortran/*.pyf``): every input is cast to the
declared kind and made Fortran-contiguous, every ``intent(out)`` array is
freshly allocated and zero-filled (numpy ``fortranobject.c``), and scalars are
passed by reference.

The libraries live in ``oracle/_ref/`` (git-ignored, built from the reference
sources).  ``available()`` says whether they were built.
"""
import ctypes
import os

import numpy as np

_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), '_ref')
_libs = {}


def _lib(name):
    if name not in _libs:
        _libs[name] = ctypes.CDLL(os.path.join(_DIR, f'lib{name}.so'))
    return _libs[name]


def available():
    return os.path.exists(os.path.join(_DIR, 'libCAPE_CODE_model_lev.so'))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float32))


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _i32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.int32))


def _rf(x):
    return ctypes.byref(ctypes.c_float(x))


def _ri(x):
    return ctypes.byref(ctypes.c_int(int(x)))


def _rd(x):
    return ctypes.byref(ctypes.c_double(x))


def loopcape_ml(p, t, td, ps, ts, tds, pinc, source, ml_depth, adiabat):
    """Reference ``loopcape_ml`` (CAPE_CODE_model_lev.f90:4-91).

    p, t, td: (nk, n2) hPa / degC; ps, ts, tds: (n2,).
    Returns cape, cin, mulvl (int32), z_out.
    """
    p, t, td = _f32(p), _f32(t), _f32(td)
    ps, ts, tds = _f32(ps), _f32(ts), _f32(tds)
    nk, n2 = t.shape
    cape = np.zeros(n2, np.float32)
    cin = np.zeros(n2, np.float32)
    mul = np.zeros(n2, np.int32)
    zo = np.zeros(n2, np.float32)
    _lib('CAPE_CODE_model_lev').loopcape_ml_(
        _p(p), _p(t), _p(td), _p(ps), _p(ts), _p(tds),
        _rf(pinc), _ri(source), _rf(ml_depth), _ri(adiabat),
        _ri(nk), _ri(n2), _p(cape), _p(cin), _p(mul), _p(zo))
    return cape, cin, mul, zo


def loopcape_pl(p, t, td, ps, ts, tds, pinc, source, ml_depth, adiabat, start_3d):
    """Reference ``loopcape_pl`` (CAPE_CODE_pressure_lev.f90:4-84): pressure
    per column, each column from its own 1-based start level.

    p, t, td: (nk, n2) hPa / degC; ps, ts, tds: (n2,); start_3d: (n2,).
    """
    p, t, td = _f32(p), _f32(t), _f32(td)
    ps, ts, tds = _f32(ps), _f32(ts), _f32(tds)
    st = _i32(start_3d)
    nk, n2 = t.shape
    cape = np.zeros(n2, np.float32)
    cin = np.zeros(n2, np.float32)
    mul = np.zeros(n2, np.int32)
    zo = np.zeros(n2, np.float32)
    _lib('CAPE_CODE_pressure_lev').loopcape_pl_(
        _p(p), _p(t), _p(td), _p(ps), _p(ts), _p(tds),
        _rf(pinc), _ri(source), _rf(ml_depth), _ri(adiabat), _p(st),
        _ri(nk), _ri(n2), _p(cape), _p(cin), _p(mul), _p(zo))
    return cape, cin, mul, zo


def loopcape_pl1d(t, td, p1d, ps, ts, tds, pinc, source, ml_depth, adiabat, start_3d):
    """Reference ``loopcape_pl1d`` (CAPE_CODE_pressure_lev.f90:88-169).

    t, td: (nk, n2); p1d: (nk,) shared pressure levels; start_3d: (n2,)
    1-based first level used per column.
    """
    t, td = _f32(t), _f32(td)
    p = _f32(np.asarray(p1d).reshape(-1, 1))
    ps, ts, tds = _f32(ps), _f32(ts), _f32(tds)
    st = _i32(start_3d)
    nk, n2 = t.shape
    cape = np.zeros(n2, np.float32)
    cin = np.zeros(n2, np.float32)
    mul = np.zeros(n2, np.int32)
    zo = np.zeros(n2, np.float32)
    _lib('CAPE_CODE_pressure_lev').loopcape_pl1d_(
        _p(t), _p(td), _p(p), _p(ps), _p(ts), _p(tds),
        _rf(pinc), _ri(source), _rf(ml_depth), _ri(adiabat), _p(st),
        _ri(nk), _ri(n2), _p(cape), _p(cin), _p(mul), _p(zo))
    return cape, cin, mul, zo


def stdheight_ml(p, t, td, ps, ts, tds, aglh0):
    """Reference ``loop_stdheight_ml`` (stdheight_2D_model_lev.f90:4-34), FP64."""
    p, t, td = _f64(p), _f64(t), _f64(td)
    ps, ts, tds = _f64(ps), _f64(ts), _f64(tds)
    nk, n2 = t.shape
    hin = _f64(np.broadcast_to(np.asarray(aglh0, np.float64), (n2,)))
    h = np.zeros((nk, n2), np.float64, order='F')
    hs = np.zeros(n2, np.float64)
    _lib('stdheight_2D_model_lev').loop_stdheight_ml_(
        _p(p), _p(t), _p(td), _p(ps), _p(ts), _p(tds), _p(hin),
        _ri(nk), _ri(n2), _p(h), _p(hs))
    return h, hs


def stdheight_pl1d(t, td, p1d, ps, ts, tds, aglh0, start_3d):
    """Reference ``loop_stdheight_pl1d`` (stdheight_2D_pressure_lev.f90:229-257).

    ``start_3d`` is declared DOUBLE PRECISION there, so it is passed as f64.
    """
    t, td = _f64(t), _f64(td)
    p = _f64(np.asarray(p1d).reshape(-1, 1))
    ps, ts, tds = _f64(ps), _f64(ts), _f64(tds)
    nk, n2 = t.shape
    hin = _f64(np.broadcast_to(np.asarray(aglh0, np.float64), (n2,)))
    st = _f64(start_3d)
    h = np.zeros((nk, n2), np.float64, order='F')
    hs = np.zeros(n2, np.float64)
    _lib('stdheight_2D_pressure_lev').loop_stdheight_pl1d_(
        _p(t), _p(td), _p(p), _p(ps), _p(ts), _p(tds), _p(hin), _p(st),
        _ri(nk), _ri(n2), _p(h), _p(hs))
    return h, hs


def _fresh_call(name, fn, *args):
    """Call ``fn`` in a freshly loaded copy of ``lib<name>.so`` and unload it.

    Used with the ``-fno-automatic`` builds: static locals start at zero on
    every load, so each call sees zero-initialised accumulators.
    """
    import _ctypes
    lib = ctypes.CDLL(os.path.join(_DIR, f'lib{name}.so'), mode=ctypes.RTLD_LOCAL)
    try:
        getattr(lib, fn)(*args)
    finally:
        _ctypes.dlclose(lib._handle)


def _bunkers(module, fn, u, v, aglh, us, vs, aglhs, start_3d=None):
    u, v, aglh = _f32(u), _f32(v), _f32(aglh)
    us, vs, aglhs = _f32(us), _f32(vs), _f32(aglhs)
    nk, n2 = u.shape
    rm = np.zeros((2, n2), np.float32, order='F')
    lm = np.zeros((2, n2), np.float32, order='F')
    mw = np.zeros((2, n2), np.float32, order='F')
    for i in range(n2):
        cols = [np.asfortranarray(a[:, i:i + 1]) for a in (u, v, aglh)]
        sfc = [np.ascontiguousarray(a[i:i + 1]) for a in (us, vs, aglhs)]
        out = [np.zeros((2, 1), np.float32, order='F') for _ in range(3)]
        args = [_p(a) for a in cols] + [_p(a) for a in sfc]
        if start_3d is not None:
            st = _f32(np.asarray(start_3d)[i:i + 1])
            args.append(_p(st))
        args += [_ri(nk), _ri(1)] + [_p(a) for a in out]
        _fresh_call(module, fn, *args)
        rm[:, i], lm[:, i], mw[:, i] = out[0][:, 0], out[1][:, 0], out[2][:, 0]
    return rm, lm, mw


def bunkers_ml(u, v, aglh, us, vs, aglhs):
    """Reference ``bunkers_loop_ml`` (Bunkers_model_lev.f90:4-69), FP32.

    Runs the ``-fno-automatic`` build one column per fresh load, so the
    accumulators read before assignment at Bunkers_model_lev.f90:144-163
    start at zero.
    """
    return _bunkers('Bunkers_model_lev_zinit', 'bunkers_loop_ml_', u, v, aglh, us, vs, aglhs)


def bunkers_pl(u, v, aglh, us, vs, aglhs, start_3d):
    """Reference ``bunkers_loop_pl`` (Bunkers_pressure_lev.f90:4-78).

    ``start_3d`` is declared REAL there, so it is passed as f32.  Same
    zero-initialisation treatment as ``bunkers_ml``.
    """
    return _bunkers('Bunkers_pressure_lev_zinit', 'bunkers_loop_pl_', u, v, aglh, us, vs, aglhs,
                    start_3d=start_3d)


def sreh_ml(u, v, aglh, us, vs, aglhs, cu_rm, cv_rm, cu_lm, cv_lm, top):
    """Reference ``loop_sreh_ml`` (SREH_model_lev.f90:4-55), FP64."""
    u, v, aglh = _f64(u), _f64(v), _f64(aglh)
    us, vs, aglhs = _f64(us), _f64(vs), _f64(aglhs)
    cu_rm, cv_rm, cu_lm, cv_lm = (_f64(a) for a in (cu_rm, cv_rm, cu_lm, cv_lm))
    nk, n2 = u.shape
    srm = np.zeros(n2, np.float64)
    slm = np.zeros(n2, np.float64)
    _lib('SREH_model_lev').loop_sreh_ml_(
        _p(u), _p(v), _p(aglh), _p(us), _p(vs), _p(aglhs),
        _p(cu_rm), _p(cv_rm), _p(cu_lm), _p(cv_lm), _rd(top),
        _ri(nk), _ri(n2), _p(srm), _p(slm))
    return srm, slm


def sreh_pl(u, v, aglh, us, vs, aglhs, cu_rm, cv_rm, cu_lm, cv_lm, top, start_3d):
    """Reference ``loop_sreh_pl`` (SREH_pressure_lev.f90:4-63), FP64.

    ``start_3d`` is declared REAL(KIND=8) there.
    """
    u, v, aglh = _f64(u), _f64(v), _f64(aglh)
    us, vs, aglhs = _f64(us), _f64(vs), _f64(aglhs)
    cu_rm, cv_rm, cu_lm, cv_lm = (_f64(a) for a in (cu_rm, cv_rm, cu_lm, cv_lm))
    st = _f64(start_3d)
    nk, n2 = u.shape
    srm = np.zeros(n2, np.float64)
    slm = np.zeros(n2, np.float64)
    _lib('SREH_pressure_lev').loop_sreh_pl_(
        _p(u), _p(v), _p(aglh), _p(us), _p(vs), _p(aglhs),
        _p(cu_rm), _p(cv_rm), _p(cu_lm), _p(cv_lm), _rd(top), _p(st),
        _ri(nk), _ri(n2), _p(srm), _p(slm))
    return srm, slm


def bunkers_ml_timing(u, v, aglh, us, vs, aglhs):
    """``bunkers_loop_ml`` from the default build in one call, as f2py would
    run it: same work, but the accumulators the reference reads before
    setting are whatever the stack holds, so the values are not meaningful.
    Used only to time the reference's CPU path."""
    u, v, aglh = _f32(u), _f32(v), _f32(aglh)
    us, vs, aglhs = _f32(us), _f32(vs), _f32(aglhs)
    nk, n2 = u.shape
    rm = np.zeros((2, n2), np.float32, order='F')
    lm = np.zeros((2, n2), np.float32, order='F')
    mw = np.zeros((2, n2), np.float32, order='F')
    _lib('Bunkers_model_lev').bunkers_loop_ml_(
        _p(u), _p(v), _p(aglh), _p(us), _p(vs), _p(aglhs),
        _ri(nk), _ri(n2), _p(rm), _p(lm), _p(mw))
    return rm, lm, mw
