"""numpy/dask API for xcape on the gfx950 kernels.

``calc_cape`` and ``calc_srh`` keep the signatures, argument checks, reshaping
and return conventions of the reference's ``xcape.core``
(src/xcape/core.py:130-542) so existing call sites run unchanged.  The
operator switch (core.py:313-324) gains ``method='gpu'``:

* ``'gpu'``   - the HIP kernels in libxcape_hip.so (default);
* ``'fortran'`` - accepted so reference call sites run unchanged; it selects
  the same HIP kernels (this package ships no Fortran);
* ``'dummy'`` - the reference's fake backend (core.py:107-121), ones;
* anything else raises ``ValueError('invalid method')``.

There is no CPU fallback: without the library or a gfx950 device the call
raises.  dask is optional: when dask arrays are passed (and dask is
importable) the work is mapped over column chunks with ``apply_gufunc`` as the
reference does (core.py:237-258, 449-469).
"""
from functools import reduce

import numpy as np

from . import _hip

_SOURCE = {'surface': 1, 'most-unstable': 2, 'mixed-layer': 3}
_ADIABAT = {'pseudo-liquid': 1, 'reversible-liquid': 2, 'pseudo-ice': 3, 'reversible-ice': 4}
_VLEV = {'sigma': 1, 'pressure': 2}
_OUTPUT = {'srh': 1, 'all': 2}
_GPU_METHODS = ('gpu', 'fortran')


def _prod(v):
    return reduce(lambda x, y: x * y, v)


def _reshape_inputs(*args):
    """(X1, .., Xn, Z) -> (prod X, Z) C-contiguous; all shapes must agree
    (core.py:31-50 builds the transposed (Z, prod X) view of the same
    memory)."""
    shape = args[0].shape
    for a in args:
        if a.shape != shape:
            raise ValueError('Input arrays must have the same shape.')
    args2 = [np.atleast_2d(a) for a in args]
    oshape = args2[0].shape
    new_shape = (_prod(oshape[:-1]), oshape[-1])
    return [np.reshape(a, new_shape) for a in args2]


def _reshape_surface_inputs(*args):
    """(X1, .., Xn) -> (prod X,) (core.py:81-100)."""
    shape = args[0].shape
    for a in args:
        if a.shape != shape:
            raise ValueError('Input arrays must have the same shape.')
    args1 = [np.atleast_1d(a) for a in args]
    new_shape = (_prod(args1[0].shape),)
    return [np.reshape(a, new_shape) for a in args1]


def _reshape_outputs(*args, shape=None):
    """(ncol,) -> shape[:-1], or (1,) for a single profile (core.py:52-63)."""
    target = (1,) if len(shape) == 1 else shape[:-1]
    return [np.reshape(a, target) for a in args]


def _reshape_outputs_uv_components(*args, shape=None):
    """(2, ncol) -> (2,) + shape[:-1] (core.py:65-79)."""
    target = (2,) if len(shape) == 1 else (2,) + shape[:-1]
    return [np.reshape(a, target) for a in args]


def _cape_dummy(*args, **kwargs):
    """Fake backend of the reference (core.py:107-121): ones, for shape tests.
    Takes the (ncol, nk) layout used here."""
    p, t, td, ps, ts, tds = args
    assert t.ndim == 2
    n = t.shape[0]
    return tuple(np.ones((1, n)) for _ in range(4))


def _dask():
    try:
        import dask.array as da
        return da
    except ImportError:
        return None


def _any_dask_array(*args):
    da = _dask()
    if da is None:
        return False
    return any(isinstance(a, da.Array) for a in args)


def pressure_start_levels(p1d, ps):
    """1-based first level with p <= ps for each column (core.py:286-289)."""
    tmp = ps[None, :] - p1d[:, None]
    return (np.ma.masked_less(tmp, 0).argmin(axis=0) + 1).astype(np.int32)


def calc_cape(*args, **kwargs):
    """Convective available potential energy and convective inhibition.

    Same contract as the reference's ``xcape.core.calc_cape``
    (src/xcape/core.py:130-235): positional ``p, t, td, ps, ts, tds`` (hPa,
    degC; profiles with levels on the last axis; for
    ``vertical_lev='pressure'`` p is the 1-D level vector), keywords
    ``source`` ('surface' | 'mixed-layer' | 'most-unstable'), ``ml_depth``
    [m], ``adiabat`` ('pseudo-liquid' | 'reversible-liquid' | 'pseudo-ice' |
    'reversible-ice'), ``pinc`` [Pa], ``method`` and ``vertical_lev`` (which,
    as in the reference, must be given).

    Returns ``(cape, cin)``, or ``(cape, cin, mulev, zmulev)`` for
    ``source='most-unstable'``, shaped like the inputs without the level axis.

    Extra keywords (not in the reference): ``fast_math`` selects the hardware
    exp2/log2/reciprocal kernel variant; ``warm_start`` the learned warm start;
    ``levels=True`` appends the level of free convection and the equilibrium
    level, ``p_lfc, p_el`` [hPa] (NaN without positive area), to the outputs.
    """
    if len(args) < 6:
        raise ValueError("Too few arguments.")
    if len(args) > 6:
        raise ValueError("Too many arguments.")
    allowed_vertical_levs = ['sigma', 'pressure']
    if kwargs['vertical_lev'] not in allowed_vertical_levs:
        raise ValueError(f"`vertical_lev` must be one of: {allowed_vertical_levs}")
    if _any_dask_array(*args):
        return _calc_cape_gufunc(*args, **kwargs)
    return _calc_cape_numpy(*args, **kwargs)


def _calc_cape_gufunc(*args, **kwargs):
    """dask path (core.py:237-258).  One scalar input slot per surface
    argument: the reference's pressure signature lists one slot too many
    (core.py:246), which this does not reproduce."""
    import dask.array as da
    signature = "(i),(i),(i),(),(),()->(),()"
    output_dtypes = ('f4', 'f4')
    if kwargs.get('vertical_lev') == 'pressure':
        # p is the shared level vector: pass it whole to every chunk
        p = np.asarray(args[0])
        rest = args[1:]

        def _fn(*a, **kw):
            return _calc_cape_numpy(p, *a, **kw)
        signature = "(i),(i),(),(),()->(),()"
        if kwargs.get('source') == 'most-unstable':
            signature += ",(),()"
            output_dtypes = output_dtypes + ('i4', 'f4')
        if kwargs.get('levels'):
            signature += ",(),()"
            output_dtypes = output_dtypes + ('f4', 'f4')
        return da.apply_gufunc(_fn, signature, *rest, output_dtypes=output_dtypes, axis=-1,
                               vectorize=False, **kwargs)
    if kwargs.get('source') == 'most-unstable':
        signature += ",(),()"
        output_dtypes = output_dtypes + ('i4', 'f4')
    if kwargs.get('levels'):
        signature += ",(),()"
        output_dtypes = output_dtypes + ('f4', 'f4')
    return da.apply_gufunc(_calc_cape_numpy, signature, *args, output_dtypes=output_dtypes,
                           axis=-1, vectorize=False, **kwargs)


def _calc_cape_numpy(*args, source='surface', ml_depth=500., adiabat='pseudo-liquid',
                     pinc=500., method='gpu', vertical_lev='sigma', fast_math=False,
                     warm_start=False, levels=False):
    """numpy path (core.py:261-332)."""
    p, t, td, ps, ts, tds = args
    original_shape = t.shape
    p_s1d, t_s1d, td_s1d = _reshape_surface_inputs(ps, ts, tds)
    start = None
    if len(p.shape) == 1:
        t_2d, td_2d = _reshape_inputs(t, td)
        p_2d = np.asarray(p)
        p_shared = True
        # computed here (as the reference does) when the inputs are not
        # float32, so the comparison happens in the caller's precision;
        # float32 inputs get the same rule on the device
        if np.asarray(p).dtype != np.float32 or np.asarray(p_s1d).dtype != np.float32:
            start = pressure_start_levels(np.asarray(p_2d, np.float64),
                                          np.asarray(p_s1d, np.float64))
    elif (p.shape == t.shape) & (vertical_lev == 'sigma'):
        p_2d, t_2d, td_2d = _reshape_inputs(p, t, td)
        p_shared = False
    elif (p.shape == t.shape) & (vertical_lev == 'pressure'):
        raise ValueError("P should be 1d")
    else:
        raise ValueError('Input arrays must have the same shape.')

    src = _SOURCE[source]
    adi = _ADIABAT[adiabat]
    _VLEV[vertical_lev]

    lv = ()
    if method in _GPU_METHODS:
        out = _hip.cape(
            p_2d, t_2d, td_2d, p_s1d, t_s1d, td_s1d, pinc=float(pinc), source=src,
            ml_depth=float(ml_depth), adiabat=adi, p_shared=p_shared, start_lev=start,
            fast_math=fast_math, warm_start=warm_start, levels=levels)
        cape_2d, cin_2d, mulev, zmulev = out[:4]
        lv = tuple(out[4:])
    elif method == 'dummy':
        cape_2d, cin_2d, mulev, zmulev = _cape_dummy(p_2d, t_2d, td_2d, p_s1d, t_s1d, td_s1d)
        if levels:
            lv = (np.full_like(cape_2d, np.nan), np.full_like(cape_2d, np.nan))
    else:
        raise ValueError('invalid method')

    if src == 2:
        return tuple(_reshape_outputs(cape_2d, cin_2d, mulev, zmulev, *lv,
                                      shape=original_shape))
    return tuple(_reshape_outputs(cape_2d, cin_2d, *lv, shape=original_shape))


def dewpoint_from_specific_humidity(q, p):
    """Dewpoint [degC] of air with specific humidity ``q`` [kg/kg] at
    pressure ``p`` [hPa], for inputs that carry humidity instead of dewpoint.

    Not in the reference (its tutorial expects q inputs, doc/tutorial.rst:
    18-23, but xcape has no converter).  It inverts the saturation formula the
    ascent itself uses for the environment's mixing ratio (getqvs, Bolton
    1980, CAPE_CODE_model_lev.f90:570-581): with r = q / (1 - q) the vapour
    pressure is e = p r / (eps + r), eps = Rd/Rv, and the dewpoint solves
    e = 6.112 exp(17.67 Td / (Td + 243.5)), so that the kernels recover r from
    the returned Td.  Broadcasts like numpy; float64 arithmetic, the result in
    the promoted input dtype (at least float32)."""
    q = np.asarray(q)
    p = np.asarray(p)
    dt = np.result_type(q, p, np.float32)
    eps = 287.04 / 461.5
    r = np.asarray(q, np.float64) / (1.0 - np.asarray(q, np.float64))
    e = np.asarray(p, np.float64) * r / (eps + r)
    x = np.log(e / 6.112)
    return (243.5 * x / (17.67 - x)).astype(dt)


def calc_srh(*args, **kwargs):
    """Storm-relative helicity with Bunkers storm motion.

    Same contract as the reference's ``xcape.core.calc_srh``
    (src/xcape/core.py:335-447): positional ``p, t, td, u, v, ps, ts, tds, us,
    vs`` (hPa, degC, m/s; levels last; p 1-D for ``vertical_lev='pressure'``),
    keywords ``depth`` [m] (default 3000), ``vertical_lev`` and ``output_var``
    ('srh' -> ``(srh_rm, srh_lm)``; 'all' -> also the Bunkers right/left-mover
    u, v and the 0-6 km mean wind u, v).  Heights above ground come from the
    hypsometric equation with a 2 m surface height (core.py:516-520).

    Extra keyword (not in the reference, whose calc_srh has no operator
    switch): ``method`` ('gpu' default, 'fortran' alias, 'dummy').
    """
    if _any_dask_array(*args):
        return _calc_srh_gufunc(*args, **kwargs)
    return _calc_srh_numpy(*args, **kwargs)


def _calc_srh_gufunc(*args, **kwargs):
    """dask path (core.py:449-469), with one input slot per argument."""
    import dask.array as da
    output_dtypes = ('f8', 'f8')
    if kwargs.get('output_var', 'srh') == 'all':
        output_dtypes = output_dtypes + ('f4',) * 6
    outs = ",".join(["()"] * len(output_dtypes))
    if kwargs.get('vertical_lev') == 'pressure':
        p = np.asarray(args[0])
        rest = args[1:]

        def _fn(*a, **kw):
            return _calc_srh_numpy(p, *a, **kw)
        signature = "(i),(i),(i),(i),(),(),(),(),()->" + outs
        return da.apply_gufunc(_fn, signature, *rest, output_dtypes=output_dtypes, axis=-1,
                               vectorize=False, **kwargs)
    signature = "(i),(i),(i),(i),(i),(),(),(),(),()->" + outs
    return da.apply_gufunc(_calc_srh_numpy, signature, *args, output_dtypes=output_dtypes,
                           axis=-1, vectorize=False, **kwargs)


def _calc_srh_numpy(*args, depth=3000, vertical_lev='sigma', output_var='srh', method='gpu'):
    """numpy path (core.py:473-542)."""
    p, t, td, u, v, ps, ts, tds, us, vs = args
    original_shape = t.shape
    p_s1d, t_s1d, td_s1d, u_s1d, v_s1d = _reshape_surface_inputs(ps, ts, tds, us, vs)
    start = None
    if len(p.shape) == 1:
        t_2d, td_2d, u_2d, v_2d = _reshape_inputs(t, td, u, v)
        p_2d = np.asarray(p)
        p_shared = True
        if p_2d.dtype != np.float32 or np.asarray(p_s1d).dtype != np.float32:
            start = pressure_start_levels(np.asarray(p_2d, np.float64),
                                          np.asarray(p_s1d, np.float64))
    elif (p.shape == t.shape) & (vertical_lev == 'sigma'):
        p_2d, t_2d, td_2d, u_2d, v_2d = _reshape_inputs(p, t, td, u, v)
        p_shared = False
    elif (p.shape == t.shape) & (vertical_lev == 'pressure'):
        raise ValueError("P should be 1d")
    else:
        raise ValueError('Input arrays must have the same shape.')
    _VLEV[vertical_lev]
    out_sel = _OUTPUT[output_var]

    if method in _GPU_METHODS:
        # float64 inputs reach the FP64 heights and helicity unrounded (the
        # reference's f2py wrappers pass them to its DOUBLE PRECISION
        # routines as given, and round them for the REAL Bunkers routine)
        f64 = any(np.asarray(a).dtype == np.float64
                  for a in (p_2d, t_2d, td_2d, u_2d, v_2d, p_s1d, t_s1d, td_s1d, u_s1d, v_s1d))
        srm, slm, rm, lm, mw = _hip.srh(p_2d, t_2d, td_2d, u_2d, v_2d, p_s1d, t_s1d, td_s1d,
                                        u_s1d, v_s1d, depth=float(depth), p_shared=p_shared,
                                        start_lev=start, inputs_f64=f64)
    elif method == 'dummy':
        n = t_2d.shape[0]
        srm, slm = np.ones(n), np.ones(n)
        rm, lm, mw = (np.ones((n, 2), np.float32) for _ in range(3))
    else:
        raise ValueError('invalid method')

    srh_rm, srh_lm = _reshape_outputs(srm, slm, shape=original_shape)
    if out_sel == 1:
        return srh_rm, srh_lm
    rm, lm, mw = _reshape_outputs_uv_components(*(np.ascontiguousarray(np.asarray(a).T)
                                                  for a in (rm, lm, mw)), shape=original_shape)
    return srh_rm, srh_lm, rm[0], rm[1], lm[0], lm[1], mw[0], mw[1]
