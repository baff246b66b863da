"""ctypes binding of ``libxcape_hip.so`` (the C-ABI in ``include/xcape_hip.h``).

This is the package's only route to the compute: there is no CPU fallback.
If the library is missing or no gfx950 device is visible the calls raise.

Arrays may be numpy arrays (host pointers: the library stages them through
device memory) or torch tensors on a HIP device (device pointers: the call is
enqueued on torch's current stream).  Profiles are (ncol, nk) C-contiguous,
levels last.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get('XCAPE_HIP_LIB', os.path.join(HERE, 'libxcape_hip.so'))

XCAPE_DEVICE_PTRS = 0x1
XCAPE_FAST_MATH = 0x2
XCAPE_WARM_START = 0x4
XCAPE_NO_SURROGATE = 0x8
XCAPE_LEVEL_MAJOR = 0x10
XCAPE_NO_PREDICTION = 0x20
XCAPE_INPUT_F64 = 0x40


class XcapeError(RuntimeError):
    """Error reported by libxcape_hip (message from xcape_last_error)."""

    def __init__(self, code, msg):
        super().__init__(f'xcape_hip error {code}: {msg}')
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [('columns', ctypes.c_int64), ('levels', ctypes.c_int64),
                ('substeps', ctypes.c_int64), ('iterations', ctypes.c_int64),
                ('nonconverged', ctypes.c_int64), ('fallbacks', ctypes.c_int64),
                ('first_substep_iterations', ctypes.c_int64), ('reserved', ctypes.c_int64 * 1),
                ('wave_trips', ctypes.c_int64), ('block_lanes', ctypes.c_int64 * 6),
                ('block_waves', ctypes.c_int64 * 6)]

    BLOCKS = ('level', 'substep', 'iteration1', 'iteration2', 'completion', 'layer')

    def as_dict(self):
        d = {n: int(getattr(self, n)) for n, _ in self._fields_[:7]}
        if self.wave_trips:
            d['wave_trips'] = int(self.wave_trips)
            d['blocks'] = {b: {'lanes': int(self.block_lanes[i]), 'waves': int(self.block_waves[i]),
                               'lane_utilisation': self.block_lanes[i] / (64.0 * self.block_waves[i])
                               if self.block_waves[i] else None}
                           for i, b in enumerate(self.BLOCKS)}
        return d


_lib = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double
_U32 = ctypes.c_uint32

# name -> (restype, argtypes); must match include/xcape_hip.h
SIGNATURES = {
    'xcape_version': (ctypes.c_char_p, []),
    'xcape_last_error': (ctypes.c_char_p, []),
    'xcape_device_count': (_I, []),
    'xcape_release_host_staging': (_I, []),
    'xcape_cape_f32': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _I64, _I64, _F, _I, _F, _I,
                            _P, _P, _P, _P, _U32, _P, _P]),
    'xcape_cape_levels_f32': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _I64, _I64, _F, _I, _F, _I,
                                   _P, _P, _P, _P, _P, _P, _U32, _P]),
    'xcape_cape_noalloc_f32': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _I64, _I64, _F, _I, _F, _I,
                                    _P, _P, _P, _P, _U32, _P, _P]),
    'xcape_cape_f64': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _I64, _I64, _D, _I, _D, _I,
                            _P, _P, _P, _P, _U32, _P, _P]),
    'xcape_cape_levels_f64': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _I64, _I64, _D, _I, _D, _I,
                                   _P, _P, _P, _P, _P, _P, _U32, _P]),
    'xcape_srh_f32': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _I64, _D,
                           _P, _P, _P, _P, _P, _U32, _P]),
    'xcape_srh_f64': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _I64, _D,
                           _P, _P, _P, _P, _P, _U32, _P]),
    'xcape_cape_srh_f32': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _I64,
                                _F, _I, _F, _I, _D, _P, _P, _P, _P, _P, _P, _P, _P, _P, _U32,
                                _P, _P]),
    'xcape_ragged_plan_create': (_I, [_P, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                      _F, _F, _P, _P, _P, _P, _I, _I, _I, _U32, _P]),
    'xcape_ragged_plan_run': (_I, [_P, _P, _P]),
    'xcape_ragged_plan_info': (_I, [_P, _P, _P]),
    'xcape_ragged_plan_destroy': (_I, [_P]),
    'xcape_mlp_nparams': (_I, []),
    'xcape_mlp_forward': (_I, [_P, _I64, _P, _P, _P]),
    'xcape_mlp_loss_grad': (_I, [_P, _P, _I64, _P, _P, _P, _P]),
    'xcape_mlp_adam': (_I, [_P, _P, _P, _P, _I64, _F, _F, _F, _F, _F, _P]),
    'xcape_surrogate_load': (_I, [_P, _P, _P]),
    'xcape_surrogate_loaded': (_I, []),
    'xcape_surrogate_eval': (_I, [_P, _I64, _P, _P]),
    'xcape_surrogate_log_f32': (_I, [_P, _P, _P, _P, _P, _P, _I64, _I64, _F, _I, _F, _I, _F,
                                     _U32, _I, _P, _I64, _P, _P]),
    'xcape_math_check': (_I, [_I, _P, _P, _I64, _P, _P]),
    'xcape_math_check_f64': (_I, [_I, _P, _I64, _P, _P]),
}


ROCM_HIP = os.environ.get('XCAPE_HIP_RUNTIME', '/opt/rocm/lib/libamdhip64.so.7')


def _hip_runtime():
    """Make exactly one HIP runtime visible with global scope.

    libxcape_hip.so carries no hard dependency on libamdhip64: its HIP calls
    bind to the runtime already in the process.  When torch (ROCm build) is
    importable its bundled runtime is the one: importing torch loads it, and
    re-opening it with RTLD_GLOBAL exposes it to the library, so streams and
    device memory are shared with torch.  Otherwise the system runtime is
    loaded with global scope.  Returns the path used.
    """
    try:
        import torch
        if getattr(torch.version, 'hip', None):
            path = os.path.join(os.path.dirname(torch.__file__), 'lib', 'libamdhip64.so')
            if os.path.exists(path):
                ctypes.CDLL(path, mode=os.RTLD_GLOBAL | os.RTLD_NOLOAD)
                return path
    except (ImportError, OSError):
        pass
    ctypes.CDLL(ROCM_HIP, mode=os.RTLD_GLOBAL)
    return ROCM_HIP


def lib():
    """Load the library (once) and declare the prototypes."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f'{LIB_PATH} is not built; run `python -m xcape_amd.build` (hipcc, gfx950)')
        _hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise XcapeError(rc, lib().xcape_last_error().decode())


def device_count():
    return lib().xcape_device_count()


def release_host_staging():
    """Frees the pinned and device staging that host-array calls keep between
    calls (the sets no call is using); returns how many sets it freed."""
    return lib().xcape_release_host_staging()


def _is_torch(a):
    return type(a).__module__.startswith('torch')


def _ptr(a):
    if a is None:
        return None
    if _is_torch(a):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


class _Args:
    """Normalise a set of arrays to one residency (host numpy or device torch)."""

    def __init__(self, arrays):
        flags = [_is_torch(a) for a in arrays if a is not None]
        if any(flags) and not all(flags):
            raise TypeError('mix of torch tensors and numpy arrays')
        self.device = bool(flags) and all(flags)
        if self.device:
            import torch
            for a in arrays:
                if a is not None and not a.is_cuda:
                    raise TypeError('torch tensors must live on a HIP device')
            self.stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        else:
            self.stream = None

    def f32(self, a):
        if a is None:
            return None
        if self.device:
            import torch
            return a.to(torch.float32).contiguous()
        return np.ascontiguousarray(a, dtype=np.float32)

    def f64(self, a):
        if a is None:
            return None
        if self.device:
            import torch
            return a.to(torch.float64).contiguous()
        return np.ascontiguousarray(a, dtype=np.float64)

    def i32(self, a):
        if a is None:
            return None
        if self.device:
            import torch
            return a.to(torch.int32).contiguous()
        return np.ascontiguousarray(a, dtype=np.int32)

    def empty(self, n, dtype, like):
        if self.device:
            import torch
            tdt = {np.float32: torch.float32, np.float64: torch.float64,
                   np.int32: torch.int32}[dtype]
            return torch.empty(n, dtype=tdt, device=like.device)
        return np.empty(n, dtype)


def _warm_flags(warm_start):
    """warm_start: False (the reference's loops, cold), True (extrapolation +
    loaded surrogate), 'extrapolation' (ignore a loaded surrogate) or 'cold'
    (the warm start's loop with every sub-step cold)."""
    if not warm_start:
        return 0
    f = XCAPE_WARM_START
    if warm_start == 'extrapolation':
        f |= XCAPE_NO_SURROGATE
    elif warm_start == 'cold':
        f |= XCAPE_NO_SURROGATE | XCAPE_NO_PREDICTION
    elif warm_start is not True:
        raise ValueError(f'warm_start={warm_start!r}')
    return f


def cape(p, t, td, ps, ts, tds, *, pinc=500., source=1, ml_depth=500., adiabat=1,
         p_shared=False, start_lev=None, fast_math=False, warm_start=False, fp64=False,
         stats=False, level_major=False, levels=False):
    """Run xcape_cape_f32 (or _f64, or with ``levels`` xcape_cape_levels_f32 /
    _f64).
    t, td: (ncol, nk); p: (ncol, nk) or, with p_shared, (nk,); a
    ``start_lev`` with (ncol, nk) p starts each column at its own level of
    its own p (the reference's loopcape_pl); with
    ``level_major`` the profile arrays are (nk, ncol).  Returns cape, cin,
    mulev, zmulev (then p_lfc, p_el [hPa] with ``levels``; then a stats dict
    with ``stats``)."""
    if levels and stats:
        raise ValueError('levels are computed without stats')
    A = _Args([p, t, td, ps, ts, tds, start_lev])
    p, t, td, ps, ts, tds = (A.f32(a) for a in (p, t, td, ps, ts, tds))
    start_lev = A.i32(start_lev)
    ncol, nk = t.shape[::-1] if level_major else t.shape
    odt = np.float64 if fp64 else np.float32
    cape_o, cin_o, z_o = (A.empty(ncol, odt, t) for _ in range(3))
    mul_o = A.empty(ncol, np.int32, t)
    flags = (XCAPE_DEVICE_PTRS if A.device else 0) | (XCAPE_FAST_MATH if fast_math else 0) \
        | _warm_flags(warm_start) | (XCAPE_LEVEL_MAJOR if level_major else 0)
    st = Stats()
    if levels:
        lfc_o, el_o = (A.empty(ncol, odt, t) for _ in range(2))
        fn = lib().xcape_cape_levels_f64 if fp64 else lib().xcape_cape_levels_f32
        rc = fn(
            _ptr(p), 1 if p_shared else 0, _ptr(start_lev), _ptr(t), _ptr(td), _ptr(ps),
            _ptr(ts), _ptr(tds), nk, ncol, pinc, source, ml_depth, adiabat, _ptr(cape_o),
            _ptr(cin_o), _ptr(mul_o), _ptr(z_o), _ptr(lfc_o), _ptr(el_o), flags, A.stream)
        check(rc)
        return cape_o, cin_o, mul_o, z_o, lfc_o, el_o
    fn = lib().xcape_cape_f64 if fp64 else lib().xcape_cape_f32
    rc = fn(_ptr(p), 1 if p_shared else 0, _ptr(start_lev), _ptr(t), _ptr(td), _ptr(ps),
            _ptr(ts), _ptr(tds), nk, ncol, pinc, source, ml_depth, adiabat,
            _ptr(cape_o), _ptr(cin_o), _ptr(mul_o), _ptr(z_o), flags,
            ctypes.byref(st) if stats else None, A.stream)
    check(rc)
    if stats:
        return cape_o, cin_o, mul_o, z_o, st.as_dict()
    return cape_o, cin_o, mul_o, z_o


def srh(p, t, td, u, v, ps, ts, tds, us, vs, *, depth=3000., p_shared=False, start_lev=None,
        motions=True, level_major=False, fp64=False, inputs_f64=False):
    """Run xcape_srh_f32 (xcape_srh_f64 when ``fp64``).  Returns srh_rm,
    srh_lm (float64) and, when ``motions``, rm, lm, mean6 as (ncol, 2) (u, v)
    pairs, float32 (float64 when ``fp64``).  ``inputs_f64``: pass the inputs
    as float64 (XCAPE_INPUT_F64), so that the FP64 heights and helicity see
    them unrounded."""
    A = _Args([p, t, td, u, v, ps, ts, tds, us, vs, start_lev])
    conv = A.f64 if inputs_f64 else A.f32
    p, t, td, u, v, ps, ts, tds, us, vs = (conv(a) for a in (p, t, td, u, v, ps, ts, tds,
                                                             us, vs))
    start_lev = A.i32(start_lev)
    ncol, nk = t.shape[::-1] if level_major else t.shape
    srm, slm = (A.empty(ncol, np.float64, t) for _ in range(2))
    if motions:
        mdt = np.float64 if fp64 else np.float32
        rm, lm, mw = (A.empty(2 * ncol, mdt, t) for _ in range(3))
    else:
        rm = lm = mw = None
    flags = (XCAPE_DEVICE_PTRS if A.device else 0) | (XCAPE_LEVEL_MAJOR if level_major else 0) \
        | (XCAPE_INPUT_F64 if inputs_f64 else 0)
    fn = lib().xcape_srh_f64 if fp64 else lib().xcape_srh_f32
    rc = fn(_ptr(p), 1 if p_shared else 0, _ptr(start_lev), _ptr(t), _ptr(td),
            _ptr(u), _ptr(v), _ptr(ps), _ptr(ts), _ptr(tds), _ptr(us), _ptr(vs), nk, ncol,
            depth, _ptr(srm), _ptr(slm), _ptr(rm), _ptr(lm), _ptr(mw), flags, A.stream)
    check(rc)
    if not motions:
        return srm, slm
    return srm, slm, rm.reshape(ncol, 2), lm.reshape(ncol, 2), mw.reshape(ncol, 2)


def cape_srh(p, t, td, u, v, ps, ts, tds, us, vs, *, pinc=500., source=1, ml_depth=500.,
             adiabat=1, depth=3000., p_shared=False, start_lev=None, fast_math=False,
             warm_start=False, stats=False, level_major=False):
    """Run the fused xcape_cape_srh_f32: (cape, cin, mulev, zmulev, srh_rm,
    srh_lm, rm, lm, mean6[, stats])."""
    A = _Args([p, t, td, u, v, ps, ts, tds, us, vs, start_lev])
    p, t, td, u, v, ps, ts, tds, us, vs = (A.f32(a) for a in (p, t, td, u, v, ps, ts, tds,
                                                              us, vs))
    start_lev = A.i32(start_lev)
    ncol, nk = t.shape[::-1] if level_major else t.shape
    cape_o, cin_o, z_o = (A.empty(ncol, np.float32, t) for _ in range(3))
    mul_o = A.empty(ncol, np.int32, t)
    srm, slm = (A.empty(ncol, np.float64, t) for _ in range(2))
    rm, lm, mw = (A.empty(2 * ncol, np.float32, t) for _ in range(3))
    flags = (XCAPE_DEVICE_PTRS if A.device else 0) | (XCAPE_FAST_MATH if fast_math else 0) \
        | _warm_flags(warm_start) | (XCAPE_LEVEL_MAJOR if level_major else 0)
    st = Stats()
    rc = lib().xcape_cape_srh_f32(
        _ptr(p), 1 if p_shared else 0, _ptr(start_lev), _ptr(t), _ptr(td), _ptr(u), _ptr(v),
        _ptr(ps), _ptr(ts), _ptr(tds), _ptr(us), _ptr(vs), nk, ncol, pinc, source, ml_depth,
        adiabat, depth, _ptr(cape_o), _ptr(cin_o), _ptr(mul_o), _ptr(z_o), _ptr(srm), _ptr(slm),
        _ptr(rm), _ptr(lm), _ptr(mw), flags, ctypes.byref(st) if stats else None, A.stream)
    check(rc)
    out = (cape_o, cin_o, mul_o, z_o, srm, slm, rm.reshape(ncol, 2), lm.reshape(ncol, 2),
           mw.reshape(ncol, 2))
    if stats:
        return out + (st.as_dict(),)
    return out


def math_check(fn, a, b=None):
    """xcape_math_check on device tensors: the accurate mode's arithmetic
    kernel `fn` applied elementwise (see include/xcape_hip.h)."""
    import torch
    a = a.to(torch.float32).contiguous()
    b = None if b is None else b.to(torch.float32).contiguous()
    y = torch.empty_like(a)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(lib().xcape_math_check(fn, _ptr(a), _ptr(b), a.numel(), _ptr(y), stream))
    return y


def math_check_f64(fn, a):
    """xcape_math_check_f64 on a device tensor, elementwise: the helicity
    chain's FP64 log kernel (fn 0), the device library's log (1), the FP64 exp
    kernel (2), the device library's exp (3); the raw division kernel (4) and
    IEEE division (5) of the pairs in the last axis of `a` (..., 2)."""
    import torch
    a = a.to(torch.float64).contiguous()
    if fn >= 4:
        if a.shape[-1] != 2:
            raise ValueError('division takes operand pairs: a.shape[-1] must be 2')
        y = torch.empty(a.shape[:-1], dtype=torch.float64, device=a.device)
    else:
        y = torch.empty_like(a)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(lib().xcape_math_check_f64(fn, _ptr(a), y.numel(), _ptr(y), stream))
    return y
