"""xcape_cape_noalloc_f32 against xcape_cape_f32, through the C-ABI.

The no-allocation entry point runs the same kernels with its staging and
warm-start scratch in the library's per-device scratch array instead of
stream-ordered allocations.  Columns are independent, so its outputs must
equal the allocating entry point's bit for bit (stronger than the CAPE/CIN
tolerance).  Two calls in a row, the second on fewer columns (the array
keeps the first call's contents), for the plain and warm-started methods on
column-contiguous sigma input (the staging path) and shared pressure levels;
and a call beyond the scratch limit fails with a clear message, writing
nothing.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def env():
    import torch

    from xcape_amd import _hip, synthetic
    assert _hip.device_count() >= 1, 'no HIP device visible'
    return torch, _hip, synthetic


def _call(H, fn, args, p_shared, nk, ncol, source, adiabat, flags, stream):
    torch = __import__('torch')
    outs = [torch.full((ncol,), -7.0, device='cuda'), torch.full((ncol,), -7.0, device='cuda'),
            torch.full((ncol,), -7, dtype=torch.int32, device='cuda'),
            torch.full((ncol,), -7.0, device='cuda')]
    P = H._ptr
    rc = fn(P(args[0]), p_shared, None, P(args[1]), P(args[2]), P(args[3]), P(args[4]),
            P(args[5]), nk, ncol, 500., source, 500., adiabat, *[P(o) for o in outs],
            flags | H.XCAPE_DEVICE_PTRS, None, stream)
    return rc, outs


@pytest.mark.parametrize('grid', ['sigma', 'pressure'])
@pytest.mark.parametrize('warm', [False, True])
def test_noalloc_matches_allocating_entry(env, grid, warm):
    torch, H, S = env
    L = H.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    flags = H._warm_flags('extrapolation' if warm else False)
    for ncol, seed in ((60_000, 71), (5_000, 72)):
        print(f'[noalloc] grid={grid} warm={warm} ncol={ncol}', flush=True)
        if grid == 'sigma':
            c = S.sigma_columns(ncol, 137, seed)
            args = [torch.from_numpy(c[k]).cuda() for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
            p_shared, nk, source, adiabat = 0, 137, 2, 3
        else:
            c = S.pressure_columns(ncol, seed)
            args = [torch.from_numpy(np.ascontiguousarray(c[k], np.float32)).cuda()
                    for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
            p_shared, nk, source, adiabat = 1, len(c['p']), 3, 1
        rc, ref = _call(H, L.xcape_cape_f32, args, p_shared, nk, ncol, source, adiabat, flags,
                        stream)
        H.check(rc)
        rc, got = _call(H, L.xcape_cape_noalloc_f32, args, p_shared, nk, ncol, source, adiabat,
                        flags, stream)
        H.check(rc)
        torch.cuda.synchronize()
        for g, r in zip(got, ref):
            assert torch.equal(g, r), (grid, warm, ncol)
        assert bool((got[0] >= 0).all())


def test_noalloc_rejects_calls_beyond_its_scratch(env):
    """A call needing more than the scratch array fails with a message that
    names its limit, writing nothing; a call of exactly that many columns
    runs and equals the allocating entry point's."""
    import re
    torch, H, S = env
    L = H.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    flags = H._warm_flags('extrapolation')
    nk = 137
    ncol = 80_000          # ~3.9 KB of scratch per column: above 256 MiB
    c = S.sigma_columns(ncol, nk, 73)
    args = [torch.from_numpy(c[k]).cuda() for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
    print(f'[noalloc] limit probe ncol={ncol}', flush=True)
    rc, outs = _call(H, L.xcape_cape_noalloc_f32, args, 0, nk, ncol, 2, 3, flags, stream)
    torch.cuda.synchronize()
    assert rc == -1
    msg = L.xcape_last_error().decode()
    m = re.search(r'at most (\d+) such columns', msg)
    assert 'scratch' in msg and m, msg
    limit = int(m.group(1))
    assert 60_000 < limit < ncol
    assert bool((outs[0] == -7.0).all())
    # exactly the limit: runs, and equals xcape_cape_f32
    sub = [a[:limit] for a in args]
    print(f'[noalloc] at the limit ncol={limit}', flush=True)
    rc, got = _call(H, L.xcape_cape_noalloc_f32, sub, 0, nk, limit, 2, 3, flags, stream)
    H.check(rc)
    rc, ref = _call(H, L.xcape_cape_f32, sub, 0, nk, limit, 2, 3, flags, stream)
    H.check(rc)
    torch.cuda.synchronize()
    for g, r in zip(got, ref):
        assert torch.equal(g, r)
    # one more column: refused
    sub = [a[:limit + 1] for a in args]
    rc, _ = _call(H, L.xcape_cape_noalloc_f32, sub, 0, nk, limit + 1, 2, 3, flags, stream)
    assert rc == -1
    # host pointers and stats are refused
    x = np.zeros((2, 3), np.float32)
    s = np.zeros(2, np.float32)
    m = np.zeros(2, np.int32)
    P = H._ptr
    assert L.xcape_cape_noalloc_f32(P(x), 0, None, P(x), P(x), P(s), P(s), P(s), 3, 2, 500., 1,
                                    500., 1, P(s), P(s), P(m), P(s), 0, None, None) == -1
    assert b'device pointers' in L.xcape_last_error()
