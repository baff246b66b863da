"""Host-array calls stream their columns through pinned staging slots, a chunk
at a time.  Every per-column result is independent of how the columns are
chunked, so a streamed host call must equal the device-pointer call on the
same inputs bit for bit, whatever the chunk size, the number of chunks (more
chunks than slots, a partial last one) and the number of threads calling at
once."""
import threading

import numpy as np
import pytest

from xcape_amd import synthetic

pytestmark = pytest.mark.gpu

NK = 60


@pytest.fixture(scope='module')
def hip():
    from xcape_amd import _hip
    assert _hip.device_count() >= 1, 'no HIP device visible'
    return _hip


def _dev(args):
    import torch
    return [None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda()
            for a in args]


def _host(x):
    if isinstance(x, dict):
        return x
    return x.cpu().numpy() if hasattr(x, 'cpu') else x


def _same(h, d):
    import torch
    torch.cuda.synchronize()
    assert len(h) == len(d)
    for a, b in zip(h, d):
        b = _host(b)
        if isinstance(a, dict):
            # the work counters; the warm loop's wave-trip counts depend on
            # how the columns were spread over the lanes (chunks differ)
            def lanes(x):
                d = {k: v for k, v in x.items() if k not in ('wave_trips', 'blocks')}
                d.update({k: v['lanes'] for k, v in x.get('blocks', {}).items()})
                return d
            assert lanes(a) == lanes(b)
        else:
            np.testing.assert_array_equal(a, b)


def _sigma(ncol, seed):
    c = synthetic.sigma_columns(ncol, NK, seed, winds=True)
    cape_args = [c[k] for k in ('p', 't', 'td', 'ps', 'ts', 'tds')]
    all_args = [c[k] for k in ('p', 't', 'td', 'u', 'v', 'ps', 'ts', 'tds', 'us', 'vs')]
    return cape_args, all_args


CAPE_CASES = [dict(), dict(warm_start=True), dict(fp64=True), dict(fast_math=True),
              dict(levels=True), dict(warm_start=True, stats=True), dict(source=3, adiabat=4),
              dict(source=1, adiabat=2, warm_start=True)]


@pytest.mark.parametrize('chunk_mb', [None, '1'])
@pytest.mark.parametrize('ncol', [1, 37, 5000])
def test_streamed_cape_matches_device_call(hip, monkeypatch, chunk_mb, ncol):
    """1 MB chunks are 1408 columns of 60-level CAPE input: 5000 columns take
    five chunks (384, 704, 1408, 1408 and 1096 columns) through three pinned
    input buffers and four compute slots."""
    if chunk_mb is None:
        monkeypatch.delenv('XCAPE_HOST_CHUNK_MB', raising=False)
    else:
        monkeypatch.setenv('XCAPE_HOST_CHUNK_MB', chunk_mb)
    args, _ = _sigma(ncol, 41 + ncol)
    dargs = _dev(args)
    for case in CAPE_CASES:
        kw = dict(pinc=500., source=2, adiabat=3)
        kw.update(case)
        print('case', kw, flush=True)
        _same(hip.cape(*args, **kw), hip.cape(*dargs, **kw))


@pytest.mark.parametrize('start', [False, True])
def test_streamed_pressure_levels(hip, monkeypatch, start):
    monkeypatch.setenv('XCAPE_HOST_CHUNK_MB', '1')
    c = synthetic.pressure_columns(6000, 77, winds=True)
    p = np.asarray(c['p'], np.float32)
    st = None
    if start:
        st = np.ones(6000, np.int32)
        st[::5] = 3
    args = [p, c['t'], c['td'], c['ps'], c['ts'], c['tds']]
    dargs = _dev(args)
    dst = _dev([st])[0]
    for warm in (False, True):
        kw = dict(pinc=500., source=2, adiabat=3, p_shared=True, warm_start=warm)
        _same(hip.cape(*args, start_lev=st, **kw), hip.cape(*dargs, start_lev=dst, **kw))
    wargs = [p, c['t'], c['td'], c['u'], c['v'], c['ps'], c['ts'], c['tds'], c['us'], c['vs']]
    dw = _dev(wargs)
    _same(hip.srh(*wargs, p_shared=True, start_lev=st),
          hip.srh(*dw, p_shared=True, start_lev=dst))


def test_streamed_srh_and_fused(hip, monkeypatch):
    """The fused call's 1216 bytes of input per column: 832-column chunks
    (the first two of 256 and 448 columns), eight of them for 5000 columns,
    so every input buffer and slot is reused."""
    monkeypatch.setenv('XCAPE_HOST_CHUNK_MB', '1')
    _, args = _sigma(5000, 5)
    dargs = _dev(args)
    for fp64 in (False, True):
        for motions in (False, True):
            _same(hip.srh(*args, fp64=fp64, motions=motions),
                  hip.srh(*dargs, fp64=fp64, motions=motions))
    for kw in (dict(), dict(warm_start=True, stats=True), dict(fast_math=True, warm_start=True)):
        kw.update(pinc=500., source=2, adiabat=3, depth=3000.)
        _same(hip.cape_srh(*args, **kw), hip.cape_srh(*dargs, **kw))


def test_concurrent_host_calls(hip, monkeypatch):
    """Four threads stream at once through the two staging sets (two wait);
    each gets its own columns' results."""
    monkeypatch.setenv('XCAPE_HOST_CHUNK_MB', '1')
    data = [_sigma(3000 + 500 * i, 300 + i)[1] for i in range(4)]
    kw = dict(pinc=500., source=2, adiabat=3, depth=3000., warm_start=True)
    refs = [[_host(x) for x in hip.cape_srh(*_dev(a), **kw)] for a in data]
    got = [None] * 4
    errs = []

    def run(i):
        try:
            got[i] = hip.cape_srh(*data[i], **kw)
        except Exception as e:   # reported below, in the main thread
            errs.append(e)

    th = [threading.Thread(target=run, args=(i,)) for i in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errs, errs
    for g, r in zip(got, refs):
        for a, b in zip(g, r):
            np.testing.assert_array_equal(a, b)


def test_release_host_staging(hip):
    args, _ = _sigma(2000, 9)
    first = hip.cape(*args, source=2, adiabat=3)
    assert hip.release_host_staging() >= 1
    assert hip.release_host_staging() == 0
    again = hip.cape(*args, source=2, adiabat=3)
    for a, b in zip(first, again):
        np.testing.assert_array_equal(a, b)


def test_out_of_core_memmap(hip, tmp_path):
    """Grids larger than host memory: the profiles as memory-mapped files go
    through calc_cape as they are (no copy into memory first); the host
    threads read each chunk from the mapping as they stage it."""
    from xcape_amd.core import calc_cape
    c = synthetic.sigma_columns(6000, NK, 12)
    mm = {}
    for k in ('p', 't', 'td'):
        m = np.lib.format.open_memmap(tmp_path / f'{k}.npy', mode='w+', dtype=np.float32,
                                      shape=(60, 100, NK))
        m[:] = c[k].reshape(60, 100, NK)
        m.flush()
        mm[k] = np.load(tmp_path / f'{k}.npy', mmap_mode='r')
    sfc = [c[k].reshape(60, 100) for k in ('ps', 'ts', 'tds')]
    kw = dict(source='most-unstable', adiabat='pseudo-ice', pinc=500., warm_start=True,
              vertical_lev='sigma')
    got = calc_cape(mm['p'], mm['t'], mm['td'], *sfc, **kw)
    ref = calc_cape(*(np.array(mm[k]) for k in ('p', 't', 'td')), *sfc, **kw)
    assert len(got) == 4
    for a, b in zip(got, ref):
        assert a.shape == (60, 100)
        np.testing.assert_array_equal(a, b)
