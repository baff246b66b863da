"""Seeded synthetic soundings for tests and benchmarks.

Every column is drawn independently:

* surface pressure ps ~ U(95000, 102000) Pa, surface temperature
  ts ~ U(285, 305) K;
* temperature falls at a lapse rate ~ U(5.5, 8.0) K/km up to a tropopause at
  z_t ~ U(11, 15) km and is isothermal above;
* dewpoint depression D(z) = D0 + g_d z with D0 ~ U(1, 15) K and
  g_d ~ U(1, 4) K/km;
* heights follow the hypsometric equation integrated upward from z = 0 at ps
  with the layer-mean virtual temperature (three fixed-point passes per layer,
  since T depends on z);
* winds (optional): u = U0 + S z cos(th) + e, v = S z sin(th) + e with
  U0 ~ U(-5, 10) m/s, S ~ U(0, 0.01) 1/s below 6 km and S/2 above (continuous
  at 6 km), th ~ U(0, 2 pi) and e ~ N(0, 1) per level; the surface wind is the
  same profile at z = 2 m.

Sigma grids put level k at sigma_k * ps with sigma = linspace(0.995, 0.1, nk).
Pressure grids use fixed levels; levels below the ground are filled with NaN
(the algorithm must never read them).

Outputs are in the units of ``calc_cape``: hPa and degC, float32, with levels
on the last axis.
"""
import numpy as np

RD = 287.04
G = 9.81


def _draw(rng, n):
    return dict(
        ps=rng.uniform(95000.0, 102000.0, n),
        ts=rng.uniform(285.0, 305.0, n),
        lapse=rng.uniform(5.5, 8.0, n) * 1e-3,
        ztrop=rng.uniform(11000.0, 15000.0, n),
        d0=rng.uniform(1.0, 15.0, n),
        gd=rng.uniform(1.0, 4.0, n) * 1e-3,
    )


def _temp(d, z):
    zc = np.minimum(z, d['ztrop'][..., None] if z.ndim > 1 else d['ztrop'])
    if z.ndim > 1:
        return d['ts'][:, None] - d['lapse'][:, None] * zc
    return d['ts'] - d['lapse'] * zc


def _dewp(d, z, t):
    if z.ndim > 1:
        dep = d['d0'][:, None] + d['gd'][:, None] * z
    else:
        dep = d['d0'] + d['gd'] * z
    return t - dep


def _tv(t, td, p):
    es = 611.2 * np.exp(17.67 * (td - 273.15) / (td - 29.65))
    r = 0.622 * es / np.maximum(p - es, 1.0)
    return t * (1.0 + 0.608 * r)


def _heights(d, p):
    """Integrate z upward along the columns of p (ncol, nk), p[:, 0] = ps."""
    ncol, nk = p.shape
    z = np.zeros((ncol, nk))
    sub = {k: v for k, v in d.items()}
    t0 = d['ts']
    tv_prev = _tv(t0, _dewp(sub, np.zeros(ncol), t0), p[:, 0])
    for k in range(1, nk):
        zk = z[:, k - 1] + RD * tv_prev / G * np.log(p[:, k - 1] / p[:, k])
        for _ in range(3):
            tk = _temp(sub, zk)
            tvk = _tv(tk, _dewp(sub, zk, tk), p[:, k])
            zk = z[:, k - 1] + RD * 0.5 * (tv_prev + tvk) / G * np.log(p[:, k - 1] / p[:, k])
        z[:, k] = zk
        tk = _temp(sub, zk)
        tv_prev = _tv(tk, _dewp(sub, zk, tk), p[:, k])
    return z


def _winds(rng, z, zs):
    n = z.shape[0]
    u0 = rng.uniform(-5.0, 10.0, n)[:, None]
    s = rng.uniform(0.0, 0.01, n)[:, None]
    th = rng.uniform(0.0, 2 * np.pi, n)[:, None]

    def prof(zz):
        return np.where(zz <= 6000.0, s * zz, s * 6000.0 + 0.5 * s * (zz - 6000.0))

    sh = prof(z)
    u = u0 + sh * np.cos(th) + rng.normal(0.0, 1.0, z.shape)
    v = sh * np.sin(th) + rng.normal(0.0, 1.0, z.shape)
    shs = prof(np.full((n, 1), zs))
    us = (u0 + shs * np.cos(th))[:, 0]
    vs = (shs * np.sin(th))[:, 0]
    return u, v, us, vs


def sigma_columns(ncol, nlev, seed, winds=False):
    """Sigma-grid soundings.  Returns a dict of float32 arrays:
    p, t, td (ncol, nlev) and ps, ts, tds (ncol,) [+ u, v, us, vs]."""
    rng = np.random.default_rng(seed)
    d = _draw(rng, ncol)
    sig = np.linspace(0.995, 0.1, nlev)
    p = np.concatenate([d['ps'][:, None], d['ps'][:, None] * sig[None, :]], axis=1)
    z = _heights(d, p)
    t = _temp(d, z)
    td = _dewp(d, z, t)
    out = dict(p=p[:, 1:] / 100.0, t=t[:, 1:] - 273.15, td=td[:, 1:] - 273.15,
               ps=d['ps'] / 100.0, ts=d['ts'] - 273.15, tds=td[:, 0] - 273.15)
    if winds:
        u, v, us, vs = _winds(rng, z[:, 1:], 2.0)
        out.update(u=u, v=v, us=us, vs=vs)
    return {k: np.ascontiguousarray(a, dtype=np.float32) for k, a in out.items()}


PRESSURE_LEVELS_HPA = np.arange(1000.0, 99.0, -25.0)   # 37 levels, 1000..100 hPa


def pressure_columns(ncol, seed, levels=PRESSURE_LEVELS_HPA, winds=False):
    """Fixed-pressure-level soundings.  p is the shared (nlev,) level vector in
    hPa; t/td below the ground are NaN."""
    rng = np.random.default_rng(seed)
    d = _draw(rng, ncol)
    lev = np.asarray(levels, np.float64)
    pl = lev * 100.0
    # integrate on a column-specific grid: ps followed by every level above ground
    pcol = np.concatenate([d['ps'][:, None], np.minimum(pl[None, :], d['ps'][:, None])], axis=1)
    # levels below ground get p == ps -> zero-thickness layers, masked below
    z = _heights(d, pcol)
    t = _temp(d, z)
    td = _dewp(d, z, t)
    below = pl[None, :] > d['ps'][:, None]
    t3 = np.where(below, np.nan, t[:, 1:] - 273.15)
    td3 = np.where(below, np.nan, td[:, 1:] - 273.15)
    out = dict(p=lev, t=t3, td=td3, ps=d['ps'] / 100.0, ts=d['ts'] - 273.15,
               tds=td[:, 0] - 273.15)
    if winds:
        u, v, us, vs = _winds(rng, z[:, 1:], 2.0)
        out.update(u=np.where(below, np.nan, u), v=np.where(below, np.nan, v), us=us, vs=vs)
    return {k: np.ascontiguousarray(a, dtype=np.float32) for k, a in out.items()}


RAGGED_NLEVS = (30, 40, 60, 91, 137)


def ragged_columns(ncol, seed):
    """Ragged batch: per-column level count from RAGGED_NLEVS, source from
    {1 surface, 2 most-unstable, 3 mixed-layer} and adiabat from {1
    pseudo-liquid, 3 pseudo-ice}, all uniform.

    Returns concatenated level data (CSR): offsets int64 (ncol+1,), p, t, td
    (sum nk,), surface arrays, nlev, source and adiabat (int32)."""
    rng = np.random.default_rng(seed)
    nlev = rng.choice(np.asarray(RAGGED_NLEVS), ncol)
    source = rng.choice(np.asarray([1, 2, 3]), ncol).astype(np.int32)
    adiabat = rng.choice(np.asarray([1, 3]), ncol).astype(np.int32)
    offsets = np.zeros(ncol + 1, np.int64)
    offsets[1:] = np.cumsum(nlev)
    ptot = np.empty(offsets[-1], np.float32)
    ttot = np.empty_like(ptot)
    tdtot = np.empty_like(ptot)
    ps = np.empty(ncol, np.float32)
    ts = np.empty_like(ps)
    tds = np.empty_like(ps)
    sub_seeds = rng.integers(0, 2**31, len(RAGGED_NLEVS))
    for j, nk in enumerate(RAGGED_NLEVS):
        idx = np.nonzero(nlev == nk)[0]
        if idx.size == 0:
            continue
        c = sigma_columns(idx.size, nk, int(sub_seeds[j]))
        rows = offsets[idx][:, None] + np.arange(nk)[None, :]
        ptot[rows] = c['p']
        ttot[rows] = c['t']
        tdtot[rows] = c['td']
        ps[idx], ts[idx], tds[idx] = c['ps'], c['ts'], c['tds']
    return dict(offsets=offsets, p=ptot, t=ttot, td=tdtot, ps=ps, ts=ts, tds=tds,
                nlev=nlev.astype(np.int32), source=source, adiabat=adiabat)


def sigma_columns_torch(ncol, nlev, seed, device='cuda', winds=False):
    """The sigma-grid generator above evaluated with torch on `device` (same
    distributions and recipe, float64 arithmetic, float32 outputs), for
    workloads too large to build quickly in numpy.  Values are not
    bit-identical to ``sigma_columns`` (different random streams)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    f64 = dict(dtype=torch.float64, device=device)

    def U(lo, hi):
        return lo + (hi - lo) * torch.rand(ncol, generator=g, **f64)
    ps, ts = U(95000.0, 102000.0), U(285.0, 305.0)
    lapse, ztrop = U(5.5, 8.0) * 1e-3, U(11000.0, 15000.0)
    d0, gd = U(1.0, 15.0), U(1.0, 4.0) * 1e-3

    def temp(z):
        return ts - lapse * torch.minimum(z, ztrop)

    def tvirt(t, td, p):
        es = 611.2 * torch.exp(17.67 * (td - 273.15) / (td - 29.65))
        r = 0.622 * es / torch.clamp(p - es, min=1.0)
        return t * (1.0 + 0.608 * r)

    sig = torch.linspace(0.995, 0.1, nlev, **f64)
    out = {k: torch.empty((ncol, nlev), dtype=torch.float32, device=device)
           for k in ('p', 't', 'td')}
    zall = torch.empty((ncol, nlev), **f64) if winds else None
    z_prev = torch.zeros(ncol, **f64)
    p_prev = ps
    t0 = ts
    tds = t0 - d0
    tv_prev = tvirt(t0, tds, ps)
    for k in range(nlev):
        pk = ps * sig[k]
        lnr = torch.log(p_prev / pk)
        zk = z_prev + RD * tv_prev / G * lnr
        for _ in range(3):
            tk = temp(zk)
            tvk = tvirt(tk, tk - (d0 + gd * zk), pk)
            zk = z_prev + RD * 0.5 * (tv_prev + tvk) / G * lnr
        tk = temp(zk)
        tdk = tk - (d0 + gd * zk)
        out['p'][:, k] = (pk / 100.0).float()
        out['t'][:, k] = (tk - 273.15).float()
        out['td'][:, k] = (tdk - 273.15).float()
        if winds:
            zall[:, k] = zk
        tv_prev = tvirt(tk, tdk, pk)
        z_prev, p_prev = zk, pk
    out['ps'] = (ps / 100.0).float()
    out['ts'] = (ts - 273.15).float()
    out['tds'] = (tds - 273.15).float()
    if winds:
        u0 = U(-5.0, 10.0)[:, None]
        s = U(0.0, 0.01)[:, None]
        th = U(0.0, 2 * np.pi)[:, None]

        def prof(zz):
            return torch.where(zz <= 6000.0, s * zz, s * 6000.0 + 0.5 * s * (zz - 6000.0))
        sh = prof(zall)
        noise = torch.randn((ncol, nlev), generator=g, **f64)
        out['u'] = (u0 + sh * torch.cos(th) + noise).float()
        noise = torch.randn((ncol, nlev), generator=g, **f64)
        out['v'] = (sh * torch.sin(th) + noise).float()
        shs = prof(torch.full((ncol, 1), 2.0, **f64))
        out['us'] = (u0 + shs * torch.cos(th))[:, 0].float()
        out['vs'] = (shs * torch.sin(th))[:, 0].float()
        del zall
    return {k: v.contiguous() for k, v in out.items()}
