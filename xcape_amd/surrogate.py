"""The warm-start surrogate: data, training on the GPU, storage, loading.

The network (3 -> 64 -> 64 -> 1, tanh-form GELU, BF16 weights, FP32
accumulation) predicts a pinc sub-step's potential-temperature rate
r = (theta2 - theta1) / ln(p2/p1) from ln p2, the parcel's theta-e and its
total water.  The warm-started ascent kernels use it where they have no
history to extrapolate from (the first sub-step of every ascent); the FP32
fixed-point iteration then runs to the reference tolerance, with the cold
start as fallback.

Everything numeric runs in libxcape_hip.so: the samples come from the
cold-start ascent kernel (``xcape_surrogate_log_f32``), training uses the
library's forward/backward/Adam kernels, evaluation its in-kernel MFMA path.
torch only holds the device buffers.

Training data recipe: the first sub-step of the cold-start ascent of
ragged-batch columns (level counts 30/40/60/91/137, surface / mixed-layer /
most-unstable parcels, pseudo-liquid and pseudo-ice adiabats, pinc 500 Pa),
generated on the GPU from a fixed seed; the validation columns come from a
different seed (a column-level split).  The first sub-step is the one the
network serves: every later sub-step has the column's own history to
extrapolate from, which predicts far closer than a BF16 network can.
"""
import ctypes
import json
import os
import time

import numpy as np

from . import _hip

NLEVS = (30, 40, 60, 91, 137)
SOURCES = (1, 2, 3)
ADIABATS = (1, 3)
FORMAT_VERSION = 1
DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data',
                            'surrogate_v1.npz')


def nparams():
    return _hip.lib().xcape_mlp_nparams()


def init_params(seed=0):
    """He-style initial parameter vector (W1, b1, W2, b2, w3, b3)."""
    h, n_in = 64, 3
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2.0 / n_in), (h, n_in))
    W2 = rng.normal(0, np.sqrt(2.0 / h), (h, h))
    w3 = rng.normal(0, np.sqrt(1.0 / h), h)
    return np.concatenate([W1.ravel(), np.zeros(h), W2.ravel(), np.zeros(h), w3,
                           [0.0]]).astype(np.float32)


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def log_samples(ncol, seed, keep_fraction=1.0, first_only=True, pinc=500., ml_depth=500.,
                chunk=1 << 20):
    """Run the cold-start ascent over `ncol` ragged-recipe columns and return
    the logged records (ln p2, theta-e, qt, rate, ln(p2/p1)) as an (n, 5)
    CUDA tensor.  first_only: one record per ascent, its first sub-step (the
    sub-step the warm start has no history for); else a hashed
    `keep_fraction` of all sub-steps.  Columns are generated on the GPU in
    chunks of `chunk`."""
    import torch

    from . import synthetic
    rng = np.random.default_rng(seed)
    per_col = 1.0 if first_only else 300.0 * keep_fraction
    capacity = int(ncol * per_col) + 1024
    out = torch.empty((capacity, 5), dtype=torch.float32, device='cuda')
    count = torch.zeros(1, dtype=torch.int64, device='cuda')
    P = _hip._ptr
    for c0 in range(0, ncol, chunk):
        n = min(chunk, ncol - c0)
        nlev = rng.choice(np.asarray(NLEVS), n)
        src = rng.choice(np.asarray(SOURCES), n)
        adi = rng.choice(np.asarray(ADIABATS), n)
        for nk in NLEVS:
            m = nlev == nk
            if not m.any():
                continue
            cols = synthetic.sigma_columns_torch(int(m.sum()), nk, int(rng.integers(2**31)),
                                                 device='cuda')
            srcs, adis = src[m], adi[m]
            for s in SOURCES:
                for a in ADIABATS:
                    idx = np.nonzero((srcs == s) & (adis == a))[0]
                    if idx.size == 0:
                        continue
                    it = torch.as_tensor(idx, device='cuda')
                    g = {k: v[it].contiguous() for k, v in cols.items()}
                    _hip.check(_hip.lib().xcape_surrogate_log_f32(
                        P(g['p']), P(g['t']), P(g['td']), P(g['ps']), P(g['ts']), P(g['tds']),
                        nk, idx.size, pinc, s, ml_depth, a, keep_fraction,
                        int(rng.integers(2**32)), 1 if first_only else 0, P(out), capacity,
                        P(count), _stream()))
            del cols
    torch.cuda.synchronize()
    n = min(int(count.item()), capacity)
    return out[:n]


def guess_error(s, samples):
    """|theta guess - converged theta| [K] of the first iterate the surrogate
    gives for logged sub-steps (rate error times the sub-step's ln-p step)."""
    import torch
    r = s.predict(samples[:, :3].contiguous())
    return (r - samples[:, 3]).abs() * samples[:, 4].abs()


class Surrogate:
    """FP32 master parameters plus input/output normalisation."""

    def __init__(self, theta, norm, meta=None):
        self.theta = np.ascontiguousarray(theta, np.float32)
        self.norm = np.ascontiguousarray(norm, np.float32)
        self.meta = dict(meta or {})
        assert self.theta.shape == (nparams(),) and self.norm.shape == (8,)

    # -- storage -------------------------------------------------------------
    def save(self, path):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        meta = dict(self.meta, format_version=FORMAT_VERSION)
        np.savez(path, theta=self.theta, norm=self.norm, meta=np.frombuffer(
            json.dumps(meta).encode(), np.uint8))

    @classmethod
    def load(cls, path=DEFAULT_PATH):
        with np.load(path, allow_pickle=False) as z:
            meta = json.loads(bytes(z['meta']).decode())
            if meta.get('format_version') != FORMAT_VERSION:
                raise ValueError(f'{path}: unsupported surrogate format {meta}')
            return cls(z['theta'], z['norm'], meta)

    # -- use ----------------------------------------------------------------
    def activate(self):
        """Make this network the one the warm-started kernels use."""
        import torch
        th = torch.as_tensor(self.theta, device='cuda')
        nm = torch.as_tensor(self.norm, device='cuda')
        _hip.check(_hip.lib().xcape_surrogate_load(_hip._ptr(th), _hip._ptr(nm), _stream()))
        torch.cuda.synchronize()

    def normalise(self, x):
        """Raw (n, 3) inputs -> the network's normalised inputs."""
        return (x - self.norm[:3]) * self.norm[3:6]

    def predict(self, x_raw):
        """Rates for raw inputs through the training (VALU) forward."""
        import torch
        x = torch.as_tensor(np.asarray(x_raw, np.float32) if not torch.is_tensor(x_raw)
                            else x_raw, device='cuda')
        u = ((x - torch.as_tensor(self.norm[:3], device='cuda')) *
             torch.as_tensor(self.norm[3:6], device='cuda')).contiguous()
        th = torch.as_tensor(self.theta, device='cuda')
        y = torch.empty(u.shape[0], dtype=torch.float32, device='cuda')
        _hip.check(_hip.lib().xcape_mlp_forward(_hip._ptr(u), u.shape[0], _hip._ptr(th),
                                                _hip._ptr(y), _stream()))
        return y * float(self.norm[6]) + float(self.norm[7])


def eval_in_kernel(x_raw):
    """The loaded surrogate through the ascent kernels' MFMA path."""
    import torch
    x = x_raw.contiguous() if torch.is_tensor(x_raw) else torch.as_tensor(
        np.asarray(x_raw, np.float32), device='cuda')
    y = torch.empty(x.shape[0], dtype=torch.float32, device='cuda')
    _hip.check(_hip.lib().xcape_surrogate_eval(_hip._ptr(x), x.shape[0], _hip._ptr(y),
                                               _stream()))
    return y


def loss_grad(theta, u, t):
    """(loss, grad) of the library's backward for normalised inputs u and
    targets t (CUDA tensors)."""
    import torch
    g = torch.empty(nparams(), dtype=torch.float32, device='cuda')
    loss = torch.empty(1, dtype=torch.float32, device='cuda')
    th = theta if torch.is_tensor(theta) else torch.as_tensor(theta, device='cuda')
    _hip.check(_hip.lib().xcape_mlp_loss_grad(_hip._ptr(u), _hip._ptr(t), u.shape[0],
                                              _hip._ptr(th), _hip._ptr(g), _hip._ptr(loss),
                                              _stream()))
    return loss, g


def fit(train, val, epochs=12, batch=16384, lr=3e-3, seed=0, log=None):
    """Train on (n, >=4) sample tensors with the library's kernels (Adam, cosine
    learning-rate decay).  Returns a Surrogate with its training record."""
    import torch

    t0 = time.perf_counter()
    x, y = train[:, :3], train[:, 3]
    mean = x.mean(0)
    scale = 1.0 / x.std(0).clamp_min(1e-12)
    out_offset = float(y.mean())
    out_scale = float(y.std().clamp_min(1e-12))
    norm = np.concatenate([mean.cpu().numpy(), scale.cpu().numpy(),
                           [out_scale, out_offset]]).astype(np.float32)
    nm = torch.as_tensor(norm, device='cuda')

    def prep(d):
        u = ((d[:, :3] - nm[:3]) * nm[3:6]).contiguous()
        t = ((d[:, 3] - nm[7]) / nm[6]).contiguous()
        return u, t
    u, t = prep(train)
    uv, tv = prep(val)
    theta = torch.as_tensor(init_params(seed), device='cuda')
    m = torch.zeros_like(theta)
    v = torch.zeros_like(theta)
    gen = torch.Generator(device='cuda')
    gen.manual_seed(seed)
    n = u.shape[0]
    steps_per_epoch = max(1, n // batch)
    total = epochs * steps_per_epoch
    step = 0
    history = []
    for ep in range(epochs):
        perm = torch.randperm(n, generator=gen, device='cuda')
        for b in range(steps_per_epoch):
            idx = perm[b * batch:(b + 1) * batch]
            ub, tb = u[idx].contiguous(), t[idx].contiguous()
            _, g = loss_grad(theta, ub, tb)
            step += 1
            lr_t = lr * 0.5 * (1.0 + np.cos(np.pi * step / total))
            _hip.check(_hip.lib().xcape_mlp_adam(_hip._ptr(theta), _hip._ptr(g), _hip._ptr(m),
                                                 _hip._ptr(v), step, lr_t, 0.9, 0.999, 1e-8,
                                                 0.0, _stream()))
        tl, _ = loss_grad(theta, u[:65536].contiguous(), t[:65536].contiguous())
        vl, _ = loss_grad(theta, uv, tv)
        rec = {'epoch': ep + 1, 'train_mse': float(tl.item()), 'val_mse': float(vl.item())}
        history.append(rec)
        if log:
            log(rec)
    torch.cuda.synchronize()
    meta = {'trained_on': 'GPU (xcape_mlp_loss_grad / xcape_mlp_adam)',
            'train_samples': int(n), 'val_samples': int(uv.shape[0]), 'epochs': epochs,
            'batch': batch, 'lr': lr, 'seed': seed, 'history': history,
            'seconds': time.perf_counter() - t0}
    return Surrogate(theta.cpu().numpy(), norm, meta)


def train(train_columns=10_000_000, val_columns=1_000_000, epochs=8, batch=16384, lr=3e-3,
          seed=2024, log=print):
    """The warm-start surrogate of the ragged-batch recipe: first sub-steps of
    `train_columns` ascents (seed), validated on `val_columns` ascents of
    another seed.  Returns the Surrogate with its record in .meta."""
    t0 = time.perf_counter()
    tr = log_samples(train_columns, seed)
    va = log_samples(val_columns, seed + 1)
    if log:
        log(f'[surrogate] {tr.shape[0]} train / {va.shape[0]} validation samples '
            f'logged in {time.perf_counter() - t0:.1f} s')
    s = fit(tr, va, epochs=epochs, batch=batch, lr=lr, seed=seed, log=log)
    err = guess_error(s, va)
    import torch
    q = torch.quantile(err[:1_000_000].double(), torch.tensor([0.5, 0.9, 0.99], dtype=torch.float64,
                                                           device=err.device))
    s.meta.update({'recipe': 'first sub-steps of cold-start ascents, ragged-batch columns '
                             '(nlev 30/40/60/91/137, SB/ML/MU, pseudo-liquid/pseudo-ice, '
                             'pinc 500 Pa)',
                   'train_columns': train_columns, 'val_columns': val_columns,
                   'val_guess_error_K': {'p50': float(q[0]), 'p90': float(q[1]),
                                         'p99': float(q[2])}})
    if log:
        log(f'[surrogate] validation first-iterate error [K]: {s.meta["val_guess_error_K"]}')
    return s


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description='train the warm-start surrogate on the GPU')
    ap.add_argument('--out', default=DEFAULT_PATH)
    ap.add_argument('--train-columns', type=int, default=10_000_000)
    ap.add_argument('--val-columns', type=int, default=1_000_000)
    ap.add_argument('--epochs', type=int, default=8)
    ap.add_argument('--seed', type=int, default=2024)
    args = ap.parse_args(argv)
    s = train(args.train_columns, args.val_columns, epochs=args.epochs, seed=args.seed)
    s.save(args.out)
    print(f'[surrogate] saved {args.out}')


if __name__ == '__main__':
    main()
