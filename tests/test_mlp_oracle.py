"""Pins of the surrogate MLP's CPU oracle (oracle/mlp_oracle.py) and of the
host-side surrogate code, without a GPU.

The oracle's backward is pinned by central finite differences of its own
forward in float64 (a dropped term or a transposed weight fails them), its
BF16 rounding and GELU against torch's independent implementations, and its
forward against a torch float64 evaluation of the same layers.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..'))
from oracle import mlp_oracle as mo  # noqa: E402


def _problem(seed=0, n=40):
    rng = np.random.default_rng(seed)
    theta = mo.init(seed)
    # non-zero biases so that their gradients are exercised too
    W1, b1, W2, b2, w3, b3 = mo.unpack(theta)
    theta = mo.pack(W1, rng.normal(0, .3, 64), W2, rng.normal(0, .3, 64), w3, 0.2)
    x = rng.normal(0, 1, (n, 3))
    t = rng.normal(0, 1, n)
    return theta, x, t


def test_param_count_and_layout():
    assert mo.N_PARAMS == 4481
    theta = np.arange(mo.N_PARAMS, dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = mo.unpack(theta)
    assert W1[1, 0] == 3 and b1[0] == 192 and W2[0, 1] == 257 and b2[0] == 256 + 4096
    assert w3[0] == 256 + 4096 + 64 and b3 == 4480
    np.testing.assert_array_equal(mo.pack(W1, b1, W2, b2, w3, b3), theta)


def test_backward_matches_finite_differences():
    theta, x, t = _problem()
    loss, g = mo.loss_and_grad(theta, x, t)
    rng = np.random.default_rng(1)
    # every bias and b3, plus a random sample of weights from each block
    idx = np.concatenate([np.arange(192, 256), np.arange(4352, 4481),
                          rng.choice(192, 24, replace=False),
                          256 + rng.choice(4096, 64, replace=False)])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(theta)
        e[i] = h
        lp = mo.loss_and_grad(theta + e, x, t)[0]
        lm = mo.loss_and_grad(theta - e, x, t)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])
    assert loss > 0


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip('torch')
    rng = np.random.default_rng(2)
    v = np.concatenate([rng.normal(0, 1, 10000), rng.normal(0, 1e3, 1000),
                        [0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8,
                         1.0 + 2.0 ** -9, 65504.0]]).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mo.bf16(v), ref)
    # ties go to even
    assert mo.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    assert mo.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6


def test_gelu_matches_torch_tanh_form():
    torch = pytest.importorskip('torch')
    z = np.linspace(-12, 12, 2001)
    ref = torch.nn.functional.gelu(torch.from_numpy(z), approximate='tanh').numpy()
    np.testing.assert_allclose(mo.gelu(z), ref, rtol=1e-12, atol=1e-14)
    zt = torch.from_numpy(z).requires_grad_(True)
    torch.nn.functional.gelu(zt, approximate='tanh').sum().backward()
    np.testing.assert_allclose(mo.gelu_grad(z), zt.grad.numpy(), rtol=1e-10, atol=1e-12)


def test_forward_matches_torch_layers():
    torch = pytest.importorskip('torch')
    theta, x, _ = _problem(3)
    W1, b1, W2, b2, w3, b3 = (torch.as_tensor(np.atleast_1d(a)) for a in mo.unpack(theta))
    xt = torch.from_numpy(x)
    g = torch.nn.functional.gelu
    y_ref = (g(g(xt @ W1.T + b1, approximate='tanh') @ W2.T + b2, approximate='tanh') @ w3
             + b3).numpy()
    np.testing.assert_allclose(mo.forward(theta, x)[0], y_ref, rtol=1e-12, atol=1e-12)


def test_init_matches_product_initialiser():
    from xcape_amd import surrogate
    np.testing.assert_array_equal(surrogate.init_params(5), mo.init(5).astype(np.float32))


def test_surrogate_file_round_trip(tmp_path):
    from xcape_amd import surrogate
    th = np.random.default_rng(0).normal(0, 1, mo.N_PARAMS).astype(np.float32)
    nm = np.arange(8, dtype=np.float32)
    s = surrogate.Surrogate(th, nm, {'note': 'x'})
    path = str(tmp_path / 's.npz')
    s.save(path)
    r = surrogate.Surrogate.load(path)
    np.testing.assert_array_equal(r.theta, th)
    np.testing.assert_array_equal(r.norm, nm)
    assert r.meta['note'] == 'x' and r.meta['format_version'] == surrogate.FORMAT_VERSION
