"""CPU oracle of the warm-start MLP -- TEST INFRASTRUCTURE ONLY.

Float64 numpy forward and backward of the 3 -> 64 -> 64 -> 1 network with
tanh-form GELU activations, used to check the GPU training kernels'
gradients and the in-kernel inference.  The GPU computes with BF16 weights
and BF16 hidden activations feeding FP32-accumulated products; ``bf16=True``
applies the same rounding here (round to nearest even on the FP32 value), so
the comparison isolates the kernels' arithmetic from the quantisation.

Parameter vector layout (float32/float64, length N_PARAMS), shared with the
library: W1 (64x3, row-major), b1 (64), W2 (64x64, row-major: W2[i, j] maps
hidden-1 unit j to hidden-2 unit i), b2 (64), w3 (64), b3 (1).
Loss: mean squared error of the scalar output.
"""
import numpy as np

H = 64
N_IN = 3
N_PARAMS = H * N_IN + H + H * H + H + H + 1
_C = np.sqrt(2.0 / np.pi)
_A = 0.044715


def unpack(theta):
    theta = np.asarray(theta, np.float64)
    o = 0
    W1 = theta[o:o + H * N_IN].reshape(H, N_IN); o += H * N_IN
    b1 = theta[o:o + H]; o += H
    W2 = theta[o:o + H * H].reshape(H, H); o += H * H
    b2 = theta[o:o + H]; o += H
    w3 = theta[o:o + H]; o += H
    b3 = theta[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, np.atleast_1d(b3)])


def bf16(x):
    """Round float64 values through float32 to bfloat16 (nearest even)."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    u = _C * (z + _A * z ** 3)
    return 0.5 * z * (1.0 + np.tanh(u))


def gelu_grad(z):
    u = _C * (z + _A * z ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * _C * (1.0 + 3.0 * _A * z * z)


def forward(theta, x, bf16_round=False):
    """Returns y (n,) and the cache for backward."""
    W1, b1, W2, b2, w3, b3 = unpack(theta)
    x = np.asarray(x, np.float64)
    if bf16_round:
        W1, W2, w3 = bf16(W1), bf16(W2), bf16(w3)
    z1 = x @ W1.T + b1
    h1 = gelu(z1)
    h1q = bf16(h1) if bf16_round else h1
    z2 = h1q @ W2.T + b2
    h2 = gelu(z2)
    y = h2 @ w3 + b3
    return y, (x, z1, h1q, z2, h2, W1, W2, w3)


def loss_and_grad(theta, x, t, bf16_round=False):
    """MSE loss and its gradient w.r.t. the parameter vector.  With bf16
    rounding the gradient is the straight-through one (the rounding is treated
    as identity), which is what the GPU kernels compute."""
    y, (x, z1, h1q, z2, h2, W1, W2, w3) = forward(theta, x, bf16_round)
    n = len(y)
    r = y - np.asarray(t, np.float64)
    loss = np.mean(r * r)
    dy = 2.0 * r / n
    dw3 = h2.T @ dy
    db3 = dy.sum()
    dz2 = (dy[:, None] * w3[None, :]) * gelu_grad(z2)
    dW2 = dz2.T @ h1q
    db2 = dz2.sum(0)
    dh1 = dz2 @ W2
    dz1 = dh1 * gelu_grad(z1)
    dW1 = dz1.T @ x
    db1 = dz1.sum(0)
    return loss, pack(dW1, db1, dW2, db2, dw3, db3)


def init(seed=0):
    """He-style initialisation (the GPU trainer starts from the same vector)."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2.0 / N_IN), (H, N_IN))
    W2 = rng.normal(0, np.sqrt(2.0 / H), (H, H))
    w3 = rng.normal(0, np.sqrt(1.0 / H), H)
    return pack(W1, np.zeros(H), W2, np.zeros(H), w3, 0.0)
