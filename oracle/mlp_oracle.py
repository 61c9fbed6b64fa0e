"""CPU oracle of the learned ground model's network (TEST INFRASTRUCTURE ONLY: tests/ use it
to check the GPU forward pass and gradients; the product trains and runs on the GPU).

The network of lee_mlp (3d_mpi_test_amd/csrc/lee_mlp.hip): layers 16->64, 64->64, 64->64,
64->16 over padded shapes (4 real inputs, 8 real outputs), GELU (tanh form) after the hidden
layers, a linear output, mean squared error over the real outputs.  Products use operands
rounded to bfloat16 (round to nearest even) and accumulate in float64 here (float32 on the
GPU), so the two agree to the bfloat16 rounding of the operands.
"""
import numpy as np

DIN = (16, 64, 64, 64)
DOUT = (64, 64, 64, 16)
N_IN, N_OUT = 4, 8


def bf16(x):
    """Round float32 values to bfloat16 (nearest even; a NaN stays a quiet NaN), returned as
    float64."""
    u32 = np.asarray(x, dtype=np.float32).view(np.uint32)
    u = u32.astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    nan = (u32 & 0x7FFFFFFF) > 0x7F800000
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(w):
    """Flat parameters -> [(W[out][in], b[out])] per layer."""
    out, off = [], 0
    for i, o in zip(DIN, DOUT):
        W = w[off:off + o * i].reshape(o, i); off += o * i
        b = w[off:off + o]; off += o
        out.append((W.astype(np.float64), b.astype(np.float64)))
    return out


def gelu(x):
    c = np.sqrt(2.0 / np.pi)
    return 0.5 * x * (1.0 + np.tanh(c * (x + 0.044715 * x ** 3)))


def gelu_d(x):
    c = np.sqrt(2.0 / np.pi)
    th = np.tanh(c * (x + 0.044715 * x ** 3))
    return 0.5 * (1 + th) + 0.5 * x * (1 - th ** 2) * c * (1 + 3 * 0.044715 * x ** 2)


def _id(x):
    return np.asarray(x, dtype=np.float64)


def forward(w, x, rounding=True):
    rnd = bf16 if rounding else _id
    L = unpack(w)
    R = x.shape[0]
    a = np.zeros((R, DIN[0])); a[:, :N_IN] = x
    zs, acts = [], [a]
    for l, (W, b) in enumerate(L):
        z = rnd(acts[-1]) @ rnd(W).T + b
        zs.append(z)
        if l < len(L) - 1:
            acts.append(gelu(z))
    return zs, acts


def predict(w, x, rounding=True):
    zs, _ = forward(w, x, rounding)
    return zs[-1][:, :N_OUT]


def loss(w, x, t, rounding=True):
    return float(np.mean((predict(w, x, rounding) - t) ** 2))


def gradient(w, x, t, rounding=True):
    """(flat gradient, mean squared error) of the loss mean((y - t)^2) over the real outputs.
    rounding=False drops the bfloat16 operand rounding (exact float64 products), which makes
    the loss differentiable so that finite differences check this backward pass."""
    rnd = bf16 if rounding else _id
    zs, acts = forward(w, x, rounding)
    L = unpack(w)
    R = x.shape[0]
    y = zs[-1]
    d = np.zeros_like(y)
    d[:, :N_OUT] = 2.0 * (y[:, :N_OUT] - t) / (R * N_OUT)
    mse = float(np.mean((y[:, :N_OUT] - t) ** 2))
    grads = [None] * len(L)
    for l in range(len(L) - 1, -1, -1):
        dz = d if l == len(L) - 1 else d * gelu_d(zs[l])
        prev = acts[l]
        gW = rnd(dz).T @ rnd(prev)
        gb = dz.sum(axis=0)
        grads[l] = (gW, gb)
        if l > 0:
            d = rnd(dz) @ rnd(L[l][0])
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return flat, mse
