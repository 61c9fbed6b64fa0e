"""CPU oracle of the learned ground network: bfloat16 rounding against known answers and
torch, the backward pass against finite differences, and the pole-fit helpers on inputs with
known answers.  No GPU."""
import os
import sys

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402


def test_bf16_rounding_known_answers():
    u = 2.0 ** -7   # bfloat16 spacing at 1
    x = np.array([1.0, 1 + u / 2, 1 + 1.5 * u, 1 + 0.75 * u, -1 - 0.25 * u, 3.0e38, 1e-40])
    y = mo.bf16(x)
    assert y[0] == 1.0
    assert y[1] == 1.0                 # tie -> even mantissa
    assert y[2] == 1 + 2 * u           # tie -> even mantissa
    assert y[3] == 1 + u
    assert y[4] == -1.0
    assert np.isfinite(y[5])
    rng = np.random.default_rng(3)
    r = (rng.standard_normal(20000) * 10.0 ** rng.uniform(-6, 6, 20000)).astype(np.float32)
    ref = torch.from_numpy(r).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16(r), ref)


def _params(seed):
    rng = np.random.default_rng(seed)
    n = sum(o * i + o for i, o in zip(mo.DIN, mo.DOUT))
    w = np.zeros(n)
    off = 0
    for l, (i, o) in enumerate(zip(mo.DIN, mo.DOUT)):
        ri = mo.N_IN if l == 0 else i
        ro = mo.N_OUT if l == len(mo.DIN) - 1 else o
        W = np.zeros((o, i)); W[:ro, :ri] = rng.uniform(-1, 1, (ro, ri)) * np.sqrt(6.0 / ri)
        w[off:off + o * i] = W.ravel(); off += o * i
        w[off:off + ro] = rng.uniform(-0.1, 0.1, ro); off += o
    return w


def test_backward_matches_finite_differences():
    """Central differences of the float64 loss (no bfloat16 rounding) against the oracle's
    backward pass, over a sample of parameters of every layer, weights and biases."""
    rng = np.random.default_rng(1)
    w = _params(2)
    x = rng.uniform(-1, 1, (23, mo.N_IN))
    t = rng.uniform(-1, 1, (23, mo.N_OUT))
    g, mse = mo.gradient(w, x, t, rounding=False)
    assert mse == pytest.approx(mo.loss(w, x, t, rounding=False), rel=1e-12)
    idx = []
    off = 0
    for l, (i, o) in enumerate(zip(mo.DIN, mo.DOUT)):
        ri = mo.N_IN if l == 0 else i
        ro = mo.N_OUT if l == len(mo.DIN) - 1 else o
        for _ in range(12):
            idx.append(off + int(rng.integers(ro)) * i + int(rng.integers(ri)))
        off += o * i
        idx += [off + int(rng.integers(ro)) for _ in range(4)]
        off += o
    h = 1e-6
    for k in idx:
        wp, wm = w.copy(), w.copy()
        wp[k] += h; wm[k] -= h
        fd = (mo.loss(wp, x, t, False) - mo.loss(wm, x, t, False)) / (2 * h)
        assert fd == pytest.approx(g[k], rel=1e-5, abs=1e-9), k


def test_padded_parameters_have_zero_gradient():
    rng = np.random.default_rng(4)
    w = _params(5)
    x = rng.uniform(-1, 1, (16, mo.N_IN)); t = rng.uniform(-1, 1, (16, mo.N_OUT))
    g, _ = mo.gradient(w, x, t)
    L = mo.unpack(g)
    assert not L[0][0][:, mo.N_IN:].any()          # padded inputs
    assert not L[-1][0][mo.N_OUT:, :].any() and not L[-1][1][mo.N_OUT:].any()  # padded outputs


def test_pole_fit_recovers_known_residues(lee):
    ground = lee.ground
    rng = np.random.default_rng(6)
    a = rng.uniform(0, 2, ground.K) * (rng.uniform(size=ground.K) > 0.3)
    zinf, tau, z0 = 900.0, 2e-4, 415.0
    f = np.linspace(5, 3000, 2000)
    Z = ground.impedance(f, a, zinf, tau)
    got, err = ground.fit_residues(f, Z, zinf, tau, z0)
    assert err < 1e-6
    assert np.allclose(got, a, atol=1e-6)


def test_reflection_of_a_scaled_echo(lee):
    """Receiver series = incident pulse + c x the rigid-floor echo gives R = c at every
    frequency (the rigid-floor run normalises the air path)."""
    ground = lee.ground
    dt, c = 1e-4, 343.0
    n = 2000
    t = np.arange(n) * dt
    z_src, z_rec = 6.0, 3.0
    pulse = lambda t0: np.exp(-((t - t0) / 4e-4) ** 2)
    inc = pulse((z_src - z_rec) / c)
    echo = pulse((z_src + z_rec) / c) * 0.9
    f, R = ground.reflection(inc - 0.4 * echo, inc + echo, dt, c, z_src, z_rec)
    band = (f > 50) & (f < 1000)
    assert np.allclose(R[band], -0.4, atol=1e-9)
