"""The learned ground network on the GPU (lee_mlp: BF16 MFMA products, FP32 accumulation)
against the CPU oracle of the same network: forward pass, loss and gradients, with row counts
that are and are not multiples of the 16-row tile; then a short training run."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("rows", [16, 50, 257])
def test_forward_matches_oracle(lee, rows):
    net = lee.ground.Network(seed=3)
    rng = np.random.default_rng(rows)
    w = net.params()
    # non-zero biases so that every term of the layer is exercised
    L = mo.unpack(w.astype(np.float64))
    off = 0
    for l, (i, o) in enumerate(zip(mo.DIN, mo.DOUT)):
        off += o * i
        ro = mo.N_OUT if l == len(mo.DIN) - 1 else o
        w[off:off + ro] = rng.uniform(-0.2, 0.2, ro)
        off += o
    net.set_params(w)
    x = rng.uniform(-1, 1, (rows, mo.N_IN)).astype(np.float32)
    got = net.predict(x)
    ref = mo.predict(w.astype(np.float64), x.astype(np.float64))
    e = rel(got, ref)
    print(f"[gpu mlp] forward rows={rows} rel-L2 {e:.2e}", flush=True)
    net.close()
    # operands are rounded to bfloat16 identically on both sides; what differs is FP32
    # accumulation and GELU (and an operand that lands on the other side of a rounding tie)
    assert e < 5e-3


@pytest.mark.parametrize("rows", [32, 45])
def test_gradient_matches_oracle(lee, rows):
    net = lee.ground.Network(seed=5)
    rng = np.random.default_rng(100 + rows)
    x = rng.uniform(-1, 1, (rows, mo.N_IN)).astype(np.float32)
    t = rng.uniform(0, 1, (rows, mo.N_OUT)).astype(np.float32)
    g, mse = net.gradient(x, t)
    w = net.params().astype(np.float64)
    gr, mr = mo.gradient(w, x.astype(np.float64), t.astype(np.float64))
    assert mse == pytest.approx(mr, rel=5e-3)
    G, Gr = mo.unpack(g.astype(np.float64)), mo.unpack(gr)
    errs = []
    for (gw, gb), (rw, rb) in zip(G, Gr):
        errs.append((rel(gw, rw), rel(gb, rb)))
    print(f"[gpu mlp] gradient rows={rows} per-layer rel-L2 (W, b): " +
          " ".join(f"({a:.1e},{b:.1e})" for a, b in errs), flush=True)
    net.close()
    for a, b in errs:
        assert a < 2e-2 and b < 2e-2
    # padded parameters receive nothing
    assert not G[0][0][:, mo.N_IN:].any() and not G[-1][0][mo.N_OUT:, :].any()


def test_training_reduces_the_loss(lee):
    """A smooth target map of the four inputs (the shape of residues against ground
    parameters): Adam on the GPU takes the mean squared error down by more than 20x."""
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, (512, 4)).astype(np.float32)
    k = np.arange(8)[None, :]
    t = (0.5 + 0.4 * np.sin(1.3 * x[:, :1] * (k + 1) / 4 + x[:, 1:2]) * np.exp(-0.3 * x[:, 2:3] ** 2)
         + 0.1 * x[:, 3:4] * (k % 3)).astype(np.float32)
    net = lee.ground.Network(seed=11)
    m0 = float(np.mean((net.predict(x) - t) ** 2))
    m1 = net.train(x, t, batch=64, epochs=150, lr=3e-3)
    m2 = float(np.mean((net.predict(x) - t) ** 2))
    print(f"[gpu mlp] training mse {m0:.3e} -> {m2:.3e} (last batch {m1:.3e})", flush=True)
    net.close()
    assert np.isfinite(m2) and m2 < m0 / 20
