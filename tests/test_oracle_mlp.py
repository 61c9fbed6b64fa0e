"""CPU oracle of the learned ground network: bfloat16 rounding against known answers and
torch, the backward pass against finite differences, and the pole-fit helpers on inputs with
known answers.  No GPU."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo


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
    # NaN (quiet and signalling, either sign) stays NaN; infinities stay infinite
    bits = np.array([0x7FC00000, 0x7F800001, 0xFFBFFFFF, 0x7FFFFFFF, 0x7F800000, 0xFF800000], dtype=np.uint32)
    z = mo.bf16(bits.view(np.float32))
    assert np.isnan(z[:4]).all() and z[4] == np.inf and z[5] == -np.inf


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


def test_moebius_map_and_residue_fit_recover_known_answers(lee):
    """A reflection built from a known Moebius map of the discrete impedance of known
    residues: the map is identified from four (Z_d, R) pairs, and the fit returns the
    residues (the model class contains the answer, so the error vanishes)."""
    G = lee.ground
    rng = np.random.default_rng(6)
    f = np.linspace(G.F_FIT[0], G.F_FIT[1], 300)
    dt, zinf, tau = 5.8e-5, 1200.0, 3e-4
    co = rng.standard_normal((1, 4)) + 1j * rng.standard_normal((1, 4))
    co = np.repeat(co, len(f), axis=0) * (1 + 0.1 * np.sin(f / 100.0))[:, None]
    Zs, Rs = [], []
    for zc, a in ((415.0, np.zeros(8)), (1660.0, np.zeros(8)), (6640.0, np.zeros(8)), (1660.0, np.eye(8)[3])):
        Z = G.discrete_impedance(f, a, zc, tau, dt)
        Zs.append(Z); Rs.append(G.moebius_apply(co, Z))
    got = G.moebius(Zs, Rs)
    Zt = G.discrete_impedance(f, rng.uniform(0, 1, 8), 900.0, tau, dt)
    assert np.allclose(G.moebius_apply(got, Zt), G.moebius_apply(co, Zt), rtol=1e-9, atol=1e-12)
    a_true = rng.uniform(0, 2, 8) * (rng.uniform(size=8) > 0.3)
    R = G.moebius_apply(co, G.discrete_impedance(f, a_true, zinf, tau, dt))
    ref = G.Reflection(G.Ground(1e5, 0.3, 2.0, 8), f, R, got, zinf, tau, dt)
    a, err = ref.fit()
    assert err < 1e-8 and np.allclose(a, a_true, atol=1e-6)


def test_discrete_impedance_is_the_recursion(lee):
    """Z_d of discrete_impedance against a time-domain run of Cal_TDBC / Update_Psi with a
    sinusoidal velocity: p_n / Zinf = v_n (1 + tb sum a) + sum_k e^{-r_k tb} Psi_k."""
    G = lee.ground
    dt, zinf, tau = 5.8e-5, 900.0, 2e-4
    a = np.array([0.3, 0.0, 0.5, 0.0, 1.2, 0.0, 0.2, 0.7])
    tb = dt / tau
    dec = np.exp(-G.R_GRID * tb)
    f0 = 400.0
    n = np.arange(50)
    v = np.exp(2j * np.pi * f0 * n * dt)
    # histories started in their periodic state (Psi = a tb v / (e^{i w dt} - e^{-r tb}))
    psi = a * tb * v[0] / (np.exp(2j * np.pi * f0 * dt) - dec)
    p = np.zeros(len(n), complex)
    for m in range(len(n)):
        p[m] = zinf * (v[m] * (1 + tb * a.sum()) + (dec * psi).sum())
        psi = dec * psi + a * tb * v[m]
    Zd = G.discrete_impedance(np.array([f0]), a, zinf, tau, dt)[0]
    assert np.abs(p / v - Zd).max() < 1e-9 * abs(Zd)


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


def test_il_errors(lee):
    G = lee.ground
    rng = np.random.default_rng(8)
    a = rng.standard_normal((4096, 3)) * np.hanning(4096)[:, None]
    n, t = G.il_errors(a, a, 1e-4)
    assert n == 0.0 and t == 0.0
    n, t = G.il_errors(a, 1.1 * a, 1e-4)
    assert n == pytest.approx(20 * np.log10(1.1), rel=1e-9) and t == pytest.approx(20 * np.log10(1.1), rel=1e-9)
    # a run that blew up (inf, then nan) matches nothing, rather than nothing to compare
    b = a.copy()
    b[3000:, 1] = np.inf
    assert G.il_errors(a, b, 1e-4) == (float("inf"), float("inf"))
    assert G.il_errors(b * 0.0, a, 1e-4) == (float("inf"), float("inf"))
