"""The CPU oracle against the reference solver's own compiled code.

oracle/_ref/ref_probe is built (oracle/Makefile) from the reference sources air.cpp,
porous.cpp, FFT.cpp, df_over_var.cpp and decomp.cpp and calls their member functions
directly.  Every comparison here is bit-for-bit.  Skipped where the probe is not built
(it needs the reference sources and the image's MPICH).
"""
import os
import tempfile

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.skipif(not O.probe_available(), reason="reference probe not built (make -C oracle ref)")

SLABS = ["x1", "x2", "y1", "y2", "z1", "z2"]
FIELDS = ["u", "v", "w", "p"]


def _fields(shape, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape) for _ in range(4)], [rng.standard_normal(shape) for _ in range(4)]


def _interior(shape):
    m = np.zeros(shape, bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


CONFIGS = {
    # name: (box kwargs, frame, gate counter)
    "all_absorbing_gate_on": (dict(pml=5, bc=(2, 2, 2, 2, 2, 2), pml_y_time=3), 0, 5),
    "all_absorbing_gate_off": (dict(pml=5, bc=(2, 2, 2, 2, 2, 2), pml_y_time=50), 0, 5),
    "shipped_bc_set": (dict(pml=5, bc=(2, 1, 5, 2, 2, 2), pml_y_time=4500), 0, 0),
    "x1_z2_only": (dict(pml=4, bc=(2, 1, 1, 1, 1, 2)), 0, 0),
    "x2_y2_only": (dict(pml=4, bc=(1, 1, 2, 1, 2, 1)), 0, 0),
    "wind1": (dict(pml=4, bc=(2, 1, 2, 2, 2, 2), velocity_method=1, velocity_coef=2.0, pml_y_time=1), 0, 0),
    "wind2": (dict(pml=4, bc=(2, 2, 2, 2, 2, 2), velocity_method=2, velocity_coef=1.0, pml_y_time=1), 0, 3),
    "wind5": (dict(pml=4, bc=(2, 2, 2, 2, 2, 2), velocity_method=5, velocity_coef=3.0, pml_y_time=1), 0, 3),
    "stretched": (dict(pml=4, bc=(2, 2, 2, 2, 2, 2), pml_y_time=1, velocity_method=5, velocity_coef=0.5), 0, 2),
}


def _deck(lee, name):
    kw, frame, counter = CONFIGS[name]
    d = lee.deck.box(18, 20, 15, 0.05, **kw)
    if name == "stretched":  # geometrically stretched axes (metrics differ point to point)
        d.air_x = 0.05 * (1.03 ** np.arange(18) - 1) / 0.03
        d.air_y = 0.05 * (1.02 ** np.arange(20) - 1) / 0.02
        d.air_z = 0.05 * (1.04 ** np.arange(15) - 1) / 0.04
    return d, frame, counter


@pytest.mark.parametrize("name", list(CONFIGS))
def test_air_stencils_pml_wind_gate_bitwise(lee, name):
    d, frame, counter = _deck(lee, name)
    shape = (15, 20, 18)
    f1, f2 = _fields(shape, hash(name) % 1000)
    res = O.run_probe(d, f1, f2, frame=frame, counter=counter)
    o = O.Oracle(d)
    o.frame_setup(0)
    for s in range(6):
        for f in range(4):
            q = res[f"q0_{SLABS[s]}_{FIELDS[f]}"]
            assert q.size == o.q_get(s, f).size, (SLABS[s], q.size)
            o.q_set(s, f, q)
    U, V, W = o.wind()
    assert np.array_equal(U.ravel(), res["U"]) and np.array_equal(V.ravel(), res["V"]) and np.array_equal(W.ravel(), res["W"])
    for _ in range(counter + 1):
        o.tick_gate()
    m = _interior(shape)
    dp = o.air_dp(*f1)
    assert np.array_equal(dp[m], res["dp"].reshape(shape)[m])
    for a, key in zip(o.air_duvw(*f1), ("du", "dv", "dw")):
        assert np.array_equal(a[m], res[key].reshape(shape)[m]), key
    o.update_q(0, p=f2[3])
    for s in range(6):
        assert np.array_equal(o.q_get(s, 3), res[f"qp_{SLABS[s]}_p"]), SLABS[s]
    o.update_q(1, u=f2[0], v=f2[1], w=f2[2])
    for s in range(6):
        for f in range(3):
            assert np.array_equal(o.q_get(s, f), res[f"qvw_{SLABS[s]}_{FIELDS[f]}"]), (SLABS[s], f)


@pytest.mark.parametrize("wind", [4, 2, 5])
def test_moving_frame_shift_and_wind_bitwise(lee, wind):
    d = lee.deck.box(16, 26, 14, 0.05, pml=4, bc=(2, 2, 2, 2, 2, 2), pml_y_time=2, velocity_method=wind,
                     velocity_coef=0.8, mf_judge=1, mf_jmax=26, mf_lead_dj=6, mf_trail_dj=6, ijkmax=66)
    shape = (14, 26, 16)
    f1, f2 = _fields(shape, 11 + wind)
    res = O.run_probe(d, f1, f2, frame=3, counter=4)
    o = O.Oracle(d)
    o.frame_setup(0)
    for s in range(6):
        for f in range(4):
            o.q_set(s, f, res[f"q0_{SLABS[s]}_{FIELDS[f]}"])
    for _ in range(4):
        o.tick_gate()
    o.frame_setup(3)
    for s in range(6):
        for f in range(4):
            assert np.array_equal(o.q_get(s, f), res[f"qshift_{SLABS[s]}_{FIELDS[f]}"]), (SLABS[s], f)
    U, V, W = o.wind()
    assert np.array_equal(U.ravel(), res["U"]) and np.array_equal(V.ravel(), res["V"]) and np.array_equal(W.ravel(), res["W"])
    o.tick_gate()
    m = _interior(shape)
    assert np.array_equal(o.air_dp(*f1)[m], res["dp"].reshape(shape)[m])
    for a, key in zip(o.air_duvw(*f1), ("du", "dv", "dw")):
        assert np.array_equal(a[m], res[key].reshape(shape)[m])


@pytest.mark.parametrize("ground", [(1.7320508075688772, 0.3, 2.0e5), (1.2, 0.45, 3.0e4), (2.0, 0.2, 2.0e6)])
def test_porous_stencils_bitwise(lee, ground):
    d = lee.deck.box(14, 12, 10, 0.05, bc=(1, 3, 1, 1, 1, 1), porous_cells=10)
    d.pore_z = d.air_z.copy()
    d.pore_dz = d.dz
    d.pore_q, d.pore_porosity, d.pore_resistivity = ground
    shape = (10, 12, 14)
    f1, f2 = _fields(shape, 5)
    res = O.run_probe(d, f1, f2, pore=ground)
    o = O.Oracle(d)
    o.frame_setup(0)
    m = _interior(shape)
    assert np.array_equal(o.pore_dp(*f1[:3])[m], res["pore_dp"].reshape(shape)[m])
    for a, key in zip(o.pore_duvw(*f1), ("pore_du", "pore_dv", "pore_dw")):
        assert np.array_equal(a[m], res[key].reshape(shape)[m])


def test_fft_and_spectrum_writer_bitwise(lee):
    d = lee.deck.box(10, 10, 10, 0.05)
    shape = (10, 10, 10)
    f1, f2 = _fields(shape, 1)
    rng = np.random.default_rng(9)
    m = 10
    re = np.zeros(1 << m)
    re[:700] = rng.standard_normal(700) * 1e-3
    im = np.zeros(1 << m)
    with tempfile.TemporaryDirectory() as td:
        ref_file = os.path.join(td, "p_f2_ref.dat")
        res = O.run_probe(d, f1, f2, fft_m=m, fft_re=re, fft_im=im, fft_file=ref_file)
        x, y = O.fft(re, im, m)
        assert np.array_equal(x, res["fft_re"]) and np.array_equal(y, res["fft_im"])
        xi, yi = O.fft(re, im, m, direction=-1)
        assert np.array_equal(xi, res["ifft_re"]) and np.array_equal(yi, res["ifft_im"])
        # the product FFT is the same algorithm: bit-for-bit
        px, py = lee.fft(re, im, m)
        assert np.array_equal(px, x) and np.array_equal(py, y)
        # the product spectrum writer against FFT_output(): feed a series whose 6-digit text
        # form is exact, so the writer's text round trip does not change it
        series = np.round(re[:700] * 1e6) / 1e6
        s2 = np.zeros(1 << m)
        s2[:700] = series
        ref2 = os.path.join(td, "p_f2_ref2.dat")
        O.run_probe(d, f1, f2, fft_m=m, fft_re=s2, fft_im=np.zeros(1 << m), fft_file=ref2)
        outdir = os.path.join(td, "out")
        os.makedirs(outdir)
        lee._check(lee.lib().lee_write_spectra(outdir.encode(), lee._p(np.ascontiguousarray(series.reshape(-1, 1))),
                                               700, 1, m, d.dt))
        with open(ref2) as a, open(os.path.join(outdir, "solution", "p_f2_0000.dat")) as b:
            assert a.read() == b.read()


def test_geo_over_var_and_decomposition(lee):
    d = lee.deck.box(10, 10, 10, 0.05)
    f1, f2 = _fields((10, 10, 10), 2)
    res = O.run_probe(d, f1, f2)
    g = res["geo"].reshape(-1, 6)
    a, b, c, e, j = g[:, 0], g[:, 1], g[:, 2], g[:, 3], g[:, 4]
    assert np.array_equal(g[:, 5], (a * b - c * e) * j)
    dec = res["decomp"].reshape(-1, 5).astype(int)
    for n, npr, myid, s, e in dec:
        # textbook MPE block split: the first n % np ranks get one extra point
        base, extra = divmod(n, npr)
        s_ref = myid * base + min(myid, extra) + 1
        e_ref = s_ref + base + (1 if myid < extra else 0) - 1
        if myid == npr - 1:
            e_ref = n
        assert (s, e) == (s_ref, min(e_ref, n) if myid != npr - 1 else n)
