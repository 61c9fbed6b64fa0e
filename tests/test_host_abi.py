"""Host side of the product: the C ABI library loads and exports what include/lee_abi.h
declares, the drop-in file contract (parser, writers) and deck validation.  No GPU needed:
nothing here launches a kernel."""
import ctypes
import os
import re
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIPPED = os.path.join(ROOT, "tests", "golden", "shipped_deck")


def test_library_exports_every_declared_entry_point(lee):
    L = lee.lib()
    hdr = open(os.path.join(ROOT, "include", "lee_abi.h")).read()
    names = set(re.findall(r"^[a-z][a-z_ ]*\**\s*\**(lee_[a-z_]+)\s*\(", hdr, flags=re.M))
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(L, n), n


def test_parser_reads_the_shipped_deck(lee):
    d = lee.load_deck(SHIPPED)
    try:
        assert d.gauss_width == pytest.approx(0.135955598689175, abs=0)
        assert (d.mf_jmax, d.mf_lead_dj, d.mf_trail_dj, d.mf_judge) == (201, 11, 11, 0)
        assert d.mf_ratio == 0.75 and d.scheme == 2
        assert list(d.bc) == [2, 1, 5, 2, 2, 2]
        assert (d.time_domain, d.out_difft, d.fft_m, d.frequency) == (1600, 100, 19, 300.0)
        assert (d.velocity_coef, d.velocity_method) == (3.0, 4)
        assert (d.gamma, d.Pav, d.c) == (1.4, 1.0e5, 340.0)
        assert list(d.pml_alpha) == [2.0, 2.0, 2.0] and list(d.pml_R) == [1000.0, 1000.0, 1000.0]
        assert d.pml_y_time == 4500
        assert d.pore_q == 1.790799820223793 and d.pore_porosity == 0.8769089596793096 and d.pore_resistivity == 51375.65
        assert list(d.src) == [0.0, -0.325, 0.0] and list(d.rcv) == [0.0, -0.23, 1.75]
        assert d.dt == 4e-6 and (d.dx, d.dy, d.dz) == (0.005, 0.005, 0.005)
        assert list(d.mpi_dims) == [2, 2, 2]
        assert (d.restart, d.restart_out, d.hill_judge) == (0, 5000, 0)
        # obstacle lines 24-29 (airpore.cpp:270-299)
        assert list(d.hill_coef) == [-1.0, 1.0, 0.0, 2.5, -0.2, 0.0, 3.0, 0.0, 0.0, 0.0, 0.0, 0.0]
        assert d.hill_points == 0
        assert [list(r) for r in d.hill_mat] == [[1.790799820223793, 0.8769089596793096, 5.137565e4],
                                                 [1.790799820223793, 0.8769089596793096, 5.137565e4],
                                                 [0.52, 0.5712, 8.0e4]]
        assert list(d.hill_box) == [-1.0, 1.0, 0.0, 2.5, -0.2, 0.0]
        assert list(d.air_hdr) == [141, 41, 41, 201, 41, 0, 441, 0, 41, 201]
        x = np.ctypeslib.as_array(d.air_x, shape=(141,))
        y = np.ctypeslib.as_array(d.air_y, shape=(201,))
        z = np.ctypeslib.as_array(d.air_z, shape=(441,))
        assert x[0] == -0.2 and x[-1] == 0.5 and y[0] == -1.0 and y[-1] == 0.0 and z[0] == 0.0 and z[-1] == 2.2
        assert d.has_pore == 0
    finally:
        lee.free_deck(d)


def test_generated_deck_round_trips_through_the_parser(lee):
    g = lee.deck.box(20, 24, 18, 0.04, pml=5, bc=(2, 3, 5, 2, 2, 2), porous_cells=7, src=(0.1, 0.3, 0.7),
                     velocity_method=5, velocity_coef=0.25, pml_y_time=123, time_domain=321, fft_m=11,
                     mf_judge=1, mf_jmax=24, mf_lead_dj=5, mf_trail_dj=5, ijkmax=60)
    g.pore_q, g.pore_porosity, g.pore_resistivity = 1.4, 0.37, 12345.5
    with tempfile.TemporaryDirectory() as td:
        g.write(td)
        d = lee.load_deck(td)
        try:
            c = g.to_c()
            for name, _ in lee.LeeDeck._fields_:
                if name.endswith(("_x", "_y", "_z")):
                    continue
                a, b = getattr(d, name), getattr(c, name)
                if hasattr(a, "__len__"):
                    assert np.array_equal(np.ctypeslib.as_array(a), np.ctypeslib.as_array(b)), name
                else:
                    assert a == b, name
            for nm, ref, n in (("air_x", g.air_x, 20), ("air_y", g.air_y, 24), ("air_z", g.air_z, 18),
                               ("pore_x", g.pore_x, 20), ("pore_y", g.pore_y, 24), ("pore_z", g.pore_z, 7)):
                assert np.array_equal(np.ctypeslib.as_array(getattr(d, nm), shape=(n,)), ref[:n]), nm
        finally:
            lee.free_deck(d)


def test_parser_reports_missing_files(lee):
    with tempfile.TemporaryDirectory() as td:
        with pytest.raises(lee.LeeError) as e:
            lee.load_deck(td)
        assert e.value.code == -3
        lee.deck.box(8, 8, 8, 0.05).write(td)
        os.remove(os.path.join(td, "coordi_air2.dat"))
        with pytest.raises(lee.LeeError) as e:
            lee.load_deck(td)
        assert "coordi_air2.dat" in str(e.value)


def test_receiver_writer_format(lee):
    rec = np.array([[1.0, -2.5e-7], [0.0, 3.14159265e2], [1e-120, 0.0]])
    with tempfile.TemporaryDirectory() as td:
        lee._check(lee.lib().lee_write_receivers(td.encode(), lee._p(np.ascontiguousarray(rec)), 3, 2))
        a = open(os.path.join(td, "solution", "pty0000.dat")).read().splitlines()
        b = open(os.path.join(td, "solution", "pty0001.dat")).read().splitlines()
        assert a == ["1.000000e+00", "0.000000e+00", "1.000000e-120"]
        assert b == ["-2.500000e-07", "3.141593e+02", "0.000000e+00"]


def test_spectrum_writer_layout(lee):
    rng = np.random.default_rng(0)
    rec = np.ascontiguousarray(rng.standard_normal((100, 2)))
    with tempfile.TemporaryDirectory() as td:
        lee._check(lee.lib().lee_write_spectra(td.encode(), lee._p(rec), 100, 2, 8, 1e-4))
        names = sorted(os.listdir(os.path.join(td, "solution")))
        # the reference always processes five receiver files (airpore.cpp:970-971)
        assert names == [f"p_f2_{i:04d}.dat" for i in range(5)]
        lines = open(os.path.join(td, "solution", "p_f2_0000.dat")).read().splitlines()
        assert len(lines) == 129 and lines[0].startswith("0.0 ")
        f = [float(l.split()[0]) for l in lines]
        assert f[1] == pytest.approx(1 / (1e-4 * 256)) and f[-1] == pytest.approx(128 / (1e-4 * 256))
        mag = np.array([float(l.split()[1]) for l in lines])
        series = np.array([float(f"{v:.6e}") for v in rec[:, 0]])
        ref = np.abs(np.fft.fft(series, 256) / 256)[:129]
        assert np.allclose(mag, ref, rtol=1e-9, atol=1e-15)
        with pytest.raises(lee.LeeError):
            lee._check(lee.lib().lee_write_spectra(td.encode(), lee._p(rec), 100, 2, 6, 1e-4))


@pytest.mark.parametrize("change,code", [
    (dict(scheme=5), -4),       # RK4 is an unimplemented enum in the reference deck path
    (dict(scheme=1), -4),
    (dict(hill_judge=6), -4),   # no obstacle mask of that kind in the reference
    (dict(hill_judge=1, integrator=1), -4),  # obstacles run with the FB_p_v step
    (dict(restart=1), -4),
    (dict(velocity_method=3), -4),
    (dict(bc=(4, 1, 1, 1, 1, 1)), -4),
    (dict(bc=(1, 1, 1, 5, 1, 1)), -4),
    (dict(mf_lead_dj=1), -1),
])
def test_unsupported_decks_are_rejected(lee, change, code):
    g = lee.deck.box(10, 10, 10, 0.05)
    for k, v in change.items():
        setattr(g, k, v)
    with pytest.raises(lee.LeeError) as e:
        lee.Solver(g)
    assert e.value.code == code, str(e.value)


@pytest.mark.parametrize("integrator,height,ok", [(0, 0.3, False), (1, 0.3, True), (1, 0.55, False)])
def test_terrain_time_step_bound(lee, integrator, height, ok):
    """Over a hill the cells shrink by 1 - h/D: the acoustic number must stay below 1 (FB_p_v)
    or sqrt(2) (RK4); beyond, the deck is refused before any device work."""
    g = lee.deck.box(24, 30, 20, 0.04, bc=(1, 1, 1, 1, 1, 1), src=(0.0, 0.5, 0.4), gauss_width=0.24)
    g.hills, g.integrator = [(0.0, 0.6, height, 0.3)], integrator
    if ok:
        return  # accepted decks are exercised on the GPU
    with pytest.raises(lee.LeeError) as e:
        lee.Solver(g)
    assert e.value.code == -1 and "time step too long" in str(e.value)


def test_porous_grid_must_share_the_air_index_space(lee):
    g = lee.deck.box(12, 10, 10, 0.05, bc=(1, 3, 1, 1, 1, 1), porous_cells=6)
    g.pore_x = g.pore_x[:-1]
    with pytest.raises(lee.LeeError) as e:
        lee.Solver(g)
    assert e.value.code == -1 and "index space" in str(e.value)


def test_fft_matches_numpy(lee):
    rng = np.random.default_rng(5)
    x = rng.standard_normal(512)
    re, im = lee.fft(x, np.zeros(512), 9)
    ref = np.fft.fft(x) / 512
    assert np.abs(re - ref.real).max() < 1e-13 and np.abs(im - ref.imag).max() < 1e-13
