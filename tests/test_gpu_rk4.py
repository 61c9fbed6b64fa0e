"""RK4, terrain-following grid and log wind on the GPU against the FP64 oracle, through the
C ABI.  These paths extend the reference (it has no RK4 and a separable grid only); the
oracle's versions are pinned by tests/test_oracle_rk4_terrain.py.  The FP64 device path
keeps the oracle's operation order, so the bar is 1e-10 relative L2 (differences, if any,
are rounding-level).  FP32 with the compensated accumulation is compared with FP64.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REC = [(0.0, 0.55, 0.45), (0.15, 0.75, 0.3), (-0.2, 0.35, 0.6), (0.05, 1.05, 0.2)]
HILLS = [(0.1, 0.7, 0.2, 0.3), (-0.3, 1.1, 0.12, 0.25)]


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def cases(lee):
    B = lee.deck.box
    out = {}
    out["pml"] = B(26, 30, 22, 0.05, pml=5, bc=(2, 1, 2, 2, 2, 2), src=(0.1, 0.7, 0.5), gauss_width=0.25,
                   time_domain=30, pml_y_time=3)
    out["porous"] = B(24, 26, 20, 0.05, pml=5, bc=(2, 3, 2, 2, 2, 2), porous_cells=8, src=(0.0, 0.6, 0.3),
                      gauss_width=0.25, time_domain=30)
    out["wind5"] = B(24, 28, 20, 0.05, pml=5, bc=(2, 1, 2, 2, 2, 2), src=(0.0, 0.6, 0.5), gauss_width=0.25,
                     time_domain=30, velocity_method=5, velocity_coef=3.0)
    t = B(26, 32, 22, 0.05, pml=5, bc=(2, 3, 2, 2, 2, 2), porous_cells=8, src=(0.0, 0.6, 0.3), gauss_width=0.25,
          time_domain=30, velocity_method=7)
    t.hills, t.wind_ustar, t.wind_z0 = HILLS, 0.4, 0.02
    out["terrain"] = t
    # x layers of 5 points on a grid 10 points wide: the slabs x1 and x2 share the points i = 4, 5
    out["overlap"] = B(10, 30, 22, 0.05, pml=5, bc=(2, 1, 2, 2, 2, 2), src=(0.0, 0.7, 0.5), gauss_width=0.25,
                       time_domain=30)
    # 70 points across x: two x tiles of 64 points, the second one partial, over a porous ground
    out["wide"] = B(70, 24, 16, 0.05, pml=5, bc=(2, 3, 2, 2, 2, 2), porous_cells=6, src=(0.4, 0.6, 0.3),
                    gauss_width=0.25, time_domain=30)
    # 200 points across x with layers of 5: tiles outside every layer (the interior launch of
    # launch shape 6) and the shell around them, over terrain with the log wind
    t = B(200, 32, 22, 0.05, pml=5, bc=(2, 3, 2, 2, 2, 2), porous_cells=8, x0=-5.0, src=(0.0, 0.6, 0.3),
          gauss_width=0.25, time_domain=30, velocity_method=7)
    t.hills, t.wind_ustar, t.wind_z0 = [(0.1, 0.7, 0.2, 0.3), (-2.0, 1.1, 0.12, 0.25)], 0.4, 0.02
    out["split"] = t
    # 128 points across x: whole 64-point tiles, so the staged tiles' x halos reach past both
    # x faces (only grid points are staged)
    t = B(128, 30, 20, 0.05, pml=5, bc=(2, 3, 2, 2, 2, 2), porous_cells=6, x0=-3.2, src=(0.0, 0.6, 0.3),
          gauss_width=0.25, time_domain=24, velocity_method=7)
    t.hills, t.wind_ustar, t.wind_z0 = [(0.2, 0.6, 0.15, 0.3)], 0.4, 0.02
    out["split128"] = t
    for d in out.values():
        d.integrator = 1
    return out


@pytest.mark.parametrize("name", ["pml", "porous", "wind5", "terrain", "overlap", "wide", "split", "split128"])
def test_rk4_fp64_parity(lee, orc, name):
    d = cases(lee)[name]
    print(f"[gpu rk4] case={name} shape={d.shape}", flush=True)
    # the split case runs launch shape 6 (interior box + layer shell), the others the tuner's
    s = lee.Solver(d, precision=64, receivers=REC, rk_tile=6 if name == "split" else 0)
    s.frame_setup(0)
    got = s.step(d.time_domain)
    o = orc.Oracle(d, receivers=REC)
    o.frame_setup(0)
    ref = o.step(d.time_domain)
    assert np.abs(ref).max() > 0
    e_rec = rel(got, ref)
    fs, fo = s.state(), o.state()
    e_f = max(rel(a, b) for a, b in zip(fs, fo))
    same = all(np.array_equal(a, b) for a, b in zip(fs, fo))
    if d.has_pore:
        e_f = max(e_f, max(rel(a, b) for a, b in zip(s.state(1), o.state(1))))
    print(f"  rec rel-L2 {e_rec:.3e}  field rel-L2 {e_f:.3e}  bitwise {same}", flush=True)
    s.close()
    assert e_rec <= 1e-10
    assert e_f <= 1e-10


def test_rk4_terrain_moving_frame(lee, orc):
    d = lee.deck.box(22, 30, 18, 0.05, pml=4, bc=(2, 3, 2, 2, 2, 2), porous_cells=6, src=(0.0, 0.5, 0.4),
                     gauss_width=0.25, time_domain=80, mf_judge=1, mf_jmax=30, mf_lead_dj=6, mf_trail_dj=6,
                     mf_ratio=0.7, ijkmax=70, c=343.0, velocity_method=7)
    d.air_y = np.arange(70) * 0.05
    d.pore_y = d.air_y.copy()
    d.integrator, d.hills, d.wind_ustar, d.wind_z0 = 1, [(0.0, 1.2, 0.15, 0.3)], 0.3, 0.02
    print("[gpu rk4] terrain + moving frame", flush=True)
    s = lee.Solver(d, precision=64, receivers=REC)
    assert s.frame_count() > 2
    got = s.run()
    o = orc.Oracle(d, receivers=REC)
    ref = o.run()
    e = rel(got, ref)
    fe = max(rel(a, b) for a, b in zip(s.state(), o.state()))
    print(f"  frames {s.frame_count()} rec rel-L2 {e:.3e} field {fe:.3e}", flush=True)
    s.close()
    assert e <= 1e-10 and fe <= 1e-10


def test_rk4_fp32_compensated_tracks_fp64(lee):
    d = cases(lee)["terrain"]
    d.time_domain = 200
    print("[gpu rk4] fp32 vs fp64, 200 steps", flush=True)
    a = lee.Solver(d, precision=64, receivers=REC)
    b = lee.Solver(d, precision=32, receivers=REC)
    a.frame_setup(0); b.frame_setup(0)
    ra, rb = a.step(d.time_domain), b.step(d.time_domain)
    e = rel(rb, ra)
    print(f"  fp32 receiver rel-L2 vs fp64 {e:.3e}", flush=True)
    a.close(); b.close()
    assert e < 1e-5


def test_fb_terrain_split_matches_oracle(lee, orc):
    """The FB_p_v step over terrain (split passes) against the oracle."""
    d = cases(lee)["terrain"]
    d.integrator = 0
    print("[gpu] FB over terrain", flush=True)
    s = lee.Solver(d, precision=64, receivers=REC)
    s.frame_setup(0)
    got = s.step(d.time_domain)
    o = orc.Oracle(d, receivers=REC)
    o.frame_setup(0)
    ref = o.step(d.time_domain)
    e = rel(got, ref)
    fe = max(rel(x, y) for x, y in zip(s.state(), o.state()))
    print(f"  rec rel-L2 {e:.3e} field {fe:.3e}", flush=True)
    s.close()
    assert e <= 1e-10 and fe <= 1e-10


def test_fb_two_x_tiles_match_oracle(lee, orc):
    """The FB_p_v step on a grid two x tiles wide (64 + 6 points; the tiles start at i = 0)."""
    d = cases(lee)["wide"]
    d.integrator = 0
    print("[gpu] FB two x tiles", flush=True)
    s = lee.Solver(d, precision=64, receivers=REC)
    s.frame_setup(0)
    got = s.step(d.time_domain)
    o = orc.Oracle(d, receivers=REC)
    o.frame_setup(0)
    ref = o.step(d.time_domain)
    assert np.abs(ref).max() > 0
    e = rel(got, ref)
    fe = max(rel(x, y) for x, y in zip(s.state(), o.state()))
    print(f"  rec rel-L2 {e:.3e} field {fe:.3e}", flush=True)
    s.close()
    assert e <= 1e-10 and fe <= 1e-10


def _launch_shapes_agree(lee, d, prec, name):
    print(f"[gpu] rk4 launch shapes fp{prec} case={name}", flush=True)
    ref = None
    for v in range(1, 12):
        s = lee.Solver(d, precision=prec, receivers=REC, integrator=1, rk_tile=v)
        s.frame_setup(0)
        got = (s.step(d.time_domain), s.state(), s.state(1))
        assert s.info().rk_tile == v
        s.close()
        if ref is None:
            ref = got
            continue
        assert np.array_equal(got[0], ref[0]), (name, v)
        for a, b in zip(got[1] + got[2], ref[1] + ref[2]):
            assert np.array_equal(a, b), (name, v)


@pytest.mark.parametrize("prec", [64, 32])
def test_rk4_launch_shapes_give_the_same_bits(lee, prec):
    """Every launch shape of the RK4 air stage (the autotuner's candidates) computes the same
    per-point arithmetic: identical receivers and fields, on a grid whose last x tile is
    partial and on one of whole tiles."""
    for name in ("split", "split128"):
        _launch_shapes_agree(lee, cases(lee)[name], prec, name)
    d = cases(lee)["split"]
    s = lee.Solver(d, precision=prec, receivers=REC, integrator=1)
    s.frame_setup(0)
    s.step(2)
    assert 1 <= s.info().rk_tile <= 11
    s.close()
