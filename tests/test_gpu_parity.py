"""FP64 GPU path vs the CPU oracle, through the C ABI (FB_p_v, reference semantics).

The bar is <= 1e-10 relative L2 on the receiver series and on the final fields; the FP64
device path keeps the reference's operation order without FMA contraction, so the
differences are expected to be zero or at rounding level.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REC = [(0.0, 0.55, 0.45), (0.15, 0.75, 0.3), (-0.2, 0.35, 0.6), (0.05, 1.05, 0.2)]


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def cases(lee):
    B = lee.deck.box
    out = {}
    out["rigid"] = B(24, 28, 20, 0.05, bc=(1, 1, 1, 1, 1, 1), src=(0.0, 0.6, 0.5), gauss_width=0.25, time_domain=40)
    out["pml5"] = B(26, 30, 22, 0.05, pml=5, bc=(2, 1, 2, 2, 2, 2), src=(0.1, 0.7, 0.5), gauss_width=0.25,
                    time_domain=40, pml_y_time=3)
    out["tdbc"] = B(24, 30, 20, 0.05, pml=5, bc=(2, 1, 5, 2, 2, 2), src=(0.0, 1.0, 0.5), gauss_width=0.25,
                    time_domain=40, pml_y_time=10)
    out["porous"] = B(24, 26, 20, 0.05, pml=5, bc=(2, 3, 2, 2, 2, 2), porous_cells=8, src=(0.0, 0.6, 0.3),
                      gauss_width=0.25, time_domain=40)
    out["wind5"] = B(24, 28, 20, 0.05, pml=5, bc=(2, 1, 2, 2, 2, 2), src=(0.0, 0.6, 0.5), gauss_width=0.25,
                     time_domain=40, velocity_method=5, velocity_coef=3.0)
    out["wind2"] = B(24, 28, 20, 0.05, bc=(1, 1, 1, 1, 1, 1), src=(0.0, 0.6, 0.5), gauss_width=0.25,
                     time_domain=40, velocity_method=2, velocity_coef=1.0)
    # long in y: the fused pass splits the planes into chunks of several planes
    out["long"] = B(24, 200, 20, 0.05, pml=5, bc=(2, 3, 5, 2, 2, 2), porous_cells=6, src=(0.0, 1.0, 0.4),
                    gauss_width=0.25, time_domain=30, pml_y_time=5)
    # 264 points across x: two x tiles of the 16-byte kernels in FP32 (256 points per
    # wavefront) and three in FP64 (128), the last one partial; x = 0 at i = 256, so the
    # pulse crosses the tile boundary into the x2 layer (i >= 259)
    out["wide"] = B(264, 30, 12, 0.05, pml=5, bc=(2, 3, 2, 2, 2, 2), porous_cells=5, x0=-12.8,
                    src=(0.0, 0.6, 0.3), gauss_width=0.25, time_domain=30)
    # 200 x 30 x 16 with layers of 5 points: two 64 x 4-point tile columns and three tile rows
    # lie outside every layer, so the split passes run their layer-free interior launch next
    # to the shell launch; over terrain, and with a mean flow
    t = B(200, 30, 16, 0.05, pml=5, bc=(2, 3, 2, 2, 2, 2), porous_cells=6, x0=-5.0, src=(0.0, 0.6, 0.3),
          gauss_width=0.25, time_domain=30)
    t.hills = [(0.3, 0.7, 0.08, 0.3), (-1.0, 0.9, 0.05, 0.4)]
    out["wide_terrain"] = t
    out["wide_wind"] = B(200, 30, 16, 0.05, pml=5, bc=(2, 1, 2, 2, 2, 2), x0=-5.0, src=(0.0, 0.6, 0.4),
                         gauss_width=0.25, time_domain=30, velocity_method=5, velocity_coef=3.0)
    return out


PATHS = {"split": 1, "fused": 2}


# the fused pass runs separable grids only (a terrain deck is refused on it)
FB_CASES = [(n, p) for n in ["rigid", "pml5", "tdbc", "porous", "wind5", "wind2", "long", "wide", "wide_terrain",
                             "wide_wind"] for p in ["split", "fused"] if not (n == "wide_terrain" and p == "fused")]


@pytest.mark.parametrize("name,path", FB_CASES)
def test_fb_fp64_parity(lee, orc, name, path):
    d = cases(lee)[name]
    print(f"[gpu parity] case={name} path={path} shape={d.shape}", flush=True)
    s = lee.Solver(d, precision=64, receivers=REC, path=PATHS[path])
    assert s.path == PATHS[path]
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
        ps, po = s.state(1), o.state(1)
        e_f = max(e_f, max(rel(a, b) for a, b in zip(ps, po)))
    if name == "tdbc":
        e_f = max(e_f, rel(s.psi(), o.psi()))
    print(f"  rec rel-L2 {e_rec:.3e}  field rel-L2 {e_f:.3e}  bitwise {same}", flush=True)
    s.close()
    assert e_rec <= 1e-10
    assert e_f <= 1e-10


@pytest.mark.parametrize("path", ["split", "fused"])
def test_moving_frame_run(lee, orc, path):
    d = lee.deck.box(22, 30, 18, 0.05, pml=4, bc=(2, 1, 2, 2, 2, 2), src=(0.0, 0.5, 0.4), gauss_width=0.25,
                     time_domain=90, mf_judge=1, mf_jmax=30, mf_lead_dj=6, mf_trail_dj=6, mf_ratio=0.7,
                     ijkmax=70, c=343.0)
    d.air_y = np.arange(70) * 0.05
    print(f"[gpu parity] moving frame path={path}", flush=True)
    s = lee.Solver(d, precision=64, receivers=REC, path=PATHS[path])
    assert s.frame_count() > 2
    got = s.run()
    o = orc.Oracle(d, receivers=REC)
    ref = o.run()
    e = rel(got, ref)
    fe = max(rel(a, b) for a, b in zip(s.state(), o.state()))
    print(f"  frames {s.frame_count()} rec rel-L2 {e:.3e} field {fe:.3e}", flush=True)
    s.close()
    assert e <= 1e-10 and fe <= 1e-10


def test_fp32_tracks_fp64(lee):
    d = cases(lee)["pml5"]
    print("[gpu] fp32 vs fp64", flush=True)
    a = lee.Solver(d, precision=64, receivers=REC)
    b = lee.Solver(d, precision=32, receivers=REC)
    a.frame_setup(0); b.frame_setup(0)
    ra, rb = a.step(d.time_domain), b.step(d.time_domain)
    e = rel(rb, ra)
    print(f"  fp32 receiver rel-L2 vs fp64 {e:.3e}", flush=True)
    a.close(); b.close()
    assert e < 1e-4


@pytest.mark.parametrize("prec", [32, 64])
def test_fused_equals_split_bitwise(lee, prec):
    """Both kernel paths give the same bits, FP32 included."""
    for name in ("pml5", "porous", "wind5", "long", "wide"):
        d = cases(lee)[name]
        print(f"[gpu] fused vs split fp{prec} case={name}", flush=True)
        a = lee.Solver(d, precision=prec, receivers=REC, path=PATHS["split"])
        b = lee.Solver(d, precision=prec, receivers=REC, path=PATHS["fused"])
        a.frame_setup(0); b.frame_setup(0)
        ra, rb = a.step(d.time_domain), b.step(d.time_domain)
        assert np.array_equal(ra, rb), name
        for x, y in zip(a.state(), b.state()):
            assert np.array_equal(x, y), name
        if d.has_pore:
            for x, y in zip(a.state(1), b.state(1)):
                assert np.array_equal(x, y), name
        a.close(); b.close()


@pytest.mark.parametrize("prec", [32, 64])
def test_fb_launch_forms_give_the_same_bits(lee, prec):
    """Every launch form of the FB_p_v air passes the autotuner picks from (split, split with
    padded tile lists, fused at four pencil depths, split with one point per lane
    throughout) gives the same bits; the tuner keeps one of the forms the path allows.  On
    the 264-point-wide deck the FP32 split passes run their interior box four points per
    lane, so forms 1 and 2 check that form against the one-point forms."""
    for name in ("pml5", "wide"):
        d = cases(lee)[name]
        ref = None
        for form in range(1, 8):
            print(f"[gpu] fb launch form {form} fp{prec} case={name}", flush=True)
            s = lee.Solver(d, precision=prec, receivers=REC, fb_form=form)
            s.frame_setup(0)
            got = (s.step(d.time_domain), s.state(), s.state(1) if d.has_pore else [])
            assert s.info().fb_form == form
            s.close()
            if ref is None:
                ref = got
                continue
            assert np.array_equal(got[0], ref[0]), (name, form)
            for x, y in zip(got[1] + got[2], ref[1] + ref[2]):
                assert np.array_equal(x, y), (name, form)
    d = cases(lee)["wide"]
    for path, allowed in (("split", (1, 2, 7)), ("fused", (3, 4, 5, 6)), (None, tuple(range(1, 8)))):
        s = lee.Solver(d, precision=prec, receivers=REC, path=PATHS[path] if path else 0)
        s.frame_setup(0)
        s.step(2)
        assert s.info().fb_form in allowed, path
        s.close()
    with pytest.raises(lee.LeeError):
        lee.Solver(d, precision=prec, receivers=REC, path=PATHS["split"], fb_form=4)
    with pytest.raises(lee.LeeError):
        lee.Solver(d, precision=prec, receivers=REC, path=PATHS["fused"], fb_form=7)
    with pytest.raises(lee.LeeError):
        lee.Solver(d, precision=prec, receivers=REC, fb_form=8)


@pytest.mark.parametrize("name", ["wide_terrain", "wide_wind"])
def test_fb_vector_interior_gives_the_same_bits(lee, name):
    """FP32 split passes: the interior box four points per lane (form 1; over terrain the
    velocity pass only) against one point per lane (form 7), fields and receivers bit for
    bit; with a mean flow both run the one-point form."""
    d = cases(lee)[name]
    got = []
    for form in (1, 7):
        print(f"[gpu] fb vector interior form {form} case={name}", flush=True)
        s = lee.Solver(d, precision=32, receivers=REC, path=PATHS["split"], fb_form=form)
        s.frame_setup(0)
        got.append((s.step(d.time_domain), s.state(), s.state(1) if d.has_pore else []))
        s.close()
    assert np.array_equal(got[0][0], got[1][0]), name
    for x, y in zip(got[0][1] + got[0][2], got[1][1] + got[1][2]):
        assert np.array_equal(x, y), name
    assert np.abs(got[0][1][3]).max() > 0, name
