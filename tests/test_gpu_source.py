"""Time-series source (calculation_2D::UpdateSource) on the GPU against the oracle through the
C ABI: FB_p_v and RK4, and across moving-frame shifts (the Gaussian follows the frame's
coordinates).  The device evaluates exp() with the GPU math library, so FP64 is held to
1e-10 relative L2 rather than bitwise."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REC = [(0.0, 0.55, 0.45), (0.15, 0.75, 0.3)]


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize("integrator", [0, 1])
def test_source_series_parity(lee, orc, integrator):
    d = lee.deck.box(24, 28, 20, 0.05, pml=5, bc=(2, 1, 2, 2, 2, 2), src=(0.0, 0.6, 0.5), gauss_width=0.25,
                     time_domain=50)
    ps = 0.5 * np.sin(np.arange(10) * 0.9)
    print(f"[gpu source] integrator={integrator}", flush=True)
    s = lee.Solver(d, precision=64, receivers=REC, integrator=integrator)
    s.frame_setup(0)
    s.set_source_series(ps, 7)
    got = s.step(d.time_domain)
    o = orc.Oracle(d, receivers=REC, integrator=integrator)
    o.frame_setup(0)
    o.set_source_series(ps, 7)
    ref = o.step(d.time_domain)
    e = rel(got, ref)
    fe = max(rel(a, b) for a, b in zip(s.state(), o.state()))
    # the source changed the answer
    o0 = orc.Oracle(d, receivers=REC, integrator=integrator)
    o0.frame_setup(0)
    assert rel(o0.step(d.time_domain), ref) > 1e-3
    print(f"  rec rel-L2 {e:.3e} field {fe:.3e}", flush=True)
    s.close()
    assert e <= 1e-10 and fe <= 1e-10


def test_source_series_moving_frame(lee, orc):
    d = lee.deck.box(22, 30, 18, 0.05, pml=4, bc=(2, 1, 2, 2, 2, 2), src=(0.0, 0.5, 0.4), gauss_width=0.25,
                     time_domain=90, mf_judge=1, mf_jmax=30, mf_lead_dj=6, mf_trail_dj=6, mf_ratio=0.7,
                     ijkmax=70, c=343.0)
    d.air_y = np.arange(70) * 0.05
    ps = np.linspace(0.2, -0.3, 20)
    print("[gpu source] moving frame", flush=True)
    s = lee.Solver(d, precision=64, receivers=REC)
    s.set_source_series(ps, 5)
    got = s.run()
    o = orc.Oracle(d, receivers=REC)
    o.set_source_series(ps, 5)
    ref = o.run()
    e = rel(got, ref)
    fe = max(rel(a, b) for a, b in zip(s.state(), o.state()))
    print(f"  frames {s.frame_count()} rec rel-L2 {e:.3e} field {fe:.3e}", flush=True)
    s.close()
    assert e <= 1e-10 and fe <= 1e-10
