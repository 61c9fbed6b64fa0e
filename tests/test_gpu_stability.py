"""Long-run stability of the RK4 path on the features the 512x256x128 benchmark workload
combines: PML on all faces under a moving frame, a strong logarithmic wind, terrain, porous ground.  The
reference's FB_p_v differencing grows in two of these settings (downwind convection of the
cross-flow velocities; y-layer auxiliaries left in place by the frame shift); the RK4 path
must stay bounded over thousands of steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def deck(lee, mf=True, pml=12, wind=True, hill=True, porous=True, steps=3000):
    d = 0.04
    kw = dict(pml=pml, bc=(2, 3 if porous else 1, 2, 2, 2, 2), porous_cells=10 if porous else 0, time_domain=steps,
              gauss_width=6 * d, src=(0.0, 1.0, 0.5), pml_y_time=0)
    if mf:
        kw.update(mf_judge=1, mf_jmax=64, mf_lead_dj=9, mf_trail_dj=9, mf_ratio=0.5, ijkmax=64 + 8 * 80)
    dk = lee.deck.box(48, 64, 40, d, **kw)
    if mf:
        dk.air_y = np.arange(64 + 8 * 80) * d
    dk.integrator = 1
    if porous:
        dk.pore_q, dk.pore_porosity, dk.pore_resistivity = 3.0 ** 0.5, 0.3, 2.0e5
    if hill:
        dk.hills = [(0.0, 2.0, 0.4, 0.6)]
    if wind:
        dk.velocity_method, dk.wind_ustar, dk.wind_z0 = 7, 0.6, 0.02
    return dk


def run(lee, dk, precision):
    s = lee.Solver(dk, precision=precision, receivers=[(0.0, 1.5, 0.6)], health_every=250)
    mf, temp = 0, 0
    left = s.frame_steps(0, 0)
    s.frame_setup(0)
    done, peak = 0, []
    while done < dk.time_domain:
        if left <= 0:
            mf += 1
            left = s.frame_steps(mf, temp)
            s.frame_setup(mf)
        k = min(left, dk.time_domain - done)
        s.step(k, record=False)
        left -= k; temp += k; done += k
    p = s.state()[3]
    s.close()
    return float(np.abs(p).max()), mf


@pytest.mark.parametrize("precision", [64, 32])
def test_rk4_moving_frame_pml_wind_terrain_stays_bounded(lee, precision):
    dk = deck(lee)
    print(f"[gpu stability] RK4 fp{precision} mf+pml+wind+terrain+porous, {dk.time_domain} steps", flush=True)
    peak, frames = run(lee, dk, precision)
    print(f"  frames {frames} final max|p| {peak:.3e}", flush=True)
    assert frames > 10
    assert np.isfinite(peak) and peak < 1.0
