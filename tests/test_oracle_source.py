"""Time-series source of calculation_2D::UpdateSource (calculation_2D.cpp:371-385) in the
oracle: when it fires, with which amplitude, and what it adds.  Pinned by the definition
itself on a linear system: the run with the series minus the run without it is zero before
step `period` and equals ps[1] times the initial Gaussian right after it."""
import numpy as np
import pytest

from oracle import oracle as O


def _pair(lee, period, ps, steps, integrator=0):
    d = lee.deck.box(16, 18, 14, 0.05, bc=(1, 1, 1, 1, 1, 1), src=(0.0, 0.4, 0.3), gauss_width=0.15)
    out = []
    for series in (None, ps):
        o = O.Oracle(d, integrator=integrator)
        o.frame_setup(0)
        if series is not None:
            o.set_source_series(series, period)
        o.step(steps)
        out.append(o.state()[3])
    return d, out


@pytest.mark.parametrize("integrator", [0, 1])
def test_source_fires_after_step_period_with_ps_of_n_over_period(lee, integrator):
    period, amp = 5, 0.37
    # ps[0] would fire at N = 0, which never happens: give it a huge value to prove it
    ps = [1e6, amp, 0.0]
    d, (p0, p1) = _pair(lee, period, ps, period - 1, integrator)
    assert np.array_equal(p0, p1)
    d, (p0, p1) = _pair(lee, period, ps, period, integrator)
    x, y, z = np.meshgrid(d.air_x, d.air_y, d.air_z, indexing="ij")
    r2 = (x - d.src[0]) ** 2 + (y - d.src[1]) ** 2 + (z - d.src[2]) ** 2
    G = np.exp(-4 * np.log(2) / d.gauss_width ** 2 * r2).transpose(2, 1, 0)
    assert np.allclose(p1 - p0, amp * G, rtol=1e-12, atol=1e-14)
    assert np.abs(p1 - p0).max() > 0.3


def test_source_series_ends_and_can_be_removed(lee):
    d, (p0, p1) = _pair(lee, 3, [0.0, 0.0], 12)   # only ps[1] = 0 inside the series
    assert np.array_equal(p0, p1)
    o = O.Oracle(d)
    o.frame_setup(0)
    o.set_source_series([0.0, 1.0], 2)
    o.set_source_series([], 2)
    o.step(4)
    o2 = O.Oracle(d)
    o2.frame_setup(0)
    o2.step(4)
    assert np.array_equal(o.state()[3], o2.state()[3])
