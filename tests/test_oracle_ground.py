"""Ground impedance (South-face TDBC with K poles) in the oracle.  Pinned by the recursion's
invariant and by the physics of a plane wave at normal incidence: a matched impedance
(Z = rho0 c, no poles' weight) absorbs it, a rigid face returns it, and an impedance with a
known pole set reflects with the coefficient R = (Z - rho0 c)/(Z + rho0 c) at low frequency."""
import numpy as np
import pytest

from oracle import oracle as O


def _column(lee, nz=160, **ground):
    d = lee.deck.box(4, 4, nz, 0.02, bc=(1, 1, 1, 1, 1, 1), src=(0.0, 0.0, -50.0))
    for k, v in ground.items():
        setattr(d, k, v)
    return d


def _plane_wave(o, d, z0, width):
    """A pulse travelling down the column: p = exp(-((z - z0)/width)^2), w = -p/(rho0 c)."""
    rho = d.gamma * d.Pav / d.c ** 2
    z = np.asarray(d.air_z)[:, None, None] * np.ones(o.shape)
    p = np.exp(-((z - z0) / width) ** 2)
    zero = np.zeros(o.shape)
    o.set_state(zero, zero, -p / (rho * d.c), p)


def _reflection(lee, **ground):
    d = _column(lee, **ground)
    o = O.Oracle(d, receivers=[(-0.01, 0.02, 1.6)])
    o.frame_setup(0)
    _plane_wave(o, d, 2.4, 0.12)
    rec = o.step(900)[:, 0]
    inc = np.abs(rec[:300]).max()      # the incident pulse passes z = 1.6 m first
    ref = np.abs(rec[450:]).max()      # the reflection returns after ~ 2 * 1.6 m / c
    return ref / inc, o, d


def test_matched_and_rigid_ground(lee):
    rho_c = 1.21 * 343.0
    r_rigid, _, _ = _reflection(lee)
    r_match, _, _ = _reflection(lee, ground_poles=[(0.0, 1.0)], ground_zinf=rho_c, ground_tau=1e-3)
    assert r_rigid > 0.9
    assert r_match < 0.05


def test_resistive_ground_reflection_coefficient(lee):
    """Z = 3 rho0 c (one pole of zero weight): R = (3 - 1)/(3 + 1) = 0.5."""
    rho_c = 1.21 * 343.0
    r, _, _ = _reflection(lee, ground_poles=[(0.0, 1.0)], ground_zinf=3 * rho_c, ground_tau=1e-3)
    assert abs(r - 0.5) < 0.03


def test_ground_recursion_invariant(lee):
    """Cal_TDBC / Update_Psi with K poles keep -w + sum_k Psi_k = p / Z_inf at every face node
    after every step (the relation the recursion is built on)."""
    rng = np.random.default_rng(3)
    poles = [(float(a), float(r)) for a, r in zip(rng.uniform(0.05, 0.8, 8), rng.uniform(0.0, 2.0, 8))]
    d = lee.deck.box(12, 14, 16, 0.05, bc=(1, 1, 1, 1, 1, 1), src=(0.0, 0.3, 0.2), gauss_width=0.2,
                     ground_poles=poles, ground_zinf=700.0, ground_tau=2e-4)
    o = O.Oracle(d)
    o.frame_setup(0)
    for _ in range(3):
        o.step(7)
        u, v, w, p = o.state()
        psi = o.ground_psi(8)
        lhs = -w[0] + psi.sum(axis=2)
        assert np.allclose(lhs, p[0] / 700.0, rtol=1e-12, atol=1e-18)
        assert np.abs(psi).max() > 0


def test_ground_impedance_is_rejected_where_undefined(lee):
    d = lee.deck.box(12, 14, 16, 0.05, bc=(1, 3, 1, 1, 1, 1), porous_cells=6, ground_poles=[(0.1, 1.0)],
                     ground_zinf=700.0, ground_tau=2e-4)
    with pytest.raises(ValueError):
        O.Oracle(d)
    d = lee.deck.box(12, 14, 16, 0.05, bc=(1, 1, 1, 1, 1, 1), ground_poles=[(-0.1, 1.0)], ground_zinf=700.0,
                     ground_tau=2e-4)
    with pytest.raises(ValueError):
        O.Oracle(d)
