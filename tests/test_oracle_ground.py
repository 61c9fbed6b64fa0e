"""Ground impedance (South-face TDBC with K poles closing a ground interface plane) in the
oracle.  Pinned by (1) an exact equivalence: with an air-like top material (q = 1, porosity 1,
no flow resistance) and an infinite impedance, the interface plane is an air plane over a
rigid wall half a cell below it -- the rigid floor of the reference one plane lower; (2) the
closure's defining relations between consecutive steps (porous momentum over half a cell
with the impedance's pressure, the recursion's history sum); (3) plane waves at normal
incidence: a matched impedance (Z = rho0 c) absorbs, a resistive one reflects with
R = (Z - rho0 c)/(Z + rho0 c)."""
import numpy as np
import pytest

from oracle import oracle as O


AIR_TOP = dict(pore_q=1.0, pore_porosity=1.0, pore_resistivity=0.0)


def _column(lee, nz=160, **ground):
    d = lee.deck.box(4, 4, nz, 0.02, bc=(1, 1, 1, 1, 1, 1), src=(0.0, 0.0, -50.0))
    if "ground_poles" in ground:
        ground = {**AIR_TOP, **ground}
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


def test_air_top_and_infinite_impedance_is_the_rigid_floor_one_plane_lower(lee):
    """The same physical nodes on two grids: (a) the ground closure at z = 0 with an air-like
    top material and Z = 1e30 (the normal velocity half a cell below the plane stays 0), and
    (b) the reference's rigid South face on the grid extended by one plane below (its wall
    sits half a cell under its plane k = 1, which is z = 0).  Receiver series agree to
    rounding (the plane's metric is 1/X_over_x where the interior forms the Jacobian)."""
    d = 0.05
    kw = dict(bc=(1, 1, 1, 1, 1, 1), src=(0.05, 0.6, 0.3), gauss_width=0.2, time_domain=120)
    a = lee.deck.box(16, 18, 14, d, ground_poles=[(0.0, 1.0)], ground_zinf=1e30, ground_tau=1e-3, **AIR_TOP, **kw)
    b = lee.deck.box(16, 18, 15, d, z0=-d, **kw)
    rec = [(0.05, 0.3, 0.05), (-0.1, 0.55, 0.1), (0.2, 0.7, 0.35)]
    oa, ob = O.Oracle(a, receivers=rec), O.Oracle(b, receivers=rec)
    oa.frame_setup(0); ob.frame_setup(0)
    ra, rb = oa.step(120), ob.step(120)
    assert np.abs(rb).max() > 1e-3
    assert np.linalg.norm(ra - rb) <= 1e-10 * np.linalg.norm(rb)
    pa, pb = oa.state()[3], ob.state()[3][1:]
    assert np.linalg.norm(pa - pb) <= 1e-10 * np.linalg.norm(pb)


def _ground_deck(lee):
    rng = np.random.default_rng(3)
    poles = [(float(a), float(r)) for a, r in zip(rng.uniform(0.05, 0.8, 8), rng.uniform(0.0, 2.0, 8))]
    return lee.deck.box(12, 14, 16, 0.05, bc=(1, 1, 1, 1, 1, 1), src=(0.0, 0.3, 0.2), gauss_width=0.2,
                        ground_poles=poles, ground_zinf=700.0, ground_tau=2e-4, pore_q=1.6, pore_porosity=0.4,
                        pore_resistivity=8.0e4), poles


def test_ground_closure_relations(lee):
    """Between consecutive steps, on the interior nodes of the plane k = 0:
      w' = ib (w - cpl (p' - p_h)),  p_h = Zinf (-w' + sum_k Psi'_k)   (half-cell momentum,
      the impedance's pressure through the recursion p_h / Zinf = v + sum_k Psi'_k),
      Psi'_k = e^{-r_k tb} Psi_k + a_k tb v,  v = -w'
      p' = p - gama (dx u + dy v + (w(k=1) - w) / dz)   (porous continuity of the top material)."""
    d, poles = _ground_deck(lee)
    o = O.Oracle(d)
    o.frame_setup(0)
    o.step(9)
    u0, v0, w0, p0 = o.state()
    psi0 = o.ground_psi(8)
    o.step(1)
    u1, v1, w1, p1 = o.state()
    psi1 = o.ground_psi(8)
    rho0 = d.gamma * d.Pav / d.c ** 2
    eff = d.pore_porosity / (d.pore_q ** 2 * rho0) * 1.0
    ib = 1.0 / (d.pore_resistivity * d.dt * eff + 1.0)
    cpl = 2.0 * d.dt * eff / d.dz
    gama = d.dt * d.gamma * d.Pav / d.pore_porosity
    sl = (slice(1, -1), slice(1, -1))
    vg = -w1[0][sl]
    ph = 700.0 * (vg + psi1[sl].sum(axis=2))
    lhs = w1[0][sl]
    rhs = ib * (w0[0][sl] - cpl * (p1[0][sl] - ph))
    assert np.abs(lhs - rhs).max() <= 1e-9 * np.abs(lhs).max() and np.abs(lhs).max() > 0
    tb = d.dt / d.ground_tau
    for q, (a, r) in enumerate(poles):
        ref = np.exp(-r * tb) * psi0[sl][..., q] + a * tb * vg
        assert np.allclose(psi1[sl][..., q], ref, rtol=1e-12, atol=1e-22)
    div = (u0[0][1:-1, 2:] - u0[0][1:-1, 1:-1]) / d.dx + (v0[0][2:, 1:-1] - v0[0][1:-1, 1:-1]) / d.dy + \
          (w0[1][sl] - w0[0][sl]) / d.dz
    assert np.allclose(p1[0][sl], p0[0][sl] - gama * div, rtol=1e-9, atol=1e-9 * np.abs(p1[0]).max())


def test_ground_impedance_is_rejected_where_undefined(lee):
    d = lee.deck.box(12, 14, 16, 0.05, bc=(1, 3, 1, 1, 1, 1), porous_cells=6, ground_poles=[(0.1, 1.0)],
                     ground_zinf=700.0, ground_tau=2e-4)
    with pytest.raises(ValueError):
        O.Oracle(d)
    d = lee.deck.box(12, 14, 16, 0.05, bc=(1, 1, 1, 1, 1, 1), ground_poles=[(-0.1, 1.0)], ground_zinf=700.0,
                     ground_tau=2e-4, **AIR_TOP)
    with pytest.raises(ValueError):
        O.Oracle(d)
    d = lee.deck.box(12, 14, 16, 0.05, bc=(1, 1, 1, 1, 1, 1), ground_poles=[(0.1, 1.0)], ground_zinf=700.0,
                     ground_tau=2e-4, pore_porosity=0.0)   # no top material
    with pytest.raises(ValueError):
        O.Oracle(d)
    d = lee.deck.box(12, 14, 16, 0.05, bc=(1, 1, 1, 1, 1, 1), ground_poles=[(0.1, 1.0)], ground_zinf=700.0,
                     ground_tau=2e-4, **AIR_TOP)
    d.integrator = 1   # discretised for the FB_p_v step only
    with pytest.raises(ValueError):
        O.Oracle(d)


def _terrain_ground_deck(lee, hill_height):
    rng = np.random.default_rng(11)
    poles = [(float(a), float(r)) for a, r in zip(rng.uniform(0.05, 0.8, 8), rng.uniform(0.0, 2.0, 8))]
    d = lee.deck.box(26, 30, 22, 0.05, pml=5, bc=(2, 1, 2, 2, 2, 2), src=(0.0, 0.7, 0.3), gauss_width=0.25,
                     time_domain=120, ground_poles=poles, ground_zinf=900.0, ground_tau=1.5e-4)
    d.hills = [(0.1, 0.7, hill_height, 0.3)]
    return d


REC_T = [(0.0, 0.55, 0.3), (0.15, 0.75, 0.1), (-0.2, 0.95, 0.05)]


def test_terrain_ground_over_a_flat_hill_is_the_flat_closure(lee):
    """The ground plane on the terrain-following grid (the plane's metric with the cross terms
    zX, zY and 1/Z_over_z per node) over a hill of zero height is the flat closure, bit for
    bit: fields, receivers and histories."""
    flat = _terrain_ground_deck(lee, 0.0)
    flat.hills = []
    ter = _terrain_ground_deck(lee, 0.0)
    out = []
    for d in (flat, ter):
        o = O.Oracle(d, receivers=REC_T)
        o.frame_setup(0)
        out.append((o.step(60), o.state(), o.ground_psi(8)))
    assert np.abs(out[0][0]).max() > 0
    assert np.array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)
    assert np.array_equal(out[0][2], out[1][2])


def test_terrain_ground_stays_bounded_and_feels_the_hill(lee):
    """Over a 0.15 m hill the ground plane's metric differs from the flat one (the receiver
    series change by more than rounding) and the run stays bounded (no growth over 400 steps
    with PML on the sides and top)."""
    flat = _terrain_ground_deck(lee, 0.0)
    flat.hills = []
    hill = _terrain_ground_deck(lee, 0.15)
    series = []
    for d in (flat, hill):
        o = O.Oracle(d, receivers=REC_T)
        o.frame_setup(0)
        r = o.step(400)
        p = o.state()[3]
        assert np.isfinite(r).all() and np.isfinite(p).all()
        series.append(r)
        # energy late in the run well below the early peak: the layers absorb, nothing grows
        assert np.abs(r[300:]).max() < 0.5 * np.abs(r[:150]).max()
    diff = np.linalg.norm(series[1] - series[0]) / np.linalg.norm(series[0])
    assert diff > 1e-3
