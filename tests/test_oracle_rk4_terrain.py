"""Pins of the oracle's extensions beyond the reference -- the RK4 integrator, the
terrain-following grid and the logarithmic wind -- by closed forms and special cases.

The reference has none of these (it parses scheme 5 = RK4 but never integrates with it,
calculation_2D.cpp:551-612; its grid is separable, calculation_2D.cpp:102-104), so each
check below is mathematics: the exact RK4 amplification of a standing mode and of the
porous drag, the observed fourth-order time convergence, the inverse of the discrete
coordinate Jacobian, the flat-terrain special case and the wind formula.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _rk4_matrix(L, dt):
    A = dt * L
    I = np.eye(L.shape[0])
    return I + A + A @ A / 2 + A @ A @ A / 6 + A @ A @ A @ A / 24


def _mode_deck(lee, axis, n, d=0.05, stretch=1.0):
    dims = [5, 5, 5]
    dims[axis] = n
    dk = lee.deck.box(*dims, d, bc=(1, 1, 1, 1, 1, 1), gauss_width=0.2, src=(0.0, 0.0, -50.0))
    coords = [dk.air_x, dk.air_y, dk.air_z]
    coords[axis] = stretch * d * np.arange(n)
    dk.air_x, dk.air_y, dk.air_z = coords
    dk.integrator = 1
    return dk


@pytest.mark.parametrize("axis,stretch", [(0, 1.0), (2, 1.0), (0, 2.0), (1, 1.0)])
def test_rk4_standing_mode_closed_form(lee, axis, stretch):
    """The standing wave p_i = P cos(t(i-1/2)), u_i = U sin(t(i-1)) of a rigid box is an
    eigenvector pair of the semi-discrete operator: dP/dt = -a U, dU/dt = b P with
    a = gP 2 sin(t/2)/h, b = 2 sin(t/2)/(rho h).  One RK4 step multiplies (P, U) by the
    degree-4 Taylor polynomial of dt [[0, -a], [b, 0]]; every face condition is met at
    every stage, so the oracle must follow that map exactly."""
    n, m_, steps = 34, 3, 25
    dk = _mode_deck(lee, axis, n, stretch=stretch)
    o = O.Oracle(dk)
    o.frame_setup(0)
    th = m_ * np.pi / (n - 2)
    idx = np.arange(n)
    prof_p = np.cos(th * (idx - 0.5))
    prof_u = np.sin(th * (idx - 1.0))
    shp = o.shape
    bshape = [1, 1, 1]
    bshape[2 - axis] = n
    P0, U0 = 1.0, 0.3
    fields = [np.zeros(shp) for _ in range(4)]
    fields[3][...] = P0 * prof_p.reshape(bshape)
    fields[axis][...] = U0 * prof_u.reshape(bshape)
    o.set_state(*fields)
    o.step(steps)
    h = stretch * 0.05
    gP = dk.gamma * dk.Pav
    rho_inv = dk.c ** 2 / gP
    a = gP / h * 2 * np.sin(th / 2)
    b = rho_inv / h * 2 * np.sin(th / 2)
    G = _rk4_matrix(np.array([[0.0, -a], [b, 0.0]]), dk.dt)
    P, U = np.linalg.matrix_power(G, steps) @ np.array([P0, U0])
    st = o.state()
    sl = [slice(None)] * 3
    keep = n - 1 if axis == 1 else n
    sl[2 - axis] = slice(0, keep)
    sl = tuple(sl)
    exp_p = np.broadcast_to((P * prof_p).reshape(bshape), shp)
    exp_u = np.broadcast_to((U * prof_u).reshape(bshape), shp)
    assert np.abs(st[3][sl] - exp_p[sl]).max() < 1e-11
    slu = list(sl)
    slu[2 - axis] = slice(1, keep)
    slu = tuple(slu)
    assert np.abs(st[axis][slu] - exp_u[slu]).max() < 1e-11
    # the RK4 map differs from FB's: the check would catch an FB step in its place
    assert abs(P - 1.0 * np.cos(0)) > 1e-6


@pytest.mark.parametrize("sigma", [2.0e5, 4.0e6])
def test_rk4_porous_drag_is_exact(lee, sigma):
    """A uniform velocity in the porous layer decays by du/dt = -(sigma/rho_e) u.  The RK4
    step takes the drag through its integrating factor, so one step multiplies it by exactly
    e^-z, z = sigma dt / rho_e (beta of porous.cpp:48) -- also far beyond the explicit RK4
    limit z < 2.79 -- at points the faces cannot reach within the steps taken."""
    dk = lee.deck.box(40, 40, 12, 0.05, bc=(1, 3, 1, 1, 1, 1), porous_cells=12, src=(0, 0, -50.0))
    dk.integrator = 1
    dk.pore_resistivity = sigma
    o = O.Oracle(dk)
    o.frame_setup(0)
    zero = [np.zeros(o.shape) for _ in range(4)]
    o.set_state(*zero)
    pshape = (o.pnz, o.pny, o.pnx)
    U = 0.2
    o.set_state(np.full(pshape, U), np.zeros(pshape), np.zeros(pshape), np.zeros(pshape), domain=1)
    steps = 3
    o.step(steps)
    eff = 1.0 / (dk.pore_q ** 2 / dk.pore_porosity / (dk.c ** 2 / dk.gamma / dk.Pav))
    z = dk.pore_resistivity * dk.dt * eff
    g = np.exp(-z)
    u = o.state(1)[0]
    c = u[o.pnz // 2, o.pny // 2, o.pnx // 2]
    assert abs(c - U * g ** steps) < 1e-14
    assert abs(g - 1 / (1 + z)) > 1e-4  # not the FB semi-implicit factor
    if sigma > 1e6:
        assert z > 2.79


def test_rk4_converges_at_fourth_order_in_time(lee):
    """At a fixed grid, halving dt shrinks the time-discretisation error 16-fold: the ratio
    of successive differences of runs at dt, dt/2, dt/4 to the same end time is ~16.  The
    pulse stays clear of the walls (face conditions re-imposed at every stage would lower
    the observed order there)."""
    base = dict(pml=0, bc=(1, 1, 1, 1, 1, 1), src=(0.0, 0.775, 0.775), gauss_width=0.15)
    ends = []
    for m in (1, 2, 4):
        dk = lee.deck.box(32, 32, 32, 0.05, dt=0.5 * 0.05 / 343.0 / m, **base)
        dk.integrator = 1
        o = O.Oracle(dk)
        o.frame_setup(0)
        o.step(12 * m)
        ends.append(o.state()[3])
    d1 = np.linalg.norm(ends[0] - ends[1])
    d2 = np.linalg.norm(ends[1] - ends[2])
    assert 12.0 < d1 / d2 < 20.0


def test_rk4_with_porous_ground_converges_at_fourth_order(lee):
    """The integrating-factor drag keeps fourth order at a moderate drag (z ~ 1): the same
    16-fold error ratio in the air and in the porous layer."""
    ends, pends = [], []
    for m in (1, 2, 4):
        dk = lee.deck.box(24, 24, 24, 0.05, dt=0.5 * 0.05 / 343.0 / m, pml=0, bc=(1, 3, 1, 1, 1, 1), porous_cells=10,
                          src=(0.0, 0.6, 0.35), gauss_width=0.15)
        dk.integrator = 1
        o = O.Oracle(dk)
        o.frame_setup(0)
        o.step(12 * m)
        ends.append(o.state()[3])
        pends.append(o.state(1)[3])
    for e in (ends, pends):
        d1, d2 = np.linalg.norm(e[0] - e[1]), np.linalg.norm(e[1] - e[2])
        assert 12.0 < d1 / d2 < 20.0


@pytest.mark.parametrize("hills", [(), ((0.0, 1.0, 0.25, 0.3),)])
def test_rk4_mean_flow_convection_is_neutral(lee, hills):
    """A strong log wind (u* = 8 m/s) over 200 steps: the RK4 path's central convective
    differences keep the field bounded, flat and over a hill, while the reference's FB_p_v
    pair (forward differences for u, v, w: downwind for V > 0) grows without bound on the
    same deck -- the instability the central form removes."""
    peaks = {}
    for integ in (1, 0):
        d = lee.deck.box(16, 40, 16, 0.05, bc=(1, 1, 1, 1, 1, 1), src=(0.0, 1.0, 0.4), gauss_width=0.2,
                         velocity_method=7)
        d.wind_ustar, d.wind_z0, d.integrator = 8.0, 0.02, integ
        d.hills = list(hills) if integ == 1 else []
        o = O.Oracle(d, integrator=integ)
        o.frame_setup(0)
        o.step(200)
        peaks[integ] = max(np.abs(a).max() for a in o.state())
    assert peaks[1] < 1.0
    assert peaks[0] > 1e3


def _terrain_deck(lee, hills, **kw):
    dk = lee.deck.box(24, 26, 18, 0.05, pml=4, bc=(2, 1, 2, 2, 2, 2), src=(0.1, 0.6, 0.3), gauss_width=0.25, **kw)
    dk.hills = hills
    return dk


def test_terrain_metrics_invert_the_discrete_jacobian(lee):
    """The nine metric terms equal the inverse of the 3x3 matrix of central differences of
    the physical coordinates X(i), Y(j), Z(i,j,k) = Z_k + h(X,Y)(top - Z_k)/(top - Z_0);
    a dropped cross term, a sign or a transposed index fails this."""
    hills = [(0.2, 0.5, 0.3, 0.25), (-0.3, 0.9, 0.15, 0.4)]
    dk = _terrain_deck(lee, hills)
    o = O.Oracle(dk)
    o.frame_setup(0)
    X, Y, Zf = dk.air_x, dk.air_y, dk.air_z
    top, D = Zf[-1], Zf[-1] - Zf[0]

    def h(x, y):
        return sum(H * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * s ** 2)) for x0, y0, H, s in hills)

    def Z(i, j, k):
        return Zf[k] + h(X[i], Y[j]) * (top - Zf[k]) / D

    dxi, dyi, dzi = 1 / dk.dx, 1 / dk.dy, 1 / dk.dz
    worst = 0.0
    for (i, j, k) in [(5, 7, 3), (12, 10, 1), (3, 20, 9), (18, 4, 15), (10, 13, 8)]:
        M = np.array([
            [(X[i + 1] - X[i - 1]) * dxi / 2, 0.0, 0.0],
            [0.0, (Y[j + 1] - Y[j - 1]) * dyi / 2, 0.0],
            [(Z(i + 1, j, k) - Z(i - 1, j, k)) * dxi / 2, (Z(i, j + 1, k) - Z(i, j - 1, k)) * dyi / 2,
             (Z(i, j, k + 1) - Z(i, j, k - 1)) * dzi / 2]])
        inv = np.linalg.inv(M)
        m, J = o.metric(i, j, k)
        worst = max(worst, np.abs(m - inv).max() / np.abs(inv).max(), abs(J - 1 / np.linalg.det(M)) * abs(np.linalg.det(M)))
        assert abs(m[2, 0]) > 1e-3 and abs(m[2, 1]) > 1e-3  # the cross terms are exercised
    assert worst < 1e-12


@pytest.mark.parametrize("integrator", [0, 1])
def test_flat_terrain_is_the_separable_grid(lee, integrator):
    """A hill of zero height gives bit-identical results to the separable grid."""
    kw = dict(time_domain=15)
    a = _terrain_deck(lee, [(0.0, 0.6, 0.0, 0.3)], **kw)
    b = _terrain_deck(lee, [], **kw)
    a.integrator = b.integrator = integrator
    rec = [(0.0, 0.7, 0.4)]
    oa, ob = O.Oracle(a, receivers=rec), O.Oracle(b, receivers=rec)
    oa.frame_setup(0); ob.frame_setup(0)
    ra, rb = oa.step(15), ob.step(15)
    assert np.array_equal(ra, rb)
    for x, y in zip(oa.state(), ob.state()):
        assert np.array_equal(x, y)


def test_terrain_shapes_the_initial_pulse_and_the_wind(lee):
    """Over a hill the Gaussian pulse is centred at the source height above the local ground,
    and the log wind is V = u*/0.41 ln(z/z0) in the height above the local ground."""
    hills = [(0.1, 0.6, 0.25, 0.3)]
    dk = _terrain_deck(lee, hills, velocity_method=7)
    dk.wind_ustar, dk.wind_z0 = 0.4, 0.02
    o = O.Oracle(dk)
    o.frame_setup(0)
    X, Y, Zf = dk.air_x, dk.air_y, dk.air_z
    top, D = Zf[-1], Zf[-1] - Zf[0]
    hh = lambda x, y: sum(H * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * s ** 2)) for x0, y0, H, s in hills)
    Hm = hh(X[None, :], Y[:, None])  # (ny, nx)
    Zp = Zf[:, None, None] + Hm[None] * ((top - Zf) / D)[:, None, None]
    zs = dk.src[2] + hh(dk.src[0], dk.src[1])
    r2 = (X[None, None, :] - dk.src[0]) ** 2 + (Y[None, :, None] - dk.src[1]) ** 2 + (Zp - zs) ** 2
    p0 = np.exp(-4 * np.log(2) / dk.gauss_width ** 2 * r2)
    assert np.abs(o.state()[3] - p0).max() < 1e-14
    U, V, W = o.wind()
    zg = Zp - Zp[0:1]
    Vexp = np.where(zg > 0.02, 0.4 / 0.41 * np.log(np.maximum(zg, 1e-300) / 0.02), 0.0)
    assert np.abs(V - Vexp).max() < 1e-13
    assert np.abs(U).max() == 0 and np.abs(W).max() == 0
