"""Pins of the CPU oracle by mathematics, for the parts the reference probe cannot run
(calculation_2D.cpp and airpore.cpp do not build here): the FB time step with its face
conditions and phase order, the East extrapolation, the TDBC recursion, the air/porous
interface, the moving frame, the receiver rule and the FFT.

Each check is chosen so that a plausible mistake (a swapped one-sided difference, a
missing coefficient, a wrong plane in the interface copy, an off-by-one in the frame
shift) fails it; none of them restates the oracle's formulas.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _mode_deck(lee, axis, n, d=0.05, stretch=1.0, porous=False):
    """A domain long along `axis` and 5 points across, rigid walls everywhere."""
    dims = [5, 5, 5]
    dims[axis] = n
    dk = lee.deck.box(*dims, d, bc=(1, 3 if porous else 1, 1, 1, 1, 1), gauss_width=0.2, src=(0.0, 0.0, -50.0))
    coords = [dk.air_x, dk.air_y, dk.air_z]
    coords[axis] = stretch * d * np.arange(n)
    dk.air_x, dk.air_y, dk.air_z = coords
    return dk


def _fb_mode(P, U, a, b, steps):
    """Closed-form FB evolution of one standing mode:  P' = P - a U ;  U' = U + b P'."""
    out = []
    for _ in range(steps):
        P = P - a * U
        U = U + b * P
        out.append((P, U))
    return out


@pytest.mark.parametrize("axis,stretch", [(0, 1.0), (2, 1.0), (0, 2.0), (1, 1.0)])
def test_fb_standing_mode_closed_form(lee, axis, stretch):
    """A discrete standing wave p_i = P cos(t(i-1/2)), u_i = U sin(t(i-1)) with t = m pi/(N-2)
    satisfies every face condition of a rigid box and evolves by the 2x2 map
    P' = P - (gP dt/h) 2 sin(t/2) U,  U' = U + (dt/(rho h)) 2 sin(t/2) P'.
    h is the physical spacing (stretch * input spacing: the metric scales the difference).
    Along y the East face is the quadratic extrapolation, which the interior never reads
    (backward differences on p), so the mode holds on j <= N-2."""
    n, m_, steps = 34, 3, 25
    dk = _mode_deck(lee, axis, n, stretch=stretch)
    o = O.Oracle(dk)
    o.frame_setup(0)
    th = m_ * np.pi / (n - 2)
    idx = np.arange(n)
    prof_p = np.cos(th * (idx - 0.5))
    prof_u = np.sin(th * (idx - 1.0))
    shp = o.shape  # (nz, ny, nx)
    bshape = [1, 1, 1]
    bshape[2 - axis] = n
    P0, U0 = 1.0, 0.3
    fields = [np.zeros(shp) for _ in range(4)]
    vel = {0: 0, 1: 1, 2: 2}[axis]
    fields[3][...] = P0 * prof_p.reshape(bshape)
    fields[vel][...] = U0 * prof_u.reshape(bshape)
    o.set_state(*fields)
    o.step(steps)
    h = stretch * 0.05
    gP = dk.gamma * dk.Pav
    rho_inv = dk.c ** 2 / gP
    a = gP * dk.dt / h * 2 * np.sin(th / 2)
    b = rho_inv * dk.dt / h * 2 * np.sin(th / 2)
    P, U = _fb_mode(P0, U0, a, b, steps)[-1]
    st = o.state()
    p, vfield = st[3], st[vel]
    sl = [slice(None)] * 3
    keep = n - 1 if axis == 1 else n
    sl[2 - axis] = slice(0, keep)
    sl = tuple(sl)
    exp_p = np.broadcast_to((P * prof_p).reshape(bshape), shp)
    exp_u = np.broadcast_to((U * prof_u).reshape(bshape), shp)
    assert np.abs(p[sl] - exp_p[sl]).max() < 1e-11 * abs(P0)
    # the normal velocity on the first plane is the wall value 0, which no interior
    # stencil reads; the mode is compared from plane 1 on
    slu = list(sl)
    slu[2 - axis] = slice(1, keep)
    slu = tuple(slu)
    assert np.abs(vfield[slu] - exp_u[slu]).max() < 1e-11 * max(abs(U), 1e-3)
    # the other two velocity components stay exactly zero
    for q in range(3):
        if q != vel:
            assert np.abs(st[q]).max() == 0.0


def test_east_extrapolation_reproduces_quadratics(lee):
    """p at j = J-1 is the quadratic through J-2, J-3, J-4 evaluated half a cell out
    (15/8, -5/4, 3/8 are the Lagrange weights at J-3/2)."""
    dk = lee.deck.box(6, 12, 6, 0.05, bc=(1, 1, 1, 1, 1, 1))
    o = O.Oracle(dk)
    o.frame_setup(0)
    j = np.arange(12.0)
    q = 0.3 - 0.07 * j + 0.011 * j ** 2
    p = np.broadcast_to(q.reshape(1, 12, 1), o.shape).copy()
    z = np.zeros(o.shape)
    o.set_state(z, z, z, p)
    o.step(1)
    pn = o.state()[3]
    assert np.allclose(pn[2:4, -1, 2:4], 0.3 - 0.07 * 10.5 + 0.011 * 10.5 ** 2, rtol=0, atol=1e-14)


def test_tdbc_recursion_invariant(lee):
    """After each step the East-face velocity and the six histories satisfy
    v + sum_k Psi_k = p/Z_inf exactly (the recursion of Cal_TDBC / Update_Psi), where
    Z_inf = rho c q/Omega with c = 340.29 and rho = gamma P0 / c^2."""
    dk = lee.deck.box(10, 16, 10, 0.05, bc=(1, 1, 5, 1, 1, 1), src=(0.0, 0.55, 0.25), gauss_width=0.2)
    o = O.Oracle(dk)
    o.frame_setup(0)
    rho = dk.gamma * dk.Pav / 340.29 / 340.29
    Z = rho * 340.29 * dk.pore_q / dk.pore_porosity
    seen = 0.0
    for _ in range(12):
        o.step(1)
        u, v, w, p = o.state()
        psi = o.psi()  # [k][i][6]
        lhs = v[:, -1, 1:-1] + psi[:, 1:-1, :].sum(axis=2)
        rhs = p[:, -1, 1:-1] / Z
        assert np.abs(lhs - rhs).max() <= 1e-15 + 1e-12 * np.abs(rhs).max()
        seen = max(seen, np.abs(psi).max())
    assert seen > 0


# The six-pole table of the TDBC (calculation_2D.h:73-75): a constant table, the model's data
TD_A = np.array([8.226223820e-01, 3.732304064e-01, 9.106232372e-02, 4.819445335e-02, 4.225250937e-01,
                 1.837970986e-01])
TD_R = np.array([8.681310138e-01, 1.953722814e-01, 1.108067950e-02, 2.566498621e-03, 1.457297589e-04,
                 4.648583789e-02])


def _tdbc_run(lee, sigma, steps):
    """East-face velocity and histories of an oracle run with the East TDBC, step by step."""
    dk = lee.deck.box(10, 14, 10, 0.05, bc=(1, 1, 5, 1, 1, 1), src=(0.0, 0.5, 0.25), gauss_width=0.2)
    dk.pore_q, dk.pore_porosity, dk.pore_resistivity = 3.0 ** 0.5, 0.3, sigma
    o = O.Oracle(dk)
    o.frame_setup(0)
    vs, ps, psis = [], [], []
    for _ in range(steps):
        o.step(1)
        u, v, w, p = o.state()
        vs.append(v[1:-1, -1, 1:-1].T.copy())  # (i, k) of the face's interior
        ps.append(p[1:-1, -1, 1:-1].T.copy())
        psis.append(o.psi()[1:-1, 1:-1, :].transpose(1, 0, 2).copy())  # (i, k, pole)
    return dk, np.array(vs), np.array(ps), np.array(psis)


def _continuous_histories(v, dt, tau, a, r):
    """The histories of the continuous impedance Z(s) = Z_inf (1 + sum_k a_k/(r_k + s tau)):
    psi_k(t) = (a_k/tau) int_0^t exp(-r_k (t - t')/tau) v(t') dt', integrated exactly for a
    velocity held at v_n over (t_{n-1}, t_n]."""
    x = r * dt / tau
    decay, gain = np.exp(-x), a / r * -np.expm1(-x)
    psi = np.zeros(v.shape[1:] + (len(a),))
    out = []
    for n in range(v.shape[0]):
        psi = decay * psi + gain * v[n][..., None]
        out.append(psi.copy())
    return np.array(out)


def test_tdbc_histories_follow_the_continuous_impedance(lee):
    """The East TDBC discretises the impedance Z(s) = Z_inf (1 + sum_k a_k/(r_k + s tau)) with
    Z_inf = rho c q/Omega and tau = gamma rho q^2/(sigma Omega), rho = gamma P0/c^2 and
    c = 340.29 m/s (airpore.cpp:252-254): p/Z_inf = v + sum_k psi_k with psi_k the memory
    integral of v.  The oracle's histories must equal the continuous memory integral of its own
    face velocity up to the discretisation error, which is first order in r_k dt/tau (the
    recursion's gain a_k dt/tau against the exact a_k/r_k (1 - exp(-r_k dt/tau)), so the
    relative error per pole is below r_k dt/tau / 2), and halve with dt/tau.  A wrong decay
    (sign, a missing gamma in tau, dt for dt/tau) or gain fails by orders of magnitude more."""
    errs = []
    for sigma in (4000.0, 2000.0):  # tau doubles: dt/tau halves
        dk, v, p, psi = _tdbc_run(lee, sigma, 60)
        rho = dk.gamma * dk.Pav / 340.29 ** 2
        zinf = rho * 340.29 * dk.pore_q / dk.pore_porosity
        tau = dk.gamma * rho * dk.pore_q ** 2 / sigma / dk.pore_porosity
        tb = dk.dt / tau
        assert 1e-3 < tb < 2e-2
        assert np.abs(v).max() > 1e-6
        cont = _continuous_histories(v, dk.dt, tau, TD_A, TD_R)
        # per pole: the relative difference stays inside the first-order bound
        for q in range(6):
            e = np.linalg.norm(psi[..., q] - cont[..., q]) / np.linalg.norm(cont[..., q])
            assert e <= 0.5 * TD_R[q] * tb * 1.02 + 1e-12, (sigma, q, e)
        errs.append(np.linalg.norm(psi - cont) / np.linalg.norm(cont))
        # the face relation of the model, with Z_inf from its closed form
        lhs = v + psi.sum(axis=-1)
        assert np.abs(lhs - p / zinf).max() <= 1e-12 * np.abs(p / zinf).max()
        # plausible mistakes in the decay or the time constant are far outside that bound
        for wrong_tau in (tau / dk.gamma, tau * dk.gamma, 1.0):
            bad = _continuous_histories(v, dk.dt, wrong_tau, TD_A, TD_R)
            assert np.linalg.norm(psi - bad) / np.linalg.norm(bad) > 20 * errs[-1]
    # first order in dt/tau
    assert 0.4 < errs[1] / errs[0] < 0.6, errs


def _interface_pair(lee, Kp=7, Ka=12, n=(14, 16)):
    d = 0.05
    common = dict(bc=(1, 3, 1, 1, 1, 1), gauss_width=0.2, src=(0.05, 0.4, 0.1))
    coupled = lee.deck.box(n[0], n[1], Ka, d, porous_cells=Kp, **common)
    # matched ground: q = 1, porosity 1, no flow resistance -> the porous equations are
    # the air equations (rho_e = rho0, Gamma = gamma P0 dt, beta = 0)
    coupled.pore_q, coupled.pore_porosity, coupled.pore_resistivity = 1.0, 1.0, 0.0
    z_single = np.concatenate([coupled.pore_z[:Kp - 1], coupled.air_z[1:]])
    single = lee.deck.box(n[0], n[1], Ka + Kp - 2, d, bc=(1, 1, 1, 1, 1, 1), gauss_width=0.2, src=(0.05, 0.4, 0.1))
    single.air_z = z_single
    return coupled, single, Kp


def test_interface_couples_the_right_planes(lee):
    """With ground parameters that make the porous equations the air equations, the
    air+porous pair (two overlapping ghost planes: air k=0 <- pore Kp-2, pore Kp-1 <- air
    k=1) must evolve exactly like one air column of Ka+Kp-2 planes."""
    coupled, single, Kp = _interface_pair(lee)
    a = O.Oracle(coupled)
    b = O.Oracle(single)
    a.frame_setup(0); b.frame_setup(0)
    a.step(30); b.step(30)
    pa, pp, pb = a.state(0)[3], a.state(1)[3], b.state(0)[3]
    scale = np.abs(pb).max()
    # porous planes 0..Kp-2 are single planes 0..Kp-2; air planes 1.. are single Kp-1..
    assert np.abs(pp[:Kp - 1] - pb[:Kp - 1]).max() < 1e-11 * scale
    assert np.abs(pa[1:] - pb[Kp - 1:]).max() < 1e-11 * scale
    wa, wb = a.state(0)[2], b.state(0)[2]
    assert np.abs(wa[1:] - wb[Kp - 1:]).max() < 1e-11 * np.abs(wb).max()


def test_porous_drag_and_compressibility(lee):
    """Increments of the porous stencils on a 1-D mode: dp = -Gamma (2/h) sin(t/2) U cos(.),
    du = (1/(1+beta) - 1) u - dt/(rho_e (1+beta)) (2/h) sin(t/2) P sin(.) with
    rho_e = q^2 rho0 / Omega, beta = sigma dt / rho_e, Gamma = gamma P0 dt / Omega."""
    dk = lee.deck.box(5, 5, 5, 0.05, bc=(1, 3, 1, 1, 1, 1), porous_cells=5)
    n, h = 30, 0.05
    dk.pore_x = h * np.arange(n); dk.air_x = dk.pore_x.copy()
    dk.pore_q, dk.pore_porosity, dk.pore_resistivity = 1.6, 0.35, 8.0e4
    o = O.Oracle(dk)
    o.frame_setup(0)
    th = 2 * np.pi / (n - 2)
    i = np.arange(n)
    shp = (o.pnz, o.pny, o.pnx)
    P, U = 0.7, 0.2
    p = np.broadcast_to(P * np.cos(th * (i - 0.5)), shp).copy()
    u = np.broadcast_to(U * np.sin(th * (i - 1.0)), shp).copy()
    z = np.zeros(shp)
    dp = o.pore_dp(u, z, z)
    du, dv, dw = o.pore_duvw(u, z, z, p)
    rho0 = dk.gamma * dk.Pav / dk.c ** 2
    rho_e = dk.pore_q ** 2 * rho0 / dk.pore_porosity
    beta = dk.pore_resistivity * dk.dt / rho_e
    gama = dk.gamma * dk.Pav * dk.dt / dk.pore_porosity
    s2 = 2 * np.sin(th / 2) / h
    exp_dp = -gama * s2 * U * np.cos(th * (i - 0.5))
    exp_du = (1 / (1 + beta) - 1) * U * np.sin(th * (i - 1.0)) + dk.dt / (rho_e * (1 + beta)) * s2 * P * np.sin(th * (i - 1.0))
    sl = (slice(1, -1), slice(1, -1), slice(1, -1))
    assert np.allclose(dp[sl], np.broadcast_to(exp_dp, shp)[sl], rtol=0, atol=1e-10 * abs(exp_dp).max())
    assert np.allclose(du[sl], np.broadcast_to(exp_du, shp)[sl], rtol=0, atol=1e-10 * abs(exp_du).max())
    assert np.abs(dv[sl]).max() == 0 and np.abs(dw[sl]).max() == 0


def test_moving_frame_matches_static_domain(lee):
    """The frame run (shift by trail-1 planes, leading planes zeroed) reproduces a static
    run over the whole propagation length in the part of the frame the frame's own walls
    have not influenced.  The frame is used as intended: open duct (x and z layers, which
    shift with the frame), wavefront well inside the frame.  The reference leaves plane
    JMAX-lead unshifted on the last rank (calculation_2D.cpp:466-476), which is harmless
    only there; the leading planes are zeroed."""
    L, Jf, lead, T = 100, 60, 7, 110
    common = dict(pml=5, bc=(2, 1, 1, 1, 2, 2), gauss_width=0.2, src=(0.0, 0.15, 0.3), time_domain=T, c=343.0)
    frame = lee.deck.box(16, L, 16, 0.05, mf_judge=1, mf_jmax=Jf, mf_lead_dj=lead, mf_trail_dj=lead,
                         mf_ratio=0.4, ijkmax=L, **common)
    static = lee.deck.box(16, L, 16, 0.05, **common)
    a = O.Oracle(frame)
    b = O.Oracle(static)
    a.run(); b.run()
    temp, mf, last = 0, 0, a.frame_count() - 1
    while True:  # replay the frame schedule to find the last frame
        lim = a.frame_steps(mf, temp)
        temp += lim
        if mf >= last or temp >= T:
            break
        mf += 1
    shift = mf * (lead - 1)
    pa = a.state()[3]
    pb = b.state()[3][:, shift:shift + Jf, :]
    scale = np.abs(pb).max()
    assert mf >= 5
    # the trailing wall's influence decays by ~1e-3 every four planes into the frame
    clean = slice(20, Jf - lead - 1)
    assert np.abs(pa[:, clean, :]).max() > 1e-2 * scale
    assert np.abs(pa[:, clean, :] - pb[:, clean, :]).max() < 1e-10 * scale
    # a frame shifted by one plane too many or too few is far off
    off = b.state()[3][:, shift + 1:shift + 1 + Jf, :]
    assert np.abs(pa[:, clean, :] - off[:, clean, :]).max() > 1e-2 * scale


def test_stale_time_level_values_never_leak(lee):
    """The reference reads not-yet-written n+1 boundary values inside its face chain
    (N before W/F ...) and overwrites them afterwards.  Filling level n+1 with NaN must
    change nothing: every point of n+1 is rewritten each step."""
    dk = lee.deck.box(16, 18, 14, 0.05, pml=4, bc=(2, 3, 5, 2, 2, 2), porous_cells=6, src=(0.0, 0.45, 0.3),
                      gauss_width=0.2, pml_y_time=2, velocity_method=5, velocity_coef=1.0)
    a = O.Oracle(dk, fill_nn_nan=True)
    b = O.Oracle(dk)
    a.frame_setup(0); b.frame_setup(0)
    ra, rb = a.step(12), b.step(12)
    for x, y in zip(a.state(0) + a.state(1), b.state(0) + b.state(1)):
        assert np.isfinite(x).all()
        assert np.array_equal(x, y)
    assert np.array_equal(ra, rb)


def test_initial_condition_is_the_gaussian_pulse(lee):
    dk = lee.deck.box(9, 10, 11, 0.05, gauss_width=0.3, src=(0.02, 0.2, 0.15))
    o = O.Oracle(dk)
    p = o.state()[3]
    X, Y, Z = np.meshgrid(dk.air_x, dk.air_y, dk.air_z, indexing="ij")
    r2 = (X - 0.02) ** 2 + (Y - 0.2) ** 2 + (Z - 0.15) ** 2
    ref = np.exp(-4 * np.log(2) * r2 / 0.3 ** 2).transpose(2, 1, 0)
    assert np.allclose(p, ref, rtol=1e-14, atol=1e-300)


def test_receiver_rule_brute_force(lee):
    """Index = first grid node with coordinate >= target - step/8 on each axis; positions
    beyond the domain map to (1,1,1) (airpore.cpp:455-456)."""
    rng = np.random.default_rng(4)
    dk = lee.deck.box(20, 24, 16, 0.05, y0=-1.0)
    pts = np.column_stack([rng.uniform(-0.5, 0.5, 40), rng.uniform(-1.1, 0.2, 40), rng.uniform(-0.1, 0.8, 40)])
    o = O.Oracle(dk, receivers=pts)
    got = o.receiver_indices()
    Yf = dk.air_y[-1] - dk.air_y[0]
    for (x, y, z), (i, j, k) in zip(pts, got):
        ii = np.nonzero(x - dk.air_x <= dk.dx / 8)[0]
        N = int(y / Yf + 1e-6)
        jj = np.nonzero(y - N * Yf - dk.air_y <= dk.dy / 8)[0] + N * (24 - 1)
        kk = np.nonzero(z - dk.air_z <= dk.dz / 8)[0]
        if len(ii) == 0 or len(jj) == 0 or len(kk) == 0 or N * 23 >= 24:
            assert (i, j, k) == (1, 1, 1)
        else:
            e = (ii[0], jj[0], kk[0])
            if e == (0, 0, 0) or e[1] > 24:
                e = (1, 1, 1)
            assert (i, j, k) == e


def test_fft_is_the_scaled_dft():
    """Forward = DFT/N (numpy), inverse unscaled.  The twiddle factor is advanced by
    repeated complex multiplication, so the error grows like N*eps (~1e-11 at N=4096)."""
    rng = np.random.default_rng(3)
    for m in (3, 8, 12):
        n = 1 << m
        x = rng.standard_normal(n)
        y = rng.standard_normal(n)
        fr, fi = O.fft(x, y, m)
        ref = np.fft.fft(x + 1j * y) / n
        tol = 1e-15 * n * np.abs(ref).max() + 1e-15
        assert np.abs(fr - ref.real).max() < tol and np.abs(fi - ref.imag).max() < tol
        br, bi = O.fft(fr, fi, m, direction=-1)
        tol = 1e-15 * n * np.abs(x).max()
        assert np.abs(br - x).max() < tol and np.abs(bi - y).max() < tol


def test_pml_absorbs_and_stays_bounded(lee):
    """A pulse launched into a fully absorbing box: the acoustic energy of the inner
    region must fall to a few per cent of its initial value once the pulse has reached
    the layers, and stay there over a long run (a sign or cross-term error in the layer
    shows up as growth).  The reference layer (no triple-corner term, kappa-stretch with
    Q integrals) leaves a small bounded residue rather than decaying to zero."""
    dk = lee.deck.box(22, 22, 22, 0.05, pml=6, bc=(2, 2, 2, 2, 2, 2), src=(0.0, 0.55, 0.55), gauss_width=0.15,
                      pml_y_time=1)
    o = O.Oracle(dk)
    o.frame_setup(0)
    rho0 = dk.gamma * dk.Pav / dk.c ** 2

    def energy():
        u, v, w, p = o.state()
        inner = (slice(7, -7),) * 3
        return 0.5 * (p[inner] ** 2 / (rho0 * dk.c ** 2) + rho0 * (u[inner] ** 2 + v[inner] ** 2 + w[inner] ** 2)).sum()

    e0 = energy()
    es = []
    for _ in range(40):
        o.step(10)
        es.append(energy())
    assert max(es[5:]) < 0.1 * e0
    assert max(es[20:]) < 3 * max(es[5:20])
