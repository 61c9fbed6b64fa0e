"""Learned ground: ground parameters -> pole/residue pairs of the ground impedance boundary.

The full solve models the ground as the porous layer of porous.cpp under the air (South face
porous, FB_p_v, the air reading the layer's top plane through the interface copy of
calculation_2D.cpp:878-886).  The learned ground drops the layer: the air's plane k = 0
becomes the ground's interface plane, which carries the porous equations of the top
material, and the layer under it is replaced by the impedance of lee_options.ground_*
(K = 8 real poles, Z(s) = Z_inf (1 + sum_k a_k / (r_k + s tau)), Z_inf = rho0 c q/porosity,
tau = rho_e / sigma, rho_e = q^2 rho0 / porosity).  The poles sit on a fixed grid in units
of 1/tau (R_GRID); a small network (lee_mlp: BF16 MFMA, trained on the GPU) maps (sigma,
porosity, tortuosity, layer depth) to the residues.

Training data come from runs of the full air + porous solve: a plane pulse in a rigid-walled
column falls on the layer, and the incident and reflected pulses at a receiver give R(f)
(divided by the same ratio over a rigid floor, which removes the air path's dispersion).
The boundary's own discretisation is taken into account exactly: for a fixed closure the
column's reflection is a Moebius function of the recursion's discrete impedance Z_d(f)
(one scalar relation closes a linear system), identified per frequency from four runs of
the impedance closure with known Z_d; inverting it turns the measured R(f) into the Z_d(f)
the boundary must realise, and Z_d is linear in the residues, so a non-negative least
squares fit (weighted by |dR/dZ_d|) gives them -- non-negative residues keep the boundary
passive and its recursion stable.

The full porous solve stays the judge and the fallback: `LearnedGround.decide` predicts the
residues, measures the reflection they give against the porous column's, and returns the
impedance deck only inside the training hull and within the error bound.
"""
from __future__ import annotations

import ctypes
import importlib
from dataclasses import dataclass

import numpy as np
from scipy.optimize import nnls

K = 8
# pole grid in units of 1/tau: one pole at zero (the layer's compliance over a rigid backing)
R_GRID = np.concatenate([[0.0], np.logspace(-2.5, 2.0, K - 1)])
# Below ~200 kPa s/m^2 the layer is extended-reacting at the grazing angles of outdoor
# propagation (its refraction index is small enough that the wave runs along inside it), which
# a locally reacting impedance cannot represent: measured in the judge box, the one-third-
# octave IL difference grows with the angle of incidence there (0.03 dB normal, 0.35 dB at
# 69 degrees for 74 kPa s/m^2) and stays <= 0.02 dB at all angles for 500-1500 kPa s/m^2.
HULL_SIGMA_MIN = 2.0e5
HULL_FIT_R = 0.05
JUDGE_BOUND_DB = 0.2          # one-third-octave IL bound of the judge box
F_BAND = (50.0, 1000.0)       # the accuracy band
F_FIT = (20.0, 1500.0)        # the fitted band
USE_GRAPH = True              # column and judge runs replay captured HIP graphs
# ground-parameter distribution of the ensemble sweep and the learned-ground workload
SIGMA_RANGE = (1.0e4, 2.0e6)
POROSITY_RANGE = (0.2, 0.5)
TORTUOSITY_RANGE = (1.5, 4.0)
DEPTH_POINTS = (4, 16)


def _lee():
    return importlib.import_module(__package__ or "3d_mpi_test_amd")


@dataclass(frozen=True)
class Ground:
    sigma: float        # flow resistivity [Pa s / m^2]
    porosity: float
    tortuosity: float   # q^2
    depth: int          # porous layer points below the interface plane


def normalise(g: Ground):
    """Network inputs in [-1, 1]."""
    ls = (np.log10(g.sigma) - np.log10(SIGMA_RANGE[0])) / (np.log10(SIGMA_RANGE[1]) - np.log10(SIGMA_RANGE[0]))
    po = (g.porosity - POROSITY_RANGE[0]) / (POROSITY_RANGE[1] - POROSITY_RANGE[0])
    to = (g.tortuosity - TORTUOSITY_RANGE[0]) / (TORTUOSITY_RANGE[1] - TORTUOSITY_RANGE[0])
    de = (g.depth - DEPTH_POINTS[0]) / (DEPTH_POINTS[1] - DEPTH_POINTS[0])
    return np.array([2 * ls - 1, 2 * po - 1, 2 * to - 1, 2 * de - 1], dtype=np.float32)


def in_hull(g: Ground):
    """Inside the sampled parameter box, with a flow resistivity the pole grid resolves."""
    return (HULL_SIGMA_MIN <= g.sigma <= SIGMA_RANGE[1] and POROSITY_RANGE[0] <= g.porosity <= POROSITY_RANGE[1]
            and TORTUOSITY_RANGE[0] <= g.tortuosity <= TORTUOSITY_RANGE[1] and DEPTH_POINTS[0] <= g.depth <= DEPTH_POINTS[1])


def zinf_tau(g: Ground, rho0: float, c: float):
    """Z_inf and tau of the porous equations (porous.cpp:40-51 with airpore.cpp:249-251)."""
    rho_e = g.tortuosity * rho0 / g.porosity
    return rho0 * c * np.sqrt(g.tortuosity) / g.porosity, rho_e / g.sigma


def sample_grounds(n: int, seed: int):
    """Grounds drawn from the sweep distribution: sigma log-uniform in SIGMA_RANGE, porosity,
    tortuosity and depth uniform in their ranges."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        out.append(Ground(float(10 ** rng.uniform(np.log10(SIGMA_RANGE[0]), np.log10(SIGMA_RANGE[1]))),
                          float(rng.uniform(*POROSITY_RANGE)), float(rng.uniform(*TORTUOSITY_RANGE)),
                          int(rng.integers(DEPTH_POINTS[0], DEPTH_POINTS[1] + 1))))
    return out


def set_top_material(deck, g: Ground):
    deck.pore_q, deck.pore_porosity, deck.pore_resistivity = float(np.sqrt(g.tortuosity)), g.porosity, g.sigma


def porous_deck(deck, g: Ground):
    """`deck` (a Deck built by deck.box with porous_cells = g.depth + 2) with g's material."""
    set_top_material(deck, g)
    return deck


def impedance_deck(deck, g: Ground, residues, zinf=None, tau=None, poles=None):
    """Set the ground impedance of `deck` (rigid South face) from residues on R_GRID (or on
    `poles`): Z_inf and tau of g's material unless given, the top material of g on the
    interface plane."""
    rho = deck.gamma * deck.Pav / deck.c ** 2
    zi, ta = zinf_tau(g, rho, deck.c)
    r = R_GRID if poles is None else poles
    deck.ground_poles = [(float(max(a, 0.0)), float(rr)) for a, rr in zip(residues, r)]
    deck.ground_zinf = float(zi if zinf is None else zinf)
    deck.ground_tau = float(ta if tau is None else tau)
    set_top_material(deck, g)
    return deck


# ---------------------------------------------------------------------- reflection data
# reflection column: plane pulse released at Z_SRC, receiver at Z_REC, N_AIR points at DZ; the
# reflected window closes at T_END_M / c, before the pulse's upward tail comes back from the top
Z_SRC, Z_REC, N_AIR, DZ, COLUMN_STEPS, T_END_M = 6.0, 3.0, 640, 0.04, 2048, 40.0


def column_deck(kind: str, g: Ground | None = None, residues=None, zinf=None, tau=None):
    """A 6 x 6 column of N_AIR points with rigid side walls and a 20-point PML at the top over
    a rigid floor (kind 'rigid'), g's porous layer ('porous') or the impedance closure with
    g's top material and the given residues ('impedance'); FB_p_v at CFL 0.5 in 3-D."""
    lee = _lee()
    porous = kind == "porous"
    d = lee.deck.box(6, 6, N_AIR, DZ, pml=20, bc=(2, 3 if porous else 1, 1, 1, 1, 1),
                     porous_cells=(g.depth + 2) if porous else 0, time_domain=COLUMN_STEPS, src=(0.0, 0.0, -100.0))
    if porous:
        porous_deck(d, g)
    elif kind == "impedance":
        impedance_deck(d, g, residues, zinf, tau)
    return d


def _plane_pulse(solver, d, z0, width):
    """A pulse travelling down: p = exp(-((z - z0)/width)^2), w = -p/(rho0 c)."""
    rho = d.gamma * d.Pav / d.c ** 2
    z = np.asarray(d.air_z)[:, None, None] * np.ones((len(d.air_z), len(d.air_y), len(d.air_x)))
    p = np.exp(-((z - z0) / width) ** 2)
    zero = np.zeros_like(p)
    solver.set_state(zero, zero, -p / (rho * d.c), p)


def column_series(d, precision=64):
    """Receiver pressure at Z_REC of the plane pulse from Z_SRC in column deck d."""
    lee = _lee()
    rec = [(float(d.air_x[2]), float(d.air_y[2]), Z_REC)]
    s = lee.Solver(d, precision=precision, receivers=rec, use_graph=USE_GRAPH)
    try:
        s.frame_setup(0)
        _plane_pulse(s, d, Z_SRC, 3 * DZ)
        return s.step(d.time_domain)[:, 0]
    finally:
        s.close()


def reflection(series, ref_series, dt, c, z_src=Z_SRC, z_rec=Z_REC, t_end=None, nfft=8192):
    """R(f) of the ground from the receiver series over it and over a rigid floor: the
    incident and reflected windows are split half-way between their arrivals; the reflected
    window ends at t_end (default: the end of the series), before any later arrival."""
    t_split = ((z_src - z_rec) / c + (z_src + z_rec) / c) / 2
    n_split = int(t_split / dt)
    n_end = len(series) if t_end is None else min(len(series), int(t_end / dt))

    def ratio(x):
        inc = np.zeros(nfft)
        ref = np.zeros(nfft)
        inc[:n_split] = x[:n_split]
        ref[:n_end - n_split] = x[n_split:n_end]
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.fft.rfft(ref) / np.fft.rfft(inc)
    f = np.fft.rfftfreq(nfft, dt)
    with np.errstate(divide="ignore", invalid="ignore"):   # beyond the pulse's band
        return f, ratio(series) / ratio(ref_series)


def discrete_impedance(f, a, zinf, tau, dt, poles=None):
    """Z_d(f) of the recursion: Zinf (1 + sum_k a_k tb / (1 - e^{-r_k tb} z^-1)), z = e^{i w dt},
    tb = dt/tau (the z-transform of Cal_TDBC / Update_Psi, calculation_2D.cpp:1012-1028);
    poles r_k default to R_GRID."""
    r = R_GRID if poles is None else np.asarray(poles, float)
    tb = dt / tau
    zi = np.exp(-2j * np.pi * np.asarray(f) * dt)
    B = tb / (1.0 - np.exp(-r[None, :] * tb) * zi[:, None])
    return zinf * (1.0 + B @ np.asarray(a, float))


# the reference's six-pole TDBC coefficients (calculation_2D.h:73-75): its fit of the
# semi-infinite layer's sqrt(1 + 1/(s tau)), the same for every ground
SIX_POLE_A = (8.226223820e-01, 3.732304064e-01, 9.106232372e-02, 4.819445335e-02, 4.225250937e-01, 1.837970986e-01)
SIX_POLE_R = (8.681310138e-01, 1.953722814e-01, 1.108067950e-02, 2.566498621e-03, 1.457297589e-04, 4.648583789e-02)


def moebius(Zs, Rs):
    """Per-frequency coefficients (al, be, ga, de) of R = (al + be Z)/(ga + de Z) from n >= 3
    pairs of (Z_d, R) arrays: the null vector of [1, Z, -R, -R Z]."""
    Zs, Rs = np.asarray(Zs), np.asarray(Rs)
    out = np.zeros((Zs.shape[1], 4), complex)
    for i in range(Zs.shape[1]):
        M = np.stack([np.ones(len(Zs)), Zs[:, i], -Rs[:, i], -Rs[:, i] * Zs[:, i]], axis=1)
        out[i] = np.linalg.svd(M)[2][-1].conj()
    return out


def moebius_apply(coef, Z):
    return (coef[:, 0] + coef[:, 1] * Z) / (coef[:, 2] + coef[:, 3] * Z)


@dataclass
class Reflection:
    """A ground's measured porous reflection and its impedance closure's Moebius map."""
    g: Ground
    f: np.ndarray        # fitted-band frequencies
    R: np.ndarray        # porous column reflection
    coef: np.ndarray     # Moebius map Z_d -> R of the closure
    zinf: float
    tau: float
    dt: float

    def predict(self, a):
        """Reflection of the impedance closure with residues a."""
        return moebius_apply(self.coef, discrete_impedance(self.f, a, self.zinf, self.tau, self.dt))

    def error(self, a, poles=None):
        """max |R_a - R_porous| over F_BAND (residues a on R_GRID, or on `poles`)."""
        b = (self.f >= F_BAND[0]) & (self.f <= F_BAND[1])
        Rp = moebius_apply(self.coef, discrete_impedance(self.f, a, self.zinf, self.tau, self.dt, poles))
        return float(np.abs(Rp[b] - self.R[b]).max())

    def fit(self, ridge=0.0):
        """Non-negative residues for the porous reflection; returns (a, error).  `ridge` adds
        ridge * (mean column norm)^2 * |a|^2 to the least squares (smoother residues across
        grounds, at some cost in fit error)."""
        co, R, f = self.coef, self.R, self.f
        Zreq = (co[:, 0] - co[:, 2] * R) / (co[:, 3] * R - co[:, 1])
        tb = self.dt / self.tau
        zi = np.exp(-2j * np.pi * f * self.dt)
        A = self.zinf * tb / (1.0 - np.exp(-R_GRID[None, :] * tb) * zi[:, None])
        w = np.abs((co[:, 1] * co[:, 2] - co[:, 0] * co[:, 3]) / (co[:, 2] + co[:, 3] * Zreq) ** 2)
        w = np.where(np.isfinite(w), w, 0.0)
        y = Zreq - self.zinf
        M = np.vstack([(A * w[:, None]).real, (A * w[:, None]).imag])
        rhs = np.concatenate([(y * w).real, (y * w).imag])
        ok = np.isfinite(rhs)
        M, rhs = M[ok], rhs[ok]
        if ridge > 0:
            scale = np.sqrt(ridge) * np.linalg.norm(M, axis=0).mean()
            M = np.vstack([M, scale * np.eye(K)])
            rhs = np.concatenate([rhs, np.zeros(K)])
        a, _ = nnls(M, rhs)
        return a, self.error(a)


class Columns:
    """Column runs on the GPU: the rigid reference once, then per ground the porous column
    and four impedance-closure columns of known Z_d."""

    def __init__(self, precision=64):
        self.precision = precision
        d = column_deck("rigid")
        self.dt, self.c = d.dt, d.c
        self.rho = d.gamma * d.Pav / d.c ** 2
        self.ref = column_series(d, precision)

    def _R(self, series):
        f, R = reflection(series, self.ref, self.dt, self.c, t_end=T_END_M / self.c)
        sel = (f >= F_FIT[0]) & (f <= F_FIT[1])
        return f[sel], R[sel]

    def porous(self, g: Ground):
        return self._R(column_series(column_deck("porous", g), self.precision))

    def impedance(self, g: Ground, a, zinf=None, tau=None):
        return self._R(column_series(column_deck("impedance", g, a, zinf, tau), self.precision))

    def measure(self, g: Ground) -> Reflection:
        z0 = self.rho * self.c
        zinf, tau = zinf_tau(g, self.rho, self.c)
        Zs, Rs = [], []
        a1 = np.zeros(K)
        a1[3] = 1.0
        for zc, a in ((z0, np.zeros(K)), (4 * z0, np.zeros(K)), (16 * z0, np.zeros(K)), (4 * z0, a1)):
            f, R = self.impedance(g, a, zc, tau)
            Zs.append(discrete_impedance(f, a, zc, tau, self.dt))
            Rs.append(R)
        f, R = self.porous(g)
        return Reflection(g, f, R, moebius(Zs, Rs), zinf, tau, self.dt)


# ridge weight of the training targets' fits: the plain non-negative fit's active set jumps
# between neighbouring grounds, which the network cannot follow; with this weight the fit
# errors stay the same (median 1.1e-3) and the network's held-out reflection error drops from
# 3.4e-3 to 1.4e-3 (median; 90th percentile 1.2e-2 to 5.5e-3) on 64 held-out grounds
TRAIN_RIDGE = 1e-4


def training_set(grounds, columns: Columns | None = None, progress=None, ridge=TRAIN_RIDGE):
    """(X [n x 4], A [n x K] fitted residues, E [n] fit errors) from the column runs."""
    cols = columns or Columns()
    X, A, E = [], [], []
    for n, g in enumerate(grounds):
        a, e = cols.measure(g).fit(ridge)
        X.append(normalise(g)); A.append(a); E.append(e)
        if progress:
            progress(n, g, e)
    return np.array(X, np.float32), np.array(A), np.array(E)


# ---------------------------------------------------------------------- network
class Network:
    """ctypes front of lee_mlp (the GPU network)."""

    def __init__(self, seed=1, device=0):
        L = _lee().lib()
        _setup(L)
        h = ctypes.c_void_p()
        rc = L.lee_mlp_create(device, seed, ctypes.byref(h))
        if rc != 0:
            raise RuntimeError("lee_mlp_create failed")
        self.h = h
        self.n = L.lee_mlp_param_count(h)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(_lee().lib().lee_mlp_error(self.h).decode())

    def params(self):
        w = np.zeros(self.n, np.float32)
        self._check(_lee().lib().lee_mlp_get_params(self.h, w.ctypes.data_as(_fp), self.n))
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._check(_lee().lib().lee_mlp_set_params(self.h, w.ctypes.data_as(_fp), w.size))

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], K), np.float32)
        self._check(_lee().lib().lee_mlp_predict(self.h, x.ctypes.data_as(_fp), x.shape[0], y.ctypes.data_as(_fp)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, np.float32); t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.n, np.float32)
        mse = ctypes.c_double()
        self._check(_lee().lib().lee_mlp_gradient(self.h, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp), x.shape[0],
                                                  g.ctypes.data_as(_fp), ctypes.byref(mse)))
        return g, mse.value

    def train(self, x, t, batch=32, epochs=200, lr=3e-3, seed=7):
        x = np.ascontiguousarray(x, np.float32); t = np.ascontiguousarray(t, np.float32)
        mse = ctypes.c_double()
        self._check(_lee().lib().lee_mlp_train(self.h, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp), x.shape[0], batch,
                                               epochs, ctypes.c_float(lr), seed, ctypes.byref(mse)))
        return mse.value

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            _lee().lib().lee_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_fp = ctypes.POINTER(ctypes.c_float)


def _setup(L):
    if getattr(L, "_mlp_ready", False):
        return
    vp = ctypes.c_void_p
    L.lee_mlp_create.argtypes = [ctypes.c_int, ctypes.c_ulonglong, ctypes.POINTER(vp)]
    L.lee_mlp_destroy.argtypes = [vp]
    L.lee_mlp_error.argtypes = [vp]
    L.lee_mlp_error.restype = ctypes.c_char_p
    L.lee_mlp_param_count.argtypes = [vp]
    L.lee_mlp_param_count.restype = ctypes.c_longlong
    L.lee_mlp_get_params.argtypes = [vp, _fp, ctypes.c_longlong]
    L.lee_mlp_set_params.argtypes = [vp, _fp, ctypes.c_longlong]
    L.lee_mlp_predict.argtypes = [vp, _fp, ctypes.c_int, _fp]
    L.lee_mlp_gradient.argtypes = [vp, _fp, _fp, ctypes.c_int, _fp, ctypes.POINTER(ctypes.c_double)]
    L.lee_mlp_train.argtypes = [vp, _fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_double)]
    L._mlp_ready = True



def il_errors(a, b, dt, fmin=F_BAND[0], fmax=F_BAND[1]):
    """Insertion-loss differences of receiver series b against a (columns = receivers, only
    receivers carrying signal), in dB over [fmin, fmax]: (largest narrow-band difference, bins
    60 dB below the band peak skipped; largest one-third-octave band-level difference, centre
    frequencies 1000 * 2^(n/3) in the band, edges at 2^(+-1/6))."""
    if not (np.isfinite(a).all() and np.isfinite(b).all()):
        return float("inf"), float("inf")  # a run that blew up matches nothing
    energy = (a ** 2).sum(axis=0)
    use = np.nonzero(energy > 1e-6 * energy.max())[0]
    n = 1 << int(np.ceil(np.log2(a.shape[0])))
    f = np.fft.rfftfreq(n, dt)
    band = (f >= fmin) & (f <= fmax)
    fc = [1000.0 * 2 ** (m / 3.0) for m in range(-30, 1)]
    fc = [x for x in fc if fmin * 0.999 <= x <= fmax * 1.001]
    narrow = third = 0.0
    for r in use:
        A = np.abs(np.fft.rfft(a[:, r], n))
        B = np.abs(np.fft.rfft(b[:, r], n))
        ok = band & (A > 1e-3 * A[band].max())
        if ok.any():
            narrow = max(narrow, float(np.abs(20 * np.log10(B[ok] / A[ok])).max()))
        for c in fc:
            sel = (f >= c * 2 ** (-1 / 6)) & (f < c * 2 ** (1 / 6))
            if sel.any():
                third = max(third, float(abs(10 * np.log10((B[sel] ** 2).sum() / (A[sel] ** 2).sum()))))
    return narrow, third


# ---------------------------------------------------------------------- 3-D judge
JUDGE_REC = [(0.0, y, h) for y in (2.6, 3.4) for h in (0.1, 0.6, 1.2)]


def judge_deck(g: Ground, kind: str, residues=None, steps=1024, poles=None, hills=None):
    """A 120 x 120 x 60 box at 0.04 m (20-point PML on the sides and top), Gaussian pulse at
    0.6 m, over g's porous layer ('porous') or the impedance closure ('impedance'); with
    `hills` ((x0, y0, height, sigma) each) terrain-following over them."""
    lee = _lee()
    kw = dict(pml=20, src=(0.0, 1.6, 0.6), gauss_width=0.24, time_domain=steps)
    if kind == "porous":
        d = lee.deck.box(120, 120, 60, DZ, bc=(2, 3, 2, 2, 2, 2), porous_cells=g.depth + 2, **kw)
        d = porous_deck(d, g)
    else:
        d = lee.deck.box(120, 120, 60, DZ, bc=(2, 1, 2, 2, 2, 2), **kw)
        d = impedance_deck(d, g, residues, poles=poles)
    if hills:
        d.hills = list(hills)
    return d


def judge_3d(g: Ground, residues, precision=64, poles=None, hills=None):
    """IL differences (narrow-band, one-third octave) of the impedance closure with
    `residues` (on R_GRID or `poles`) against the full porous solve in the judge box."""
    lee = _lee()
    out = []
    for kind in ("porous", "impedance"):
        d = judge_deck(g, kind, residues, poles=poles, hills=hills)
        s = lee.Solver(d, precision=precision, receivers=JUDGE_REC, use_graph=USE_GRAPH)
        try:
            s.frame_setup(0)
            out.append(s.step(d.time_domain))
        finally:
            s.close()
    return il_errors(out[0], out[1], d.dt)


class LearnedGround:
    """The trained network with its decision rule.  Targets are sqrt(a_k) (a = t^2 >= 0 keeps
    the boundary passive)."""

    def __init__(self, seed=1, device=0):
        self.net = Network(seed=seed, device=device)

    def train(self, X, A, epochs=600, batch=32, lr=2e-3, seed=7):
        """`epochs` at lr, then as many at lr/4 (Adam's moments carry over)."""
        t = np.sqrt(np.maximum(A, 0.0)).astype(np.float32)
        self.net.train(X, t, batch=batch, epochs=epochs, lr=lr, seed=seed)
        return self.net.train(X, t, batch=batch, epochs=epochs, lr=lr / 4, seed=seed + 1)

    def residues(self, grounds):
        t = self.net.predict(np.array([normalise(g) for g in grounds], np.float32))
        return t.astype(np.float64) ** 2

    def decide(self, g: Ground, columns: Columns, bound_r=HULL_FIT_R, bound_db=JUDGE_BOUND_DB):
        """(residues or None, report): the full porous solve judges the predicted residues --
        the column reflection error and the one-third-octave IL difference in the judge box;
        None (keep the full porous solve) outside the hull or above either bound."""
        rep = {"in_hull": in_hull(g)}
        if not rep["in_hull"]:
            return None, rep
        a = self.residues([g])[0]
        rep["reflection_error"] = columns.measure(g).error(a)
        if rep["reflection_error"] > bound_r:
            return None, rep
        rep["il_narrow_dB"], rep["il_third_octave_dB"] = judge_3d(g, a)
        return (a if rep["il_third_octave_dB"] <= bound_db else None), rep

    def close(self):
        self.net.close()
