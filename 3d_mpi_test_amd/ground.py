"""Learned ground: ground parameters -> pole/residue pairs of the ground impedance boundary.

The full solve models the ground as the porous layer of porous.cpp under the air (South
face porous, FB_p_v).  The learned ground replaces that layer by the South-face impedance of
lee_options.ground_* (K = 8 real poles, Z(s) = Z_inf (1 + sum_k a_k / (r_k + s tau)),
Z_inf = rho0 c q / porosity, tau = rho_e / sigma with rho_e = q^2 rho0 / porosity) whose
residues a_k on a fixed pole grid r_k come from a small network trained on the GPU
(lee_mlp, BF16 MFMA).  The training data are reflection coefficients measured in runs of the
full air + porous solve: a plane pulse in a rigid-walled column falls on the layer, the
reflected and incident pulses at a receiver give R(f) (divided by the same ratio over a rigid
ground, which removes the air path's numerical dispersion), and Z(f) = rho c (1 + R)/(1 - R)
is fitted with non-negative residues (passive, so the boundary recursion stays stable).

Outside the training hull (flow resistivity below HULL_SIGMA_MIN, or a fit worse than
HULL_FIT_R in reflection coefficient) the model declines and the caller keeps the full porous solve.
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
HULL_SIGMA_MIN = 3.0e4
HULL_FIT_R = 0.05
# C3 / C4 ground-parameter distribution
SIGMA_RANGE = (1.0e4, 2.0e6)
POROSITY_RANGE = (0.2, 0.5)
TORTUOSITY_RANGE = (1.5, 4.0)
DEPTH_POINTS = (4, 16)


def _lee():
    return importlib.import_module(__package__ or "3d_mpi_test_amd")


@dataclass
class Ground:
    sigma: float        # flow resistivity [Pa s / m^2]
    porosity: float
    tortuosity: float   # q^2
    depth: int          # porous layer points


def normalise(g: Ground):
    """Network inputs in [-1, 1]."""
    ls = (np.log10(g.sigma) - np.log10(SIGMA_RANGE[0])) / (np.log10(SIGMA_RANGE[1]) - np.log10(SIGMA_RANGE[0]))
    po = (g.porosity - POROSITY_RANGE[0]) / (POROSITY_RANGE[1] - POROSITY_RANGE[0])
    to = (g.tortuosity - TORTUOSITY_RANGE[0]) / (TORTUOSITY_RANGE[1] - TORTUOSITY_RANGE[0])
    de = (g.depth - DEPTH_POINTS[0]) / (DEPTH_POINTS[1] - DEPTH_POINTS[0])
    return np.array([2 * ls - 1, 2 * po - 1, 2 * to - 1, 2 * de - 1], dtype=np.float32)


def zinf_tau(g: Ground, rho0: float, c: float):
    """Z_inf and tau of the porous equations (porous.cpp:40-51 with airpore.cpp:249-251)."""
    rho_e = g.tortuosity * rho0 / g.porosity
    return rho0 * c * np.sqrt(g.tortuosity) / g.porosity, rho_e / g.sigma


def sample_grounds(n: int, seed: int):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        out.append(Ground(float(10 ** rng.uniform(np.log10(SIGMA_RANGE[0]), np.log10(SIGMA_RANGE[1]))),
                          float(rng.uniform(*POROSITY_RANGE)), float(rng.uniform(*TORTUOSITY_RANGE)),
                          int(rng.integers(DEPTH_POINTS[0], DEPTH_POINTS[1] + 1))))
    return out


# ---------------------------------------------------------------------- reflection data
def column_deck(g: Ground | None, dz=0.04, n_air=320, steps=2048):
    """A 6 x 6 column of n_air points with rigid side walls and a 20-point PML at the top,
    over the ground (the porous layer of g, or a rigid floor for g = None); FB_p_v at
    CFL 0.5 in 3-D."""
    lee = _lee()
    kw = dict(pml=20, bc=(2, 3 if g else 1, 1, 1, 1, 1), porous_cells=(g.depth + 2) if g else 0,
              time_domain=steps, src=(0.0, 0.0, -100.0))
    d = lee.deck.box(6, 6, n_air, dz, **kw)
    if g:
        d.pore_q, d.pore_porosity, d.pore_resistivity = float(np.sqrt(g.tortuosity)), g.porosity, g.sigma
    return d


def _plane_pulse(solver, d, z0, width):
    rho = d.gamma * d.Pav / d.c ** 2
    z = np.asarray(d.air_z)[:, None, None] * np.ones((len(d.air_z), len(d.air_y), len(d.air_x)))
    p = np.exp(-((z - z0) / width) ** 2)
    zero = np.zeros_like(p)
    solver.set_state(zero, zero, -p / (rho * d.c), p)


def column_series(g: Ground | None, dz=0.04, n_air=320, z_src=6.0, z_rec=3.0, steps=2048, precision=64):
    """Pressure at z_rec of a plane pulse from z_src falling on the ground."""
    lee = _lee()
    d = column_deck(g, dz, n_air, steps)
    rec = [(float(d.air_x[2]), float(d.air_y[2]), z_rec)]
    s = lee.Solver(d, precision=precision, receivers=rec, use_graph=True)
    s.frame_setup(0)
    _plane_pulse(s, d, z_src, 3 * dz)
    out = s.step(steps)[:, 0]
    s.close()
    return out, d.dt


def reflection(series, ref_series, dt, c, z_src, z_rec, nfft=4096):
    """R(f) of the ground from the receiver series over it and over a rigid floor: the
    incident and reflected windows are split half-way between their arrivals."""
    t_split = ((z_src - z_rec) / c + (z_src + z_rec) / c) / 2
    n_split = int(t_split / dt)
    def ratio(x):
        inc = np.zeros(nfft); ref = np.zeros(nfft)
        inc[:n_split] = x[:n_split]
        ref[:len(x) - n_split] = x[n_split:]
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.fft.rfft(ref) / np.fft.rfft(inc)
    f = np.fft.rfftfreq(nfft, dt)
    with np.errstate(divide="ignore", invalid="ignore"):   # beyond the pulse's band
        return f, ratio(series) / ratio(ref_series)


def fit_residues(f, Z, zinf, tau, z0, fmin=20.0, fmax=2000.0):
    """Non-negative residues on R_GRID for Z(f) / Z_inf - 1, each frequency weighted by the
    sensitivity of the reflection coefficient to Z (|dR/dZ| = 2 z0 / |Z + z0|^2, z0 = rho0 c);
    returns (a, max |R_fit - R| over 50-1000 Hz)."""
    sel = (f >= fmin) & (f <= fmax)
    st = 2j * np.pi * f[sel] * tau
    y = Z[sel] / zinf - 1.0
    A = 1.0 / (R_GRID[None, :] + st[:, None])
    w = 2.0 * z0 * zinf / np.abs(Z[sel] + z0) ** 2
    M = np.vstack([(A * w[:, None]).real, (A * w[:, None]).imag])
    b = np.concatenate([(y * w).real, (y * w).imag])
    a, _ = nnls(M, b)
    return a, reflection_error(f, Z, impedance(f, a, zinf, tau), z0)


def impedance(f, a, zinf, tau):
    """Z(f) of residues a on R_GRID."""
    return zinf * (1 + (1.0 / (R_GRID[None, :] + 2j * np.pi * np.asarray(f)[:, None] * tau)) @ np.asarray(a, float))


def reflection_error(f, Z, Zf, z0):
    """max |R_fit - R| of the plane-wave reflection coefficient over 50-1000 Hz."""
    band = (f >= 50) & (f <= 1000)
    R = (Z[band] - z0) / (Z[band] + z0)
    Rf = (Zf[band] - z0) / (Zf[band] + z0)
    return float(np.abs(Rf - R).max())


def training_set(grounds, dz=0.04, precision=64):
    """Reflection data of the full air + porous solve for each ground, fitted residues and
    fit errors max |R_fit - R|: (X [n x 4], A [n x 8], err [n])."""
    lee = _lee()
    d0 = column_deck(None, dz)
    rho = d0.gamma * d0.Pav / d0.c ** 2
    ref, dt = column_series(None, dz, precision=precision)
    X, A, E = [], [], []
    for g in grounds:
        s, _ = column_series(g, dz, precision=precision)
        f, R = reflection(s, ref, dt, d0.c, 6.0, 3.0)
        Z = rho * d0.c * (1 + R) / (1 - R)
        zinf, tau = zinf_tau(g, rho, d0.c)
        a, e = fit_residues(f, Z, zinf, tau, rho * d0.c)
        X.append(normalise(g)); A.append(a.astype(np.float32)); E.append(e)
    return np.array(X), np.array(A), np.array(E)


def in_hull(g: Ground):
    return g.sigma >= HULL_SIGMA_MIN


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


def ground_options(deck, g: Ground, residues):
    """Set the ground impedance fields of `deck` (a Deck) from residues on R_GRID."""
    rho = deck.gamma * deck.Pav / deck.c ** 2
    zinf, tau = zinf_tau(g, rho, deck.c)
    deck.ground_poles = [(float(max(a, 0.0)), float(r)) for a, r in zip(residues, R_GRID)]
    deck.ground_zinf, deck.ground_tau = float(zinf), float(tau)
    return deck
