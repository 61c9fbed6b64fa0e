"""Deck model of the drop-in file contract, and a generator of decks.

A deck is what the reference solver reads at start-up: the positional ``input.txt``
(parsed by ``airpore::airpore``, 3D_MPI_TEST/airpore.cpp:17-316) and the grid files
``coordi_air{1,2,3}.dat`` / ``coordi_pore{1,2,3}.dat`` (read by
``calculation_2D::calculation_2D``, 3D_MPI_TEST/calculation_2D.cpp:44-68, 154-204).

``Deck.write(dir)`` produces files in exactly that layout (same line structure as the
shipped ``3D_MPI_TEST/input.txt``, so the reference parser and ours read the same values),
and ``Deck.to_c()`` builds the ``lee_deck`` struct of ``include/lee_abi.h`` directly.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
from typing import Optional, Sequence

import numpy as np

c_double_p = ctypes.POINTER(ctypes.c_double)


class LeeDeck(ctypes.Structure):
    """ctypes mirror of ``lee_deck`` (include/lee_abi.h)."""

    _fields_ = [
        ("gauss_width", ctypes.c_double),
        ("mf_jmax", ctypes.c_int), ("mf_lead_dj", ctypes.c_int), ("mf_trail_dj", ctypes.c_int),
        ("mf_judge", ctypes.c_int),
        ("mf_ratio", ctypes.c_double),
        ("scheme", ctypes.c_int),
        ("bc", ctypes.c_int * 6),
        ("time_domain", ctypes.c_int), ("out_difft", ctypes.c_int), ("fft_m", ctypes.c_int),
        ("frequency", ctypes.c_double),
        ("velocity_coef", ctypes.c_double),
        ("velocity_method", ctypes.c_int),
        ("gamma", ctypes.c_double), ("Pav", ctypes.c_double), ("c", ctypes.c_double),
        ("pml_alpha", ctypes.c_double * 3), ("pml_R", ctypes.c_double * 3),
        ("pml_y_time", ctypes.c_int),
        ("pore_q", ctypes.c_double), ("pore_porosity", ctypes.c_double),
        ("pore_resistivity", ctypes.c_double),
        ("src", ctypes.c_double * 3), ("rcv", ctypes.c_double * 3),
        ("dt", ctypes.c_double),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("mpi_dims", ctypes.c_int * 3),
        ("restart", ctypes.c_int), ("restart_out", ctypes.c_int),
        ("hill_judge", ctypes.c_int),
        ("hill_coef", ctypes.c_double * 12),
        ("hill_points", ctypes.c_int),
        ("hill_mat", (ctypes.c_double * 3) * 3),
        ("hill_box", ctypes.c_double * 6),
        ("pore_dx", ctypes.c_double), ("pore_dy", ctypes.c_double), ("pore_dz", ctypes.c_double),
        ("air_hdr", ctypes.c_int * 10),
        ("air_x", c_double_p), ("air_y", c_double_p), ("air_z", c_double_p),
        ("has_pore", ctypes.c_int),
        ("pore_hdr", ctypes.c_int * 10),
        ("pore_x", c_double_p), ("pore_y", c_double_p), ("pore_z", c_double_p),
    ]


class LeeOptions(ctypes.Structure):
    """ctypes mirror of ``lee_options`` (include/lee_abi.h): the sidecar of a deck."""

    _fields_ = [
        ("precision", ctypes.c_int),
        ("n_receivers", ctypes.c_int),
        ("receivers", c_double_p),
        ("health_every", ctypes.c_int),
        ("use_graph", ctypes.c_int),
        ("device", ctypes.c_int),
        ("path", ctypes.c_int),
        ("integrator", ctypes.c_int),
        ("n_hills", ctypes.c_int),
        ("hills", ctypes.c_double * 32),
        ("wind_ustar", ctypes.c_double),
        ("wind_z0", ctypes.c_double),
        ("ground_poles", ctypes.c_int),
        ("ground_a", ctypes.c_double * 8),
        ("ground_r", ctypes.c_double * 8),
        ("ground_zinf", ctypes.c_double),
        ("ground_tau", ctypes.c_double),
        ("rk_tile", ctypes.c_int),
        ("fb_form", ctypes.c_int),
    ]


INT_DECK, INT_RK4 = 0, 1


def make_options(deck: Optional["Deck"] = None, precision: int = 64, receivers=None, health_every: int = 0,
                 device: int = 0, path: int = 0, integrator: Optional[int] = None, use_graph: int = 0,
                 rk_tile: int = 0, fb_form: int = 0) -> LeeOptions:
    """``lee_options`` for a deck: its extensions (integrator, hills, log wind) plus the run
    settings.  The receiver array is kept alive on the returned struct."""
    o = LeeOptions()
    o.precision, o.health_every, o.device, o.path, o.use_graph = precision, health_every, device, path, use_graph
    o.rk_tile, o.fb_form = rk_tile, fb_form
    if receivers is not None:
        rec = np.ascontiguousarray(receivers, dtype=np.float64).reshape(-1)
        o._rec = rec
        o.n_receivers = rec.size // 3
        o.receivers = rec.ctypes.data_as(c_double_p)
    if deck is not None:
        o.integrator = deck.integrator if integrator is None else integrator
        hills = list(deck.hills)
        if len(hills) > 8:
            raise ValueError("at most 8 hills")
        o.n_hills = len(hills)
        for n, h in enumerate(hills):
            for q in range(4):
                o.hills[4 * n + q] = float(h[q])
        o.wind_ustar, o.wind_z0 = float(deck.wind_ustar), float(deck.wind_z0)
        poles = list(deck.ground_poles)
        if len(poles) > 8:
            raise ValueError("at most 8 ground poles")
        o.ground_poles = len(poles)
        for n, (a, r) in enumerate(poles):
            o.ground_a[n], o.ground_r[n] = float(a), float(r)
        o.ground_zinf, o.ground_tau = float(deck.ground_zinf), float(deck.ground_tau)
    elif integrator is not None:
        o.integrator = integrator
    return o


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return c_double_p()
    return a.ctypes.data_as(c_double_p)


def _fmt(v) -> str:
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return repr(float(v))


@dataclasses.dataclass
class Deck:
    """All values of input.txt plus the grid arrays (air, and porous when South = 3)."""

    air_x: np.ndarray
    air_y: np.ndarray
    air_z: np.ndarray
    # PML point counts of the coordi_air1.dat header: IMAX1 IMAX2 JMAX1 JMAX2 KMAX1 KMAX2
    pml_cells: Sequence[int] = (0, 0, 0, 0, 0, 0)
    ijkmax: Optional[int] = None
    gauss_width: float = 0.3
    mf_jmax: Optional[int] = None
    mf_lead_dj: int = 11
    mf_trail_dj: int = 11
    mf_judge: int = 0
    mf_ratio: float = 0.75
    scheme: int = 2
    bc: Sequence[int] = (1, 1, 1, 1, 1, 1)  # N S E W F B
    time_domain: int = 100
    out_difft: int = 100
    fft_m: int = 12
    frequency: float = 300.0
    velocity_coef: float = 0.0
    velocity_method: int = 4
    gamma: float = 1.4
    Pav: float = 1.0e5
    c: float = 340.0
    pml_alpha: Sequence[float] = (2.0, 2.0, 2.0)
    pml_R: Sequence[float] = (1000.0, 1000.0, 1000.0)
    pml_y_time: int = 4500
    pore_q: float = 1.7320508075688772
    pore_porosity: float = 0.3
    pore_resistivity: float = 2.0e5
    src: Sequence[float] = (0.0, 0.0, 0.0)
    rcv: Sequence[float] = (0.0, -0.23, 1.75)
    dt: float = 4e-6
    dx: float = 0.005
    dy: float = 0.005
    dz: float = 0.005
    mpi_dims: Sequence[int] = (1, 1, 1)
    restart: int = 0
    restart_out: int = 5000
    restart_infile: str = "ct003"
    hill_judge: int = 0
    # embedded porous obstacle: curve coefficients a..l, curve points, hill / canopy / trunk
    # materials (q, porosity, resistivity) and bounding box (the shipped deck's values)
    hill_coef: Sequence[float] = (-1.0, 1.0, 0.0, 2.5, -0.2, 0.0, 3.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    hill_points: int = 0
    hill_mat: Sequence[Sequence[float]] = ((1.790799820223793, 0.8769089596793096, 5.137565e4),
                                           (1.790799820223793, 0.8769089596793096, 5.137565e4),
                                           (0.52, 0.5712, 8.0e4))
    hill_box: Sequence[float] = (-1.0, 1.0, 0.0, 2.5, -0.2, 0.0)
    pore_dx: float = 0.005
    pore_dy: float = 0.005
    pore_dz: float = 0.005
    pore_x: Optional[np.ndarray] = None
    pore_y: Optional[np.ndarray] = None
    pore_z: Optional[np.ndarray] = None
    # sidecar extensions (lee_options; not part of input.txt): time integrator, Gaussian
    # hills (x0, y0, height, sigma) of a terrain-following grid, log wind of velocity_method 7,
    # ground impedance (pole/residue pairs (a_k, r_k), Z_inf, tau) on the South face
    integrator: int = 0
    hills: Sequence[Sequence[float]] = ()
    wind_ustar: float = 0.0
    wind_z0: float = 0.0
    ground_poles: Sequence[Sequence[float]] = ()
    ground_zinf: float = 0.0
    ground_tau: float = 0.0

    def __post_init__(self):
        self.air_x = np.ascontiguousarray(self.air_x, dtype=np.float64)
        self.air_y = np.ascontiguousarray(self.air_y, dtype=np.float64)
        self.air_z = np.ascontiguousarray(self.air_z, dtype=np.float64)
        for name in ("pore_x", "pore_y", "pore_z"):
            v = getattr(self, name)
            if v is not None:
                setattr(self, name, np.ascontiguousarray(v, dtype=np.float64))
        if self.mf_jmax is None:
            self.mf_jmax = len(self.air_y)
        if self.ijkmax is None:
            self.ijkmax = len(self.air_y)
        self._keep = []

    # ------------------------------------------------------------------ shape helpers
    @property
    def shape(self):
        return (len(self.air_x), len(self.air_y), len(self.air_z))

    @property
    def has_pore(self) -> bool:
        return self.pore_x is not None

    def air_hdr(self):
        i1, i2, j1, j2, k1, k2 = self.pml_cells
        return [len(self.air_x), i1, i2, len(self.air_y), j1, j2, len(self.air_z), k1, k2, int(self.ijkmax)]

    def pore_hdr(self):
        if not self.has_pore:
            return [0] * 10
        return [len(self.pore_x), 0, 0, len(self.pore_y), 0, 0, len(self.pore_z), 0, 0, len(self.pore_y)]

    # ------------------------------------------------------------------ C struct
    def to_c(self) -> LeeDeck:
        d = LeeDeck()
        d.gauss_width = self.gauss_width
        d.mf_jmax, d.mf_lead_dj, d.mf_trail_dj = int(self.mf_jmax), self.mf_lead_dj, self.mf_trail_dj
        d.mf_judge, d.mf_ratio = self.mf_judge, self.mf_ratio
        d.scheme = self.scheme
        for q in range(6):
            d.bc[q] = int(self.bc[q])
        d.time_domain, d.out_difft, d.fft_m, d.frequency = self.time_domain, self.out_difft, self.fft_m, self.frequency
        d.velocity_coef, d.velocity_method = self.velocity_coef, self.velocity_method
        d.gamma, d.Pav, d.c = self.gamma, self.Pav, self.c
        for q in range(3):
            d.pml_alpha[q] = self.pml_alpha[q]
            d.pml_R[q] = self.pml_R[q]
            d.src[q] = self.src[q]
            d.rcv[q] = self.rcv[q]
            d.mpi_dims[q] = self.mpi_dims[q]
        d.pml_y_time = self.pml_y_time
        d.pore_q, d.pore_porosity, d.pore_resistivity = self.pore_q, self.pore_porosity, self.pore_resistivity
        d.dt, d.dx, d.dy, d.dz = self.dt, self.dx, self.dy, self.dz
        d.restart, d.restart_out, d.hill_judge = self.restart, self.restart_out, self.hill_judge
        for q in range(12):
            d.hill_coef[q] = float(self.hill_coef[q])
        d.hill_points = int(self.hill_points)
        for a in range(3):
            for b in range(3):
                d.hill_mat[a][b] = float(self.hill_mat[a][b])
        for q in range(6):
            d.hill_box[q] = float(self.hill_box[q])
        d.pore_dx, d.pore_dy, d.pore_dz = self.pore_dx, self.pore_dy, self.pore_dz
        for q, v in enumerate(self.air_hdr()):
            d.air_hdr[q] = v
        # the reader takes JMAX values from coordi_air2.dat where JMAX is the frame length
        ny = int(self.mf_jmax) if self.mf_judge == 1 else len(self.air_y)
        self._ay = np.ascontiguousarray(self.air_y[:ny])
        d.air_hdr[3] = len(self.air_y)
        d.air_x, d.air_y, d.air_z = _ptr(self.air_x), _ptr(self._ay), _ptr(self.air_z)
        d.has_pore = 1 if self.has_pore else 0
        if self.has_pore:
            for q, v in enumerate(self.pore_hdr()):
                d.pore_hdr[q] = v
            self._py = np.ascontiguousarray(self.pore_y[:ny])
            d.pore_x, d.pore_y, d.pore_z = _ptr(self.pore_x), _ptr(self._py), _ptr(self.pore_z)
        return d

    # ------------------------------------------------------------------ files
    def write(self, dirpath: str) -> None:
        """Write input.txt and coordi_*.dat with the line structure of the shipped deck."""
        os.makedirs(dirpath, exist_ok=True)
        f = _fmt
        lines = [
            f"\t{f(self.gauss_width)}//gaussian_width",
            f"\t{self.mf_jmax} {self.mf_lead_dj} {self.mf_trail_dj} {self.mf_judge} {f(self.mf_ratio)} //Moving domain:MF_JMAX; lead; trail; judge; ratio",
            f"\t{self.scheme} //scheme:1.FB_v_p 2.FB_p_v 5.RK4",
            "\t" + " ".join(str(int(b)) for b in self.bc) + " //NorthBC SouthBC EastBC WestBC FrontBC BackBC and",
            "\t\t\t//B.C type: 1.rigid 2.absorbing layer 3.porous media 4.radiation 5. TDBC",
            f"\t{self.time_domain} {self.out_difft} {self.fft_m} {f(self.frequency)} //time_domain;out_difft;FFT_m;frequency;",
            "",
            f"\t{f(self.velocity_coef)} {self.velocity_method} //b; velocity method",
            f"\t{f(self.gamma)} {f(self.Pav)} {f(self.c)} //adiabatic_coef;Pav;sound_speed;",
            f"\t{f(self.pml_alpha[0])} {f(self.pml_R[0])}//PMLx: alpha,resistivity;",
            f"\t{f(self.pml_alpha[1])} {f(self.pml_R[1])} {self.pml_y_time} //PMLy: alpha,resistivity and time_step;",
            f"\t{f(self.pml_alpha[2])} {f(self.pml_R[2])} //PMLz: alpha,resistivity;",
            f"\t{f(self.pore_q)} {f(self.pore_porosity)} {f(self.pore_resistivity)} //PorePara.q;porosity;resistivity;",
            "\t" + " ".join(f(v) for v in (*self.src, *self.rcv)) + " //source.xyz;receiver.xyz;",
            "",
            f"\t{f(self.dt)} //AirStep.diff_t;",
            f"\t{f(self.dx)} {f(self.dy)} {f(self.dz)} //AirStep.diff_x;diff_y;diff_z;",
            "",
            "\t" + " ".join(str(int(v)) for v in self.mpi_dims) + " //mpi_xarea,mpi_yarea,mpi_zarea",
            "",
            "\toutput.data;",
            f"\t{self.restart} {self.restart_out} {self.restart_infile}",
            "",
            f"\t{self.hill_judge} " + " ".join(f(x) for x in self.hill_coef) + f" {int(self.hill_points)} //Curve",
            "\t" + " ".join(f(x) for x in self.hill_mat[0]) + " //hill.q;porosity;resistivity;",
            "\t" + " ".join(f(x) for x in self.hill_mat[1]) + " //canopy.q;porosity;resistivity;",
            "\t" + " ".join(f(x) for x in self.hill_mat[2]) + " //trunk.q;porosity;resistivity;",
            "",
            "\t" + " ".join(f(x) for x in self.hill_box) + " //xstar,xstop,ystar,ystop,zstar and zstop;",
            "",
            "\tporous_media;",
            f"\t{f(self.pore_dx)} {f(self.pore_dy)} {f(self.pore_dz)} //SouthStep.diff_x;diff_y;diff_z;",
        ]
        with open(os.path.join(dirpath, "input.txt"), "w") as fh:
            fh.write("\n".join(lines) + "\n")
        _write_grid(dirpath, "coordi_air", self.air_hdr(), self.air_x, self.air_y, self.air_z)
        if self.has_pore:
            _write_grid(dirpath, "coordi_pore", self.pore_hdr(), self.pore_x, self.pore_y, self.pore_z)


def _write_grid(dirpath, stem, hdr, x, y, z):
    with open(os.path.join(dirpath, stem + "1.dat"), "w") as fh:
        fh.write(" ".join(str(int(v)) for v in hdr) + "\n")
        fh.write(" ".join(repr(float(v)) for v in x) + "\n")
    with open(os.path.join(dirpath, stem + "2.dat"), "w") as fh:
        fh.write(" ".join(repr(float(v)) for v in y) + "\n")
    with open(os.path.join(dirpath, stem + "3.dat"), "w") as fh:
        fh.write(" ".join(repr(float(v)) for v in z) + "\n")


# ---------------------------------------------------------------------- generators

def uniform_axis(n: int, d: float, start: float = 0.0) -> np.ndarray:
    """n coordinates start + i*d, computed per index (no accumulated rounding)."""
    return start + d * np.arange(n, dtype=np.float64)


def stretched_axis(n: int, d: float, ratio: float, start: float = 0.0) -> np.ndarray:
    """n coordinates whose spacing starts at d and grows geometrically by `ratio` per cell:
    start + d (ratio^i - 1)/(ratio - 1) (uniform for ratio 1)."""
    if ratio == 1.0:
        return uniform_axis(n, d, start)
    i = np.arange(n, dtype=np.float64)
    return start + d * (ratio ** i - 1.0) / (ratio - 1.0)


def box(nx: int, ny: int, nz: int, d: float = 0.05, *, pml: int = 0, bc=None, dt=None,
        cfl: float = 0.5, c: float = 343.0, gamma: float = 1.4, rho0: float = 1.21,
        porous_cells: int = 0, pore_dz: Optional[float] = None, x0: Optional[float] = None,
        y0: float = 0.0, z0: float = 0.0, spacing: Optional[Sequence[float]] = None,
        stretch: Optional[Sequence[float]] = None, **kw) -> Deck:
    """A rectangular air box of nx*ny*nz points, optional PML on the sides and top (`pml`
    points each) and an optional porous ground layer of `porous_cells` points.

    The spacing is d on every axis, or spacing = (dx, dy, dz) per axis; stretch = (rx, ry, rz)
    grows each axis's spacing geometrically from its first cell (the deck's diff_x ... stay the
    first spacing, as the reference takes the computational step from input.txt and the
    metric from the coordinate files).  P0 = rho0*c^2/gamma (so that c^2/(gamma*P0) = 1/rho0),
    dt = cfl*min(spacing)/c unless given.
    """
    dx, dy, dz = (d, d, d) if spacing is None else (float(spacing[0]), float(spacing[1]), float(spacing[2]))
    rx, ry, rz = (1.0, 1.0, 1.0) if stretch is None else (float(stretch[0]), float(stretch[1]), float(stretch[2]))
    x0 = -0.5 * (nx - 1) * dx if x0 is None else x0
    ax = stretched_axis(nx, dx, rx, x0)
    ay = stretched_axis(ny, dy, ry, y0)
    az = stretched_axis(nz, dz, rz, z0)
    Pav = rho0 * c * c / gamma
    if dt is None:
        dt = cfl * min(dx, dy, dz) / c
    if bc is None:
        # N S E W F B: absorbing top and sides, ground rigid (or porous)
        bc = (2 if pml else 1, 3 if porous_cells else 1, 2 if pml else 1, 2 if pml else 1,
              2 if pml else 1, 2 if pml else 1)
    pml_cells = (pml if bc[5] == 2 else 0, pml if bc[4] == 2 else 0,
                 pml if bc[3] == 2 else 0, pml if bc[2] == 2 else 0,
                 pml if bc[1] == 2 else 0, pml if bc[0] == 2 else 0)
    kwargs = dict(air_x=ax, air_y=ay, air_z=az, pml_cells=pml_cells, bc=tuple(bc), dt=dt,
                  dx=dx, dy=dy, dz=dz, gamma=gamma, Pav=Pav, c=c)
    if porous_cells:
        pdz = dz if pore_dz is None else pore_dz
        # porous column: Kp points ending one plane above the air ground (two planes overlap
        # through the ghost copy of calculation_2D.cpp:878-886)
        pz = z0 + pdz * (np.arange(porous_cells, dtype=np.float64) - (porous_cells - 2))
        kwargs.update(pore_x=ax.copy(), pore_y=ay.copy(), pore_z=pz, pore_dx=dx, pore_dy=dy, pore_dz=pdz)
    kwargs.update(kw)
    return Deck(**kwargs)


# ---------------------------------------------------------------------- benchmark workloads

def config_c0(time_domain: int = 2000, reference_scheme: bool = False) -> Deck:
    """96x64x48 air points at 0.05 m over an 8-point Zwikker-Kosten ground (sigma 200
    kPa s/m^2, porosity 0.3, tortuosity 3), c0 = 343 m/s, rho0 = 1.21 kg/m^3, no wind,
    Gaussian pulse of half-width 3 dx at 1.0 m, 20-point PML on the sides and top, CFL 0.5,
    RK4, five receivers at 1.5 m height 1-4 m from the source (the top layer starts at 1.4 m,
    so the receivers sit in it: run as specified).  reference_scheme=True keeps the deck's
    FB_p_v step, the only integrator of the reference."""
    d = 0.05
    dk = box(96, 64, 48, d, pml=20, bc=(2, 3, 2, 2, 2, 2), porous_cells=8 + 2, y0=-0.75, time_domain=time_domain,
             gauss_width=2 * 3 * d, src=(0.0, 0.0, 1.0), rcv=(0.0, 1.0, 1.5), fft_m=12, pml_y_time=0)
    dk.pore_q, dk.pore_porosity, dk.pore_resistivity = 3.0 ** 0.5, 0.3, 2.0e5
    if not reference_scheme:
        dk.integrator = INT_RK4
    return dk


def receivers_c0():
    return [(0.0, float(r), 1.5) for r in (1.0, 1.75, 2.5, 3.25, 4.0)]


def config_c1(time_domain: int = 10000, reference_scheme: bool = False) -> Deck:
    """512x256x128 air points at 0.04 m (x lateral, y propagation, z up) over a 16-point
    porous layer (200 kPa s/m^2, porosity 0.3, tortuosity 3), 20-point PML on the sides and
    top, moving frame on (frame = the 256 y-planes, shifted by 32 planes; the first frame runs
    half a frame transit), Gaussian pulse of half-width 3 dx at 0.5 m, CFL 0.5 with c0 = 343 m/s, RK4,
    terrain-following over a Gaussian hill of height 2 m and sigma 5 m (centred on the line
    of receivers at y = 30 m), logarithmic wind u(z) = (0.4/0.41) ln(z/0.02) along +y.
    Heights of the source and receivers are above the local ground.  reference_scheme=True
    gives the same grid with the reference's capabilities only: FB_p_v, flat, no wind."""
    d = 0.04
    lead = 33
    ijk = 256 + 32 * 160
    dk = box(512, 256, 128, d, pml=20, bc=(2, 3, 2, 2, 2, 2), porous_cells=16 + 2, time_domain=time_domain,
             gauss_width=2 * 3 * d, src=(0.0, 2.0, 0.5), rcv=(0.0, 6.0, 1.2), fft_m=14, pml_y_time=0,
             mf_judge=1, mf_jmax=256, mf_lead_dj=lead, mf_trail_dj=lead, mf_ratio=0.5, ijkmax=ijk)
    dk.pore_q, dk.pore_porosity, dk.pore_resistivity = 3.0 ** 0.5, 0.3, 2.0e5
    if not reference_scheme:
        dk.integrator = INT_RK4
        dk.hills = [(0.0, 30.0, 2.0, 5.0)]
        dk.velocity_method, dk.wind_ustar, dk.wind_z0 = 7, 0.4, 0.02
    return dk


def config_c2(time_domain: int = 2000, seed: int = 20261202) -> Deck:
    """Precision study: 768x384x192 air points at 0.04 m (the spacing is not specified; the
    512x256x128 benchmark's is used) over a 24-point porous layer (200 kPa s/m^2, porosity 0.3,
    tortuosity 3), terrain-following over a random
    smooth terrain -- a sum of 6 Gaussians with heights U[0, 1.5] m and widths U[3, 10] m at
    centres uniform over the footprint, drawn from numpy's PCG64 with the given seed --
    20-point PML on the sides and top, no wind, no moving frame, Gaussian pulse of half-width
    3 dx at 0.5 m, RK4 at CFL 0.5, 2000 steps (the step count is not specified)."""
    d = 0.04
    dk = box(768, 384, 192, d, pml=20, bc=(2, 3, 2, 2, 2, 2), porous_cells=24 + 2, time_domain=time_domain,
             gauss_width=2 * 3 * d, src=(0.0, 3.0, 0.5), rcv=(0.0, 6.0, 1.2), fft_m=14, pml_y_time=0)
    dk.pore_q, dk.pore_porosity, dk.pore_resistivity = 3.0 ** 0.5, 0.3, 2.0e5
    rng = np.random.default_rng(seed)
    x0, x1, y0, y1 = dk.air_x[0], dk.air_x[-1], dk.air_y[0], dk.air_y[-1]
    hills = []
    for _ in range(6):
        hx, hy = rng.uniform(x0, x1), rng.uniform(y0, y1)
        hills.append((float(hx), float(hy), float(rng.uniform(0.0, 1.5)), float(rng.uniform(3.0, 10.0))))
    dk.integrator, dk.hills = INT_RK4, hills
    return dk


def config_c3(time_domain: int = 3000, seed: int = 20261202, n: int = 64):
    """Batched ensemble: n domains with sides drawn from {64..128} x {48..96} x {48..64} at
    0.05 m, each over its own porous layer -- flow resistivity log-uniform in 10-2000
    kPa s/m^2, porosity U[0.2, 0.5], tortuosity U[1.5, 4] (q = sqrt(tortuosity)), depth 4-16
    points -- a source at a height U[0.3, 2] m, a logarithmic wind with u* U[0, 0.5] m/s
    (z0 = 0.02 m), RK4 at CFL 0.5 for 3000 steps.  Not specified and chosen here: 10-point
    PML on the sides and top, the source at a quarter of the domain length on the centre
    line, four receivers at 1.2 m along the centre line, numpy's PCG64 with the given seed.
    Returns (decks, receivers)."""
    rng = np.random.default_rng(seed)
    d = 0.05
    decks, recs = [], []
    for _ in range(n):
        nx, ny, nz = (int(rng.integers(64, 129)), int(rng.integers(48, 97)), int(rng.integers(48, 65)))
        depth = int(rng.integers(4, 17))
        sigma = float(10.0 ** rng.uniform(np.log10(1.0e4), np.log10(2.0e6)))
        poro = float(rng.uniform(0.2, 0.5))
        tort = float(rng.uniform(1.5, 4.0))
        hs = float(rng.uniform(0.3, 2.0))
        ustar = float(rng.uniform(0.0, 0.5))
        ylen = (ny - 1) * d
        dk = box(nx, ny, nz, d, pml=10, bc=(2, 3, 2, 2, 2, 2), porous_cells=depth + 2, time_domain=time_domain,
                 gauss_width=2 * 3 * d, src=(0.0, 0.25 * ylen, hs), rcv=(0.0, 0.5 * ylen, 1.2), fft_m=12,
                 velocity_method=7)
        dk.pore_q, dk.pore_porosity, dk.pore_resistivity = tort ** 0.5, poro, sigma
        dk.integrator, dk.wind_ustar, dk.wind_z0 = INT_RK4, ustar, 0.02
        decks.append(dk)
        recs.append([(0.0, f * ylen, 1.2) for f in (0.5, 0.625, 0.75, 0.875)])
    return decks, recs


C4_HILL = (0.0, 12.0, 1.0, 5.0)  # x0, y0, height, sigma [m]: a hill between the C4 receivers


def config_c4(ground, kind: str, residues=None, time_domain: int = 8000, flat: bool = False) -> Deck:
    """Learned ground: 1024x512x160 air points at 0.04 m (the spacing is not specified; the
    512x256x128 benchmark's is used), curvilinear: terrain-following over a Gaussian hill of
    height 1 m and sigma 5 m at y = 12 m on the receiver line (the height is not specified;
    2 m, as in the 512x256x128 workload, breaks the FB_p_v step bound at CFL 0.5 on this
    160-point column), 20-point PML on the sides and top, no wind, no moving frame, Gaussian
    pulse of half-width 3 dx at 0.5 m above the ground, 3 m into the domain, FB_p_v (the
    ground impedance is discretised for it) at CFL 0.5, 8000 steps.  kind 'porous': the full
    solve over the porous layer of `ground` (a ground.Ground, depth points + the two interface
    planes, following the terrain); kind 'impedance': air only, the ground interface plane on
    the terrain closed by the impedance with `residues` on ground.R_GRID.  flat=True drops the
    hill."""
    from . import ground as G
    d = 0.04
    kw = dict(pml=20, time_domain=time_domain, gauss_width=2 * 3 * d, src=(0.0, 3.0, 0.5), rcv=(0.0, 6.0, 1.2),
              fft_m=14, pml_y_time=0)
    if kind == "porous":
        dk = box(1024, 512, 160, d, bc=(2, 3, 2, 2, 2, 2), porous_cells=ground.depth + 2, **kw)
        dk = G.porous_deck(dk, ground)
    else:
        dk = box(1024, 512, 160, d, bc=(2, 1, 2, 2, 2, 2), **kw)
        dk = G.impedance_deck(dk, ground, residues)
    if not flat:
        dk.hills = [C4_HILL]
    return dk


def receivers_c4():
    """Sixteen receivers 1-15 m from the source at 0.3 m and 1.2 m height (not specified)."""
    return [(0.0, 3.0 + r, h) for r in (1.0, 2.0, 3.0, 4.5, 6.0, 8.0, 11.0, 15.0) for h in (0.3, 1.2)]


def receivers_c2():
    """Thirty-two receivers on a line at 1.2 m along the propagation axis (not specified)."""
    return [(0.0, 4.0 + 0.35 * r, 1.2) for r in range(32)]


def receivers_c1():
    """Thirty-two receivers on a line at 1.2 m height along the propagation axis."""
    return [(0.0, 2.0 + 4.0 + 6.0 * r, 1.2) for r in range(32)]
