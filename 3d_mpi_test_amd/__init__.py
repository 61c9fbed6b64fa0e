"""MI355X-native time-stepping hot path of the 3D_MPI_TEST outdoor-sound FDTD solver.

The compute path is ``liblee_hip.so`` (hand-written HIP kernels for gfx950 behind the C ABI
of ``include/lee_abi.h``).  This module is a thin ctypes front end over that ABI:

* :class:`Solver` -- one domain on one GPU (``lee_create`` ... ``lee_destroy``);
* :class:`AirPore` -- mirror of the reference driver ``airpore`` (airpore.h:18-67):
  ``AirPore(input_dir).get_output()`` runs the whole time loop and writes
  ``solution/pty%04d.dat``; ``get_FFT_y1()`` writes ``solution/p_f2_%04d.dat``.

There is no CPU fallback: if the library is missing, or no GPU is visible, the calls fail.
"""
from __future__ import annotations

import ctypes
import importlib
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblee_hip.so")

deck = importlib.import_module(__name__ + ".deck")
LeeDeck = deck.LeeDeck
ground = importlib.import_module(__name__ + ".ground")

LEE_OK = 0
PATH_AUTO, PATH_SPLIT, PATH_FUSED = 0, 1, 2
ERRORS = {-1: "LEE_ERR_ARG", -2: "LEE_ERR_DEVICE", -3: "LEE_ERR_IO", -4: "LEE_ERR_UNSUPPORTED",
          -5: "LEE_ERR_NONFINITE", -6: "LEE_ERR_STATE"}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


LeeOptions = deck.LeeOptions
INT_DECK, INT_RK4 = deck.INT_DECK, deck.INT_RK4


class LeeInfo(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("pnx", ctypes.c_int), ("pny", ctypes.c_int), ("pnz", ctypes.c_int),
        ("ijkmax", ctypes.c_int), ("n_receivers", ctypes.c_int), ("precision", ctypes.c_int),
        ("mf_count", ctypes.c_int),
        ("steps_done", ctypes.c_longlong), ("pml_y1_counter", ctypes.c_longlong),
        ("path", ctypes.c_int),
        ("rk_tile", ctypes.c_int),
        ("fb_form", ctypes.c_int),
    ]


class LeeKstat(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 32),
        ("launches", ctypes.c_longlong),
        ("total_ms", ctypes.c_double),
        ("bytes_per_launch", ctypes.c_double),
        ("points_per_launch", ctypes.c_double),
    ]


class LeeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load the HIP library (fails loudly when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.lee_deck_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(LeeDeck)]
        L.lee_deck_free.argtypes = [ctypes.POINTER(LeeDeck)]
        L.lee_deck_free.restype = None
        L.lee_options_default.argtypes = [ctypes.POINTER(LeeOptions)]
        L.lee_options_default.restype = None
        L.lee_create.argtypes = [ctypes.POINTER(LeeDeck), ctypes.POINTER(LeeOptions), ctypes.POINTER(vp)]
        L.lee_destroy.argtypes = [vp]
        L.lee_last_error.argtypes = [vp]
        L.lee_last_error.restype = ctypes.c_char_p
        L.lee_global_error.restype = ctypes.c_char_p
        L.lee_get_info.argtypes = [vp, ctypes.POINTER(LeeInfo)]
        L.lee_receiver_indices.argtypes = [vp, _ip, ctypes.c_int]
        L.lee_frame_setup.argtypes = [vp, ctypes.c_int]
        L.lee_step.argtypes = [vp, ctypes.c_int, _dp]
        L.lee_frame_steps.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.lee_frame_count.argtypes = [vp]
        L.lee_run.argtypes = [vp, _dp, ctypes.c_longlong]
        L.lee_get_state.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.lee_set_state.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.lee_get_psi.argtypes = [vp, _dp, ctypes.c_longlong]
        L.lee_last_step_ms.argtypes = [vp]
        L.lee_last_step_ms.restype = ctypes.c_double
        L.lee_write_receivers.argtypes = [ctypes.c_char_p, _dp, ctypes.c_int, ctypes.c_int]
        L.lee_write_spectra.argtypes = [ctypes.c_char_p, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.lee_fft.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.lee_profile.argtypes = [vp, ctypes.c_int]
        L.lee_kernel_stats.argtypes = [vp, ctypes.POINTER(LeeKstat), ctypes.c_int]
        L.lee_copy_bench.argtypes = [vp, ctypes.c_longlong, ctypes.c_int]
        L.lee_stream_bench.argtypes = [vp, ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.lee_stream_bench.restype = ctypes.c_double
        L.lee_save_restart.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
        L.lee_load_restart.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
        L.lee_write_contour.argtypes = [vp, ctypes.c_char_p]
        L.lee_write_contour_noalloc.argtypes = [vp, ctypes.c_char_p]
        L.lee_set_source_series.argtypes = [vp, _dp, ctypes.c_longlong, ctypes.c_int]
        L.lee_get_ground_psi.argtypes = [vp, _dp, ctypes.c_longlong]
        L.lee_ensemble_create.argtypes = [ctypes.c_int, ctypes.POINTER(LeeDeck), ctypes.POINTER(LeeOptions),
                                          ctypes.POINTER(ctypes.c_void_p)]
        L.lee_ensemble_destroy.argtypes = [vp]
        L.lee_ensemble_size.argtypes = [vp]
        L.lee_ensemble_member.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.lee_ensemble_receivers.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.lee_ensemble_step.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_int]
        L.lee_ensemble_last_ms.argtypes = [vp]
        L.lee_ensemble_last_ms.restype = ctypes.c_double
        L.lee_ensemble_error.argtypes = [vp]
        L.lee_ensemble_error.restype = ctypes.c_char_p
        L.lee_ensemble_set_launch.argtypes = [vp, ctypes.c_int]
        L.lee_ensemble_launch.argtypes = [vp, ctypes.c_int]
        L.lee_read_source_series.argtypes = [ctypes.c_char_p, _dp, ctypes.c_longlong]
        L.lee_read_source_series.restype = ctypes.c_longlong
        L.lee_copy_bench.restype = ctypes.c_double
        _lib = L
    return _lib


class Ensemble:
    """Many independent domains stepped together (lee_ensemble_*): each RK4 stage is one
    launch per kernel kind over every member.  `decks` and per-member `receivers` lists;
    members share precision, wind kind and grid kind."""

    def __init__(self, decks, precision: int = 32, receivers=None, use_graph: bool = True, launch: int = 0):
        L = lib()
        n = len(decks)
        self._cdecks = (LeeDeck * n)()
        self._opts = (LeeOptions * n)()
        keep = []
        for d, dk in enumerate(decks):
            c = dk.to_c() if isinstance(dk, deck.Deck) else dk
            keep.append(c)
            ctypes.memmove(ctypes.byref(self._cdecks, d * ctypes.sizeof(LeeDeck)), ctypes.byref(c), ctypes.sizeof(LeeDeck))
            o = deck.make_options(dk if isinstance(dk, deck.Deck) else None, precision=precision,
                                  receivers=None if receivers is None else receivers[d], integrator=deck.INT_RK4)
            keep.append(o)
            ctypes.memmove(ctypes.byref(self._opts, d * ctypes.sizeof(LeeOptions)), ctypes.byref(o), ctypes.sizeof(LeeOptions))
        self._keep = keep
        h = ctypes.c_void_p()
        rc = L.lee_ensemble_create(n, self._cdecks, self._opts, ctypes.byref(h))
        if rc != LEE_OK:
            raise LeeError(rc, (L.lee_global_error() or b"").decode(errors="replace"))
        self.h = h
        self.n = n
        self.use_graph = use_graph
        off = (ctypes.c_int * (n + 1))()
        L.lee_ensemble_receivers(self.h, off, n + 1)
        self.offsets = list(off)
        self.nrec = self.offsets[-1]
        self.members = []
        for d in range(n):
            m = ctypes.c_void_p()
            self._check(L.lee_ensemble_member(self.h, d, ctypes.byref(m)))
            self.members.append(Solver._member(m, keep[2 * d]))
        if launch:
            self.set_launch(launch)

    def _check(self, rc):
        if rc != LEE_OK:
            raise LeeError(rc, (lib().lee_ensemble_error(self.h) or b"").decode(errors="replace"))

    def step(self, n: int, record: bool = True):
        """n steps of every member; returns a list of (n, nrec_d) receiver series."""
        rec = np.zeros((n, max(self.nrec, 1))) if record else None
        self._check(lib().lee_ensemble_step(self.h, n, _p(rec), 1 if self.use_graph else 0))
        if rec is None:
            return None
        return [rec[:, self.offsets[d]:self.offsets[d + 1]].copy() for d in range(self.n)]

    def last_ms(self) -> float:
        return lib().lee_ensemble_last_ms(self.h)

    def set_launch(self, form: int):
        """Launch form of the air and porous stages: 0 autotune (default), 1 a block per
        tile, 2 tiles padded per plane to multiples of 8, 3 persistent."""
        self._check(lib().lee_ensemble_set_launch(self.h, form))

    def launch(self, porous: bool = False) -> int:
        """The launch form in use (1..3) of the air stage, or of the porous stage; 0 before
        the first step."""
        return lib().lee_ensemble_launch(self.h, 1 if porous else 0)

    def points(self) -> int:
        return sum(m.nx * m.ny * m.nz + m.pnx * m.pny * m.pnz for m in self.members)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            for m in self.members:
                m.h = ctypes.c_void_p()
            lib().lee_ensemble_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_source_series(dirpath: str, cap: int = 100001) -> np.ndarray:
    """Values of dirpath/source_input.dat (calculation_2D.cpp:220-235), at most cap."""
    out = np.zeros(cap)
    n = lib().lee_read_source_series(dirpath.encode(), _p(out), cap)
    if n < 0:
        raise LeeError(int(n), lib().lee_global_error().decode())
    return out[:n].copy()


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else _dp()


def _check(rc, ctx=None):
    if rc != LEE_OK:
        L = lib()
        msg = (L.lee_last_error(ctx) if ctx else L.lee_global_error()) or b""
        raise LeeError(rc, msg.decode(errors="replace"))
    return rc


def load_deck(dirpath: str) -> LeeDeck:
    """Parse input.txt + coordi_* with the library's parser (a ``LeeDeck`` that owns its arrays;
    free with ``free_deck``)."""
    d = LeeDeck()
    _check(lib().lee_deck_load(dirpath.encode(), ctypes.byref(d)))
    return d


def free_deck(d: LeeDeck) -> None:
    lib().lee_deck_free(ctypes.byref(d))


class Solver:
    """One solver context on one GPU (C ABI of include/lee_abi.h)."""

    def __init__(self, d, precision: int = 64, receivers=None, health_every: int = 0, device: int = 0,
                 path: int = PATH_AUTO, integrator: Optional[int] = None, use_graph: bool = False,
                 rk_tile: int = 0, fb_form: int = 0):
        L = lib()
        if isinstance(d, deck.Deck):
            self._deck_obj = d
            self._cdeck = d.to_c()
        else:
            self._deck_obj = None
            self._cdeck = d
        opt = deck.make_options(self._deck_obj, precision=precision, receivers=receivers,
                                health_every=health_every, device=device, path=path, integrator=integrator,
                                use_graph=1 if use_graph else 0, rk_tile=rk_tile, fb_form=fb_form)
        self._opt = opt
        h = ctypes.c_void_p()
        _check(L.lee_create(ctypes.byref(self._cdeck), ctypes.byref(opt), ctypes.byref(h)))
        self.h = h
        self._owned = True
        self._describe()

    def _describe(self):
        info = self.info()
        self.nx, self.ny, self.nz = info.nx, info.ny, info.nz
        self.pnx, self.pny, self.pnz = info.pnx, info.pny, info.pnz
        self.nrec = info.n_receivers
        self.path = info.path
        self.time_domain = self._cdeck.time_domain

    @classmethod
    def _member(cls, handle, cdeck):
        """A view of a context owned by an Ensemble (closing it does not destroy it)."""
        s = cls.__new__(cls)
        s._deck_obj, s._cdeck, s._opt = None, cdeck, None
        s.h = handle
        s._owned = False
        s._describe()
        return s

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            if getattr(self, "_owned", True):
                lib().lee_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> LeeInfo:
        i = LeeInfo()
        _check(lib().lee_get_info(self.h, ctypes.byref(i)), self.h)
        return i

    def receiver_indices(self):
        out = (ctypes.c_int * (3 * max(self.nrec, 1)))()
        lib().lee_receiver_indices(self.h, out, 3 * self.nrec)
        return np.array(out, dtype=np.int64)[: 3 * self.nrec].reshape(-1, 3)

    def frame_setup(self, n: int) -> None:
        _check(lib().lee_frame_setup(self.h, n), self.h)

    def frame_steps(self, mf: int, temp_time: int) -> int:
        return lib().lee_frame_steps(self.h, mf, temp_time)

    def frame_count(self) -> int:
        return lib().lee_frame_count(self.h)

    def step(self, n: int, record: bool = True) -> Optional[np.ndarray]:
        rec = np.zeros((n, self.nrec)) if record else None
        _check(lib().lee_step(self.h, n, _p(rec)), self.h)
        return rec

    def run(self) -> np.ndarray:
        rec = np.zeros((self.time_domain, self.nrec))
        _check(lib().lee_run(self.h, _p(rec), rec.size), self.h)
        return rec

    def state(self, domain: int = 0):
        shp = (self.nz, self.ny, self.nx) if domain == 0 else (self.pnz, self.pny, self.pnx)
        arrs = [np.zeros(shp) for _ in range(4)]
        _check(lib().lee_get_state(self.h, domain, *[_p(a) for a in arrs]), self.h)
        return arrs

    def set_state(self, u, v, w, p, domain: int = 0):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None for a in (u, v, w, p)]
        _check(lib().lee_set_state(self.h, domain, *[_p(a) for a in arrs]), self.h)

    def psi(self):
        out = np.zeros((self.nz, self.nx, 6))
        _check(lib().lee_get_psi(self.h, _p(out), out.size), self.h)
        return out

    def save_restart(self, path: str, index: int) -> None:
        _check(lib().lee_save_restart(self.h, path.encode(), index), self.h)

    def load_restart(self, path: str) -> int:
        idx = ctypes.c_int(0)
        _check(lib().lee_load_restart(self.h, path.encode(), ctypes.byref(idx)), self.h)
        return idx.value

    def write_contour(self, dirpath: str) -> None:
        _check(lib().lee_write_contour(self.h, dirpath.encode()), self.h)

    def write_contour_noalloc(self, dirpath: str) -> None:
        """The same file through lee_write_contour_noalloc (no device allocation: the subsample
        is staged in the library's fixed per-GPU array)."""
        _check(lib().lee_write_contour_noalloc(self.h, dirpath.encode()), self.h)

    def ground_psi(self, K: Optional[int] = None) -> np.ndarray:
        """Ground impedance histories, shape (ny, nx, K), K = the context's pole count (a
        different K is an error: the library writes nx*ny*K values node-major)."""
        have = self._opt.ground_poles if self._opt is not None else 0
        if K is None:
            K = have
        if K != have or K <= 0:
            raise ValueError(f"the context has {have} ground poles, asked for {K}")
        out = np.zeros((self.ny, self.nx, K))
        _check(lib().lee_get_ground_psi(self.h, _p(out), out.size), self.h)
        return out

    def set_source_series(self, ps, period: int = 23) -> None:
        """Time-series source: every `period` steps the pressure gains ps[N/period] times the
        initial Gaussian (calculation_2D::UpdateSource); an empty series removes it."""
        a = np.ascontiguousarray(np.asarray(ps, dtype=np.float64).ravel())
        _check(lib().lee_set_source_series(self.h, _p(a) if a.size else None, a.size, period), self.h)

    def last_ms(self) -> float:
        return lib().lee_last_step_ms(self.h)


class AirPore:
    """Mirror of the reference driver class ``airpore`` (airpore.h:18-67).

    ``AirPore(dir)`` reads ``dir/input.txt`` and the ``coordi_*`` files like
    ``airpore::airpore`` (airpore.cpp:17-316); ``get_output()`` runs the time loop of
    ``airpore::get_output`` (airpore.cpp:400-641) on the GPU and writes the receiver series;
    ``get_FFT_y1()`` writes the spectra like ``airpore::get_FFT_y1`` (airpore.cpp:966-1011).
    """

    def __init__(self, input_dir: str, precision: int = 64, receivers=None):
        self.dir = input_dir
        self._deck = load_deck(input_dir)
        self.solver = Solver(self._deck, precision=precision, receivers=receivers)
        self.series = None

    def get_output(self, contour_dir: Optional[str] = None, restart_dir: Optional[str] = None) -> np.ndarray:
        """The time loop of airpore::get_output (airpore.cpp:400-641).  With `contour_dir` a
        contour file is written whenever the steps done are a multiple of out_difft
        (airpore.cpp:615-618); with `restart_dir` a restart file ctNNN_0.dat whenever they are a
        multiple of restart_out (airpore.cpp:620, 664-700 name the files this way)."""
        if contour_dir is None and restart_dir is None:
            self.series = self.solver.run()
        else:
            self.series = self._loop(contour_dir, restart_dir)
        s = np.ascontiguousarray(self.series)
        _check(lib().lee_write_receivers(self.dir.encode(), _p(s), s.shape[0], s.shape[1]))
        return self.series

    def _loop(self, contour_dir, restart_dir):
        sv, d = self.solver, self._deck
        T = d.time_domain
        out = np.zeros((T, sv.nrec))
        every = [x for x in (d.out_difft if contour_dir else 0, d.restart_out if restart_dir else 0) if x > 0]
        mf, done = 0, 0
        while done < T:
            if mf >= sv.frame_count():
                break
            lim = sv.frame_steps(mf, done)
            sv.frame_setup(mf)
            left = max(lim, 0)
            while left > 0 and done < T:
                nxt = min([e - done % e for e in every] + [left, T - done])
                out[done:done + nxt] = sv.step(nxt)
                done += nxt
                left -= nxt
                if contour_dir and done % d.out_difft == 0:
                    sv.write_contour(contour_dir)
                if restart_dir and done % d.restart_out == 0:
                    n = done // d.restart_out
                    sv.save_restart(os.path.join(restart_dir, f"ct{n:03d}_0.dat"), done)
            mf += 1
            if d.mf_judge != 1:
                break
        return out

    def get_FFT_y1(self) -> None:
        if self.series is None:
            raise RuntimeError("get_output must run first")
        if self._deck.rcv[1] == 0:  # the reference writes spectra only when receiver.y != 0
            return
        s = np.ascontiguousarray(self.series)
        _check(lib().lee_write_spectra(self.dir.encode(), _p(s), s.shape[0], s.shape[1],
                                       self._deck.fft_m, self._deck.dt))

    def close(self):
        self.solver.close()
        free_deck(self._deck)


def fft(re, im, m: int, direction: int = 1):
    re = np.array(re, dtype=np.float64, copy=True)
    im = np.array(im, dtype=np.float64, copy=True)
    _check(lib().lee_fft(direction, m, _p(re), _p(im)))
    return re, im
