"""FP64 GPU path vs the CPU oracle on the decks that matter beyond small uniform boxes: an
anisotropic, geometrically stretched separable grid (every metric differs per axis and per
point, so a swapped axis or a wrong reciprocal table fails), the reference's own shipped deck
(tests/golden/shipped_deck: 141x201x441, FB_p_v, East TDBC, PML on x and z, wind method 4,
y1 gate at 4500) and the 96x64x48 porous-ground workload (RK4 and the reference's FB_p_v),
all through the C ABI.  Bar: <= 1e-10 relative L2 on the receiver series and on every final
field (the FP64 device path keeps the oracle's operation order, so the differences are zero
or at rounding level).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIPPED = os.path.join(ROOT, "tests", "golden", "shipped_deck")
PATHS = {"split": 1, "fused": 2}


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def compare(s, o, got, ref, psi=False):
    """Largest relative L2 over the receiver series and every field of both domains."""
    e_rec = rel(got, ref)
    e_f = max(rel(a, b) for a, b in zip(s.state(), o.state()))
    if s.pnx:
        e_f = max(e_f, max(rel(a, b) for a, b in zip(s.state(1), o.state(1))))
    if psi:
        e_f = max(e_f, rel(s.psi(), o.psi()))
    return e_rec, e_f


def stretched(lee, integrator):
    """dx != dy != dz, each axis stretched by its own ratio, PML on every air face but South,
    a porous ground (FB) or a porous ground under two hills (RK4)."""
    d = lee.deck.box(26, 30, 22, spacing=(0.04, 0.05, 0.035), stretch=(1.03, 1.02, 1.04), pml=5,
                     bc=(2, 3, 2, 2, 2, 2), porous_cells=7, src=(0.1, 0.6, 0.35), gauss_width=0.25,
                     time_domain=40, pml_y_time=4)
    if integrator == "rk4":
        d.integrator = 1
        d.hills = [(0.15, 0.7, 0.12, 0.3)]
    return d


REC_S = [(0.0, 0.55, 0.45), (0.15, 0.75, 0.3), (-0.2, 0.35, 0.6), (0.05, 1.05, 0.2)]


@pytest.mark.parametrize("kind", ["fb-split", "fb-fused", "rk4"])
def test_stretched_anisotropic_grid(lee, orc, kind):
    d = stretched(lee, "rk4" if kind == "rk4" else "fb")
    dx = np.diff(d.air_x)
    assert dx.max() / dx.min() > 1.5 and len({d.dx, d.dy, d.dz}) == 3
    path = PATHS[kind.split("-")[1]] if kind.startswith("fb") else 0
    print(f"[gpu decks] stretched anisotropic {kind} shape={d.shape}", flush=True)
    s = lee.Solver(d, precision=64, receivers=REC_S, path=path)
    s.frame_setup(0)
    got = s.step(d.time_domain)
    o = orc.Oracle(d, receivers=REC_S)
    o.frame_setup(0)
    ref = o.step(d.time_domain)
    assert np.abs(ref).max() > 0
    e_rec, e_f = compare(s, o, got, ref)
    print(f"  rec rel-L2 {e_rec:.3e}  field rel-L2 {e_f:.3e}", flush=True)
    s.close()
    assert e_rec <= 1e-10 and e_f <= 1e-10


@pytest.fixture(scope="module")
def shipped(lee, orc):
    """The shipped deck advanced 20 steps by the oracle (receivers next to the source, which
    sits on the ground at y = -0.325 m: the deck's own receivers are 0.1-2.5 m away and see
    nothing in 20 steps of 4 us)."""
    d = lee.load_deck(SHIPPED)
    rec = [(0.0, -0.30, 0.02), (0.01, -0.34, 0.03), (0.0, -0.325, 0.05), (0.02, -0.31, 0.0)]
    print("[gpu decks] shipped deck: oracle, 20 steps", flush=True)
    o = orc.Oracle(d, receivers=rec)
    o.frame_setup(0)
    ref = o.step(20)
    yield d, rec, o, ref
    lee.free_deck(d)


@pytest.mark.parametrize("path", ["split", "fused"])
def test_shipped_deck(lee, shipped, path):
    d, rec, o, ref = shipped
    print(f"[gpu decks] shipped deck 141x201x441 path={path}", flush=True)
    s = lee.Solver(d, precision=64, receivers=rec, path=PATHS[path])
    assert (s.nx, s.ny, s.nz) == (141, 201, 441)
    s.frame_setup(0)
    got = s.step(20)
    assert np.abs(ref).max() > 0
    e_rec, e_f = compare(s, o, got, ref, psi=True)
    print(f"  rec rel-L2 {e_rec:.3e}  field rel-L2 {e_f:.3e}", flush=True)
    s.close()
    assert e_rec <= 1e-10 and e_f <= 1e-10


@pytest.mark.parametrize("scheme", ["rk4", "fb"])
def test_c0_workload(lee, orc, scheme):
    """96x64x48 at 0.05 m over the 8-point porous ground, 20-point PML on the sides and top,
    five receivers at 1.5 m: 200 steps (14.6 ms, so the pulse has passed the receivers 1 m and
    1.75 m away)."""
    d = lee.deck.config_c0(time_domain=200, reference_scheme=scheme == "fb")
    rec = lee.deck.receivers_c0()
    print(f"[gpu decks] C0 {scheme} 200 steps", flush=True)
    s = lee.Solver(d, precision=64, receivers=rec)
    s.frame_setup(0)
    got = s.step(200)
    o = orc.Oracle(d, receivers=rec)
    o.frame_setup(0)
    ref = o.step(200)
    # receivers 1 and 2 (1 m and 1.75 m from the source) see the pulse; the others lie beyond
    # the grid's y extent (3.15 m) and are clamped to its last node, inside the layers
    assert (np.abs(ref).max(axis=0)[:2] > 1e-3).all()
    e_rec, e_f = compare(s, o, got, ref)
    print(f"  rec rel-L2 {e_rec:.3e}  field rel-L2 {e_f:.3e}", flush=True)
    s.close()
    assert e_rec <= 1e-10 and e_f <= 1e-10
