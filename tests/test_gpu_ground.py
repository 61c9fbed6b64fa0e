"""Ground impedance (South-face TDBC with K poles, FB_p_v) on the GPU against the oracle
through the C ABI: fields, receivers and histories, FP64 to 1e-10 (same arithmetic, so
expected bitwise), with PML on the sides and top and with graph replay."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REC = [(0.0, 0.55, 0.3), (0.15, 0.75, 0.1)]


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def deck(lee, pml, K=8):
    rng = np.random.default_rng(11)
    poles = [(float(a), float(r)) for a, r in zip(rng.uniform(0.05, 0.8, K), rng.uniform(0.0, 2.0, K))]
    bc = (2, 1, 2, 2, 2, 2) if pml else (1, 1, 1, 1, 1, 1)
    return lee.deck.box(24, 28, 20, 0.05, pml=pml, bc=bc, src=(0.0, 0.6, 0.25), gauss_width=0.25, time_domain=60,
                        ground_poles=poles, ground_zinf=900.0, ground_tau=1.5e-4)


@pytest.mark.parametrize("pml,graph,K", [(0, False, 8), (5, False, 8), (5, True, 8), (5, False, 3)])
def test_ground_impedance_parity(lee, orc, pml, graph, K):
    d = deck(lee, pml, K)
    print(f"[gpu ground] pml={pml} graph={graph} poles={K}", flush=True)
    s = lee.Solver(d, precision=64, receivers=REC, use_graph=graph)
    s.frame_setup(0)
    got = s.step(d.time_domain)
    o = orc.Oracle(d, receivers=REC)
    o.frame_setup(0)
    ref = o.step(d.time_domain)
    e = rel(got, ref)
    fe = max(rel(a, b) for a, b in zip(s.state(), o.state()))
    pe = rel(s.ground_psi(K), o.ground_psi(K))
    print(f"  rec {e:.2e} fields {fe:.2e} histories {pe:.2e}", flush=True)
    s.close()
    assert e <= 1e-10 and fe <= 1e-10 and pe <= 1e-10


@pytest.mark.parametrize("graph", [False, True])
def test_ground_impedance_on_terrain_parity(lee, orc, graph):
    """The ground plane on the terrain-following grid (a 0.15 m hill under the source), FP64
    against the oracle: fields, receivers and histories."""
    d = deck(lee, 5)
    d.hills = [(0.1, 0.7, 0.15, 0.3)]
    print(f"[gpu ground] terrain graph={graph}", flush=True)
    s = lee.Solver(d, precision=64, receivers=REC, use_graph=graph)
    s.frame_setup(0)
    got = s.step(d.time_domain)
    o = orc.Oracle(d, receivers=REC)
    o.frame_setup(0)
    ref = o.step(d.time_domain)
    e = rel(got, ref)
    fe = max(rel(a, b) for a, b in zip(s.state(), o.state()))
    pe = rel(s.ground_psi(), o.ground_psi(8))
    print(f"  rec {e:.2e} fields {fe:.2e} histories {pe:.2e}", flush=True)
    s.close()
    assert np.abs(ref).max() > 0
    assert e <= 1e-10 and fe <= 1e-10 and pe <= 1e-10


def test_ground_impedance_fp32(lee):
    d = deck(lee, 5)
    a = lee.Solver(d, precision=64, receivers=REC)
    b = lee.Solver(d, precision=32, receivers=REC)
    a.frame_setup(0); b.frame_setup(0)
    e = rel(b.step(d.time_domain), a.step(d.time_domain))
    print(f"[gpu ground] fp32 vs fp64 receivers {e:.2e}", flush=True)
    a.close(); b.close()
    assert e < 1e-4


@pytest.mark.parametrize("K", [8, 3])
def test_ground_histories_survive_a_restart(lee, tmp_path, K):
    """The restart file carries the ground histories node-major ([i + nx*j][K]) whatever the
    device layout (pole-major, stride nx*ny): a fresh context that loads it holds the saved
    histories (to the file's ten significant digits) and continues the run, for the full
    eight poles and for three."""
    d = deck(lee, 5, K)
    print(f"[gpu ground] restart poles={K}", flush=True)
    a = lee.Solver(d, precision=64, receivers=REC)
    a.frame_setup(0)
    a.step(30)
    path = str(tmp_path / "ct001_0.dat")
    a.save_restart(path, 30)
    saved = a.ground_psi(K)
    rest = a.step(30)
    b = lee.Solver(d, precision=64, receivers=REC)
    b.frame_setup(0)
    assert b.load_restart(path) == 30
    loaded = b.ground_psi(K)
    cont = b.step(30)
    e_psi, e_run = rel(loaded, saved), rel(cont, rest)
    print(f"  histories {e_psi:.2e} continuation {e_run:.2e}", flush=True)
    a.close()
    b.close()
    assert np.abs(saved).max() > 0
    assert e_psi <= 1e-9 and e_run <= 1e-6
