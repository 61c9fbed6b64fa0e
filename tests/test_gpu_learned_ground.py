"""The learned ground end to end on the GPU: reflection data from column runs of the full
air + porous solve, the closure's calibration, residue fits, the network trained by the
BF16-MFMA kernels; then, on held-out grounds, what the network gains over the method
without it (the reference's fixed six-pole TDBC for every ground), the full porous solve as
judge (column reflection and the 3-D judge box), and the fallback outside the hull."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained(lee):
    G = lee.ground
    cols = G.Columns()
    grounds = [g for g in G.sample_grounds(400, 7) if G.in_hull(g)][:40]
    X, A, E = G.training_set(grounds[:32], cols)
    print(f"[gpu learned] 32 training grounds, fit error median {np.median(E):.4f} max {E.max():.4f}", flush=True)
    lg = G.LearnedGround(seed=3)
    keep = E <= G.HULL_FIT_R
    mse = lg.train(X[keep], A[keep], epochs=600, batch=16, lr=2e-3)
    print(f"[gpu learned] trained on {keep.sum()} grounds, mse {mse:.2e}", flush=True)
    yield G, cols, lg, grounds[32:]
    lg.close()


def test_fit_reproduces_the_closure_run(trained):
    """The calibrated closure predicts a run of the impedance column with the fitted residues
    (the Moebius identification is exact up to the runs' rounding)."""
    G, cols, _, held = trained
    g = held[0]
    r = cols.measure(g)
    a, e = r.fit()
    f, RT = cols.impedance(g, a)
    b = (f >= G.F_BAND[0]) & (f <= G.F_BAND[1])
    dev = float(np.abs(RT - r.predict(a))[b].max())
    print(f"[gpu learned] {g}: fit error {e:.4f}, run vs prediction {dev:.2e}", flush=True)
    assert dev < 1e-3 and e < G.HULL_FIT_R


def test_network_beats_the_six_pole_tdbc(trained):
    G, cols, lg, held = trained
    pred = lg.residues(held)
    net, six = [], []
    for g, a in zip(held, pred):
        r = cols.measure(g)
        net.append(r.error(a))
        six.append(r.error(np.array(G.SIX_POLE_A), poles=G.SIX_POLE_R))
        print(f"[gpu learned] {g}: |dR| network {net[-1]:.4f}  six-pole {six[-1]:.4f}", flush=True)
    net, six = np.array(net), np.array(six)
    assert net.max() < 0.05
    assert np.mean(net < six) >= 0.75 and net.mean() < 0.5 * six.mean()


def test_judge_box_and_fallback(trained, lee):
    G, cols, lg, held = trained
    g = held[1]
    a = lg.residues([g])[0]
    n_net, t_net = G.judge_3d(g, a)
    n_six, t_six = G.judge_3d(g, np.array(G.SIX_POLE_A), poles=G.SIX_POLE_R)
    print(f"[gpu learned] judge box {g}: network IL {n_net:.3f} dB (1/3 oct {t_net:.3f}); "
          f"six-pole {n_six:.3f} dB (1/3 oct {t_six:.3f})", flush=True)
    assert t_net <= max(G.JUDGE_BOUND_DB, t_six)
    out, rep = lg.decide(G.Ground(1.5e4, 0.3, 2.0, 8), cols)
    assert out is None and rep["in_hull"] is False


def test_judge_box_over_a_hill(trained):
    """Over terrain the impedance plane stands for the porous layer's top plane, which the
    layer, shifted rigidly with the ground, meshes without the air column's compression: with
    the fitted residues the closure tracks the full porous solve over a hill (0.35 m, sigma
    1.8 m: the air cells over it 15 % shorter, slopes up to 0.12) as closely as on flat ground.
    Taking the air plane's compressed metric instead misses by 0.8 dB here."""
    G, cols, _, held = trained
    g = held[0]
    a, e = cols.measure(g).fit()
    n, t = G.judge_3d(g, a, hills=[(0.0, 2.6, 0.35, 1.8)])
    print(f"[gpu learned] judge box over a hill, {g}: fit {e:.4f}, IL {n:.3f} dB (1/3 oct {t:.3f})", flush=True)
    assert t <= G.JUDGE_BOUND_DB and n < 1.0
