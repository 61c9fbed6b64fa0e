"""Ensemble runner (lee_ensemble_*): uneven domains with different ground, wind and source
parameters stepped by one launch per kernel kind.  Each member must reproduce, bit for bit,
the same deck run alone through lee_step (same kernels, same arithmetic), in FP64 and FP32,
eager and graph-replayed; FP64 members also match the oracle (<= 1e-10)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def members(lee):
    B = lee.deck.box
    out = []
    # porous ground, log wind, PML gate opening mid-run
    d = B(20, 26, 18, 0.05, pml=4, bc=(2, 3, 2, 2, 2, 2), porous_cells=6, src=(0.0, 0.5, 0.35), gauss_width=0.2,
          time_domain=40, velocity_method=7, pml_y_time=9)
    d.wind_ustar, d.wind_z0 = 0.3, 0.02
    d.pore_q, d.pore_porosity, d.pore_resistivity = 1.6, 0.35, 8.0e4
    out.append(d)
    # rigid ground, no PML in y, different size and wind
    d = B(34, 22, 16, 0.05, pml=5, bc=(2, 1, 1, 1, 2, 2), src=(0.1, 0.4, 0.3), gauss_width=0.25, time_domain=40,
          velocity_method=7)
    d.wind_ustar, d.wind_z0 = 0.5, 0.05
    out.append(d)
    # porous with a thicker layer, calm air
    d = B(18, 30, 20, 0.05, pml=4, bc=(2, 3, 2, 2, 2, 2), porous_cells=9, src=(-0.1, 0.8, 0.5), gauss_width=0.2,
          time_domain=40, velocity_method=7, pml_y_time=1)
    d.wind_ustar, d.wind_z0 = 0.05, 0.02
    d.pore_q, d.pore_porosity, d.pore_resistivity = 1.2, 0.45, 2.0e4
    out.append(d)
    for d in out:
        d.integrator = 1
    return out


def receivers(d):
    x, y, z = d.air_x, d.air_y, d.air_z
    return [(x[len(x) // 2], y[len(y) // 2], z[len(z) // 2]), (x[3], y[-5], z[5])]


@pytest.mark.parametrize("precision,graph", [(64, False), (64, True), (32, True)])
def test_ensemble_members_match_single_runs(lee, orc, precision, graph):
    ds = members(lee)
    recs = [receivers(d) for d in ds]
    print(f"[gpu ensemble] fp{precision} graph={graph}", flush=True)
    e = lee.Ensemble(ds, precision=precision, receivers=recs, use_graph=graph)
    got = [np.concatenate(parts) for parts in zip(*[e.step(n) for n in (7, 33)])]
    for d, dk in enumerate(ds):
        s = lee.Solver(dk, precision=precision, receivers=recs[d], integrator=1)
        s.frame_setup(0)
        ref = s.step(40)
        assert np.array_equal(got[d], ref), d
        for a, b in zip(e.members[d].state(), s.state()):
            assert np.array_equal(a, b), d
        if dk.has_pore:
            for a, b in zip(e.members[d].state(1), s.state(1)):
                assert np.array_equal(a, b), d
        s.close()
        if precision == 64:
            o = orc.Oracle(dk, receivers=recs[d], integrator=1)
            o.frame_setup(0)
            r = o.step(40)
            err = rel(got[d], r)
            print(f"  member {d}: vs oracle {err:.2e}", flush=True)
            assert err <= 1e-10
    e.close()


@pytest.mark.parametrize("form", [1, 2, 3])
def test_ensemble_launch_forms_match_single_runs(lee, form):
    """Every launch form of the air and porous stages (a block per tile, tiles padded per
    plane, persistent grid) steps each member bit for bit as the deck run alone."""
    ds = members(lee)
    recs = [receivers(d) for d in ds]
    for precision, graph in [(64, False), (32, True)]:
        print(f"[gpu ensemble] launch form {form} fp{precision} graph={graph}", flush=True)
        e = lee.Ensemble(ds, precision=precision, receivers=recs, use_graph=graph, launch=form)
        assert e.launch() == 0
        got = [np.concatenate(parts) for parts in zip(*[e.step(n) for n in (5, 24)])]
        assert e.launch() == form and e.launch(porous=True) == form
        for d, dk in enumerate(ds):
            s = lee.Solver(dk, precision=precision, receivers=recs[d], integrator=1)
            s.frame_setup(0)
            assert np.array_equal(got[d], s.step(29)), d
            for a, b in zip(e.members[d].state(), s.state()):
                assert np.array_equal(a, b), d
            if dk.has_pore:
                for a, b in zip(e.members[d].state(1), s.state(1)):
                    assert np.array_equal(a, b), d
            s.close()
        e.close()


def test_ensemble_autotuner_keeps_a_form(lee):
    ds = members(lee)
    e = lee.Ensemble(ds, precision=32, receivers=[receivers(d) for d in ds], use_graph=True)
    e.step(3)
    assert e.launch() in (1, 2, 3) and e.launch(porous=True) in (1, 2, 3)
    with pytest.raises(lee.LeeError):
        e.set_launch(4)
    e.set_launch(2)
    assert e.launch() == 0
    e.step(2)
    assert e.launch() == 2 and e.launch(porous=True) == 2
    e.close()


def kind_members(lee, kind):
    """Members sharing a grid and wind kind other than the flat log-wind set above: over
    terrain (log wind), with U V W wind fields (method 5), and in calm air."""
    B = lee.deck.box
    out = []
    if kind == "terrain":
        d = B(22, 26, 18, 0.05, pml=4, bc=(2, 3, 2, 2, 2, 2), porous_cells=6, src=(0.0, 0.5, 0.35), gauss_width=0.2,
              time_domain=30, velocity_method=7)
        d.hills, d.wind_ustar, d.wind_z0 = [(0.1, 0.6, 0.15, 0.3)], 0.3, 0.02
        out.append(d)
        d = B(30, 20, 16, 0.05, pml=4, bc=(2, 1, 2, 2, 2, 2), src=(0.1, 0.4, 0.3), gauss_width=0.25,
              time_domain=30, velocity_method=7)
        d.hills, d.wind_ustar, d.wind_z0 = [(-0.2, 0.5, 0.1, 0.25), (0.3, 0.8, 0.08, 0.2)], 0.5, 0.05
        out.append(d)
    elif kind == "uvw":
        out.append(B(24, 28, 18, 0.05, pml=4, bc=(2, 3, 2, 2, 2, 2), porous_cells=6, src=(0.0, 0.6, 0.4),
                     gauss_width=0.25, time_domain=30, velocity_method=5, velocity_coef=3.0))
        out.append(B(18, 22, 16, 0.05, pml=4, bc=(2, 1, 2, 2, 2, 2), src=(0.05, 0.5, 0.3), gauss_width=0.2,
                     time_domain=30, velocity_method=5, velocity_coef=2.0))
    else:
        out.append(B(20, 24, 18, 0.05, pml=4, bc=(2, 3, 2, 2, 2, 2), porous_cells=7, src=(0.0, 0.5, 0.4),
                     gauss_width=0.25, time_domain=30))
        out.append(B(26, 18, 16, 0.05, pml=4, bc=(2, 1, 2, 2, 2, 2), src=(0.1, 0.4, 0.3), gauss_width=0.2,
                     time_domain=30))
    for d in out:
        d.integrator = 1
    return out


@pytest.mark.parametrize("kind", ["terrain", "uvw", "calm"])
@pytest.mark.parametrize("form", [0, 3])
def test_ensemble_grid_and_wind_kinds_match_single_runs(lee, orc, kind, form):
    """The batched kernels of the other grid and wind kinds (terrain, U V W wind, no wind),
    autotuned and persistent: each member bit for bit as the deck run alone, FP64 members
    within 1e-10 of the oracle."""
    ds = kind_members(lee, kind)
    recs = [receivers(d) for d in ds]
    for precision, graph in [(64, False), (32, True)]:
        print(f"[gpu ensemble] {kind} launch {form} fp{precision} graph={graph}", flush=True)
        e = lee.Ensemble(ds, precision=precision, receivers=recs, use_graph=graph, launch=form)
        got = [np.concatenate(parts) for parts in zip(*[e.step(n) for n in (4, 21)])]
        for d, dk in enumerate(ds):
            s = lee.Solver(dk, precision=precision, receivers=recs[d], integrator=1)
            s.frame_setup(0)
            assert np.array_equal(got[d], s.step(25)), (kind, d)
            for a, b in zip(e.members[d].state(), s.state()):
                assert np.array_equal(a, b), (kind, d)
            s.close()
            if precision == 64:
                o = orc.Oracle(dk, receivers=recs[d], integrator=1)
                o.frame_setup(0)
                err = rel(got[d], o.step(25))
                print(f"  member {d}: vs oracle {err:.2e}", flush=True)
                assert err <= 1e-10
        e.close()
