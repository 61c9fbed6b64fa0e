"""Embedded porous obstacles in the oracle against the reference's own air::cal_hill_var and
stencils (oracle/_ref/ref_probe), bit for bit: the index box, the material mask of all five
curve kinds, and the pressure and velocity increments with the obstacle overriding PML and
wind terms (air.cpp:1907-1945, 2294-2318)."""
import numpy as np
import pytest

from oracle import oracle as O
from obstacle_decks import obstacle_deck

needs_probe = pytest.mark.skipif(not O.probe_available(), reason="reference probe not built (make -C oracle ref)")


def _interior(shape):
    m = np.zeros(shape, bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


@needs_probe
@pytest.mark.parametrize("kind,wind", [(1, 0), (2, 5), (3, 0), (4, 4), (5, 0)])
def test_obstacle_mask_and_stencils_match_the_reference(lee, kind, wind):
    kw = dict(velocity_method=wind, velocity_coef=2.0) if wind else {}
    d = obstacle_deck(lee, kind, **kw)
    nz, ny, nx = len(d.air_z), len(d.air_y), len(d.air_x)
    shape = (nz, ny, nx)
    rng = np.random.default_rng(40 + kind)
    f1 = [rng.standard_normal(shape) for _ in range(4)]
    f2 = [rng.standard_normal(shape) for _ in range(4)]
    res = O.run_probe(d, f1, f2, counter=3, obstacle=True)
    o = O.Oracle(d)
    o.frame_setup(0)
    kind_o, box, mask = o.obstacle()
    assert kind_o == kind
    assert list(box) == [int(v) for v in res["obox"]]
    assert np.array_equal(mask.ravel(), res["omask"].astype(np.int32))
    assert (mask == 1).sum() > 0
    if kind == 5:
        assert (mask == 2).sum() > 0   # both trunk passes mark nodes
    # same PML auxiliaries as the probe's fresh air object
    slabs, fields = ["x1", "x2", "y1", "y2", "z1", "z2"], ["u", "v", "w", "p"]
    for s in range(6):
        for f in range(4):
            o.q_set(s, f, res[f"q0_{slabs[s]}_{fields[f]}"])
    for _ in range(4):
        o.tick_gate()
    m = _interior(shape)
    dp = o.air_dp(*f1)
    assert np.array_equal(dp[m], res["odp"].reshape(shape)[m])
    for a, key in zip(o.air_duvw(*f1), ("odu", "odv", "odw")):
        assert np.array_equal(a[m], res[key].reshape(shape)[m]), key
    # the override changed the increments inside the obstacle and nowhere else
    inside = np.zeros(shape, bool)
    inside[box[4]:box[5] + 1, box[2]:box[3] + 1, box[0]:box[1] + 1] = mask > 0
    inside &= m
    assert not np.array_equal(res["odp"].reshape(shape)[inside], res["dp"].reshape(shape)[inside])
    assert np.array_equal(res["odp"].reshape(shape)[m & ~inside], res["dp"].reshape(shape)[m & ~inside])


def test_obstacle_material_update_closed_form(lee):
    """Inside a box obstacle with uniform fields the velocity relaxes by exactly invBeta per
    step and the pressure is frozen: u' = u / (1 + sigma dt / rho_e) (air.cpp:1910-1914)."""
    d = obstacle_deck(lee, 2, bc=(1, 1, 1, 1, 1, 1), pml=0)
    o = O.Oracle(d)
    o.frame_setup(0)
    _, box, mask = o.obstacle()
    shape = o.shape
    u = np.full(shape, 0.7); v = np.full(shape, -0.3); w = np.full(shape, 0.2); p = np.full(shape, 5.0)
    du, dv, dw = o.air_duvw(u, v, w, p)
    dp = o.air_dp(u, v, w, p)
    q, por, res = d.hill_mat[1]
    Aav = d.c ** 2 / d.gamma / d.Pav
    rho_e = q * q / por / Aav
    ib = 1.0 / (1.0 + res * d.dt / rho_e)
    k, j, i = np.argwhere(mask == 1)[len(np.argwhere(mask == 1)) // 2]
    k, j, i = k + box[4], j + box[2], i + box[0]
    assert du[k, j, i] == pytest.approx(0.7 * ib - 0.7, rel=1e-14)
    assert dv[k, j, i] == pytest.approx(-0.3 * ib + 0.3, rel=1e-14)
    assert dp[k, j, i] == 0.0


@pytest.mark.parametrize("change,msg", [
    (dict(hill_judge=7), "no mask"),
    (dict(hill_box=(5.0, 6.0, 5.0, 6.0, 5.0, 6.0)), None),
])
def test_oracle_rejects_bad_obstacles(lee, change, msg):
    d = obstacle_deck(lee, 1)
    for k, v in change.items():
        setattr(d, k, v)
    if msg is None:
        # bounds that match no node leave the reference's box at node 0 (air.cpp:987-1005)
        o = O.Oracle(d)
        kind, box, mask = o.obstacle()
        assert list(box) == [0, 0, 0, 0, 0, 0] and mask.shape == (1, 1, 1)
        return
    with pytest.raises(Exception, match=msg):
        O.Oracle(d)
