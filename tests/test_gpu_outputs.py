"""Checkpoint / restart and contour files in the reference's text formats (airpore.cpp:664-862,
calculation_2D.cpp:497-545, air.cpp:1305-1474), through the C ABI.

* Layout: the header line, then u v w p per node in the reference's f[i][j][k] loop order,
  the mean flow, the PML auxiliaries slab by slab in the reference's loop orders, the porous
  fields -- checked value by value against the FP64 state and the oracle's auxiliaries.
* Round trip: a run continued from the file follows the uninterrupted run to the 10-digit
  precision of the text format.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REC = [(0.0, 0.55, 0.45), (0.15, 0.75, 0.3)]


def _deck(lee, bc=(2, 3, 5, 2, 2, 2)):
    return lee.deck.box(16, 18, 14, 0.05, pml=4, bc=bc, porous_cells=6, src=(0.0, 0.45, 0.3),
                        gauss_width=0.25, time_domain=60, pml_y_time=2, velocity_method=5, velocity_coef=2.0)


def _tokens(path):
    with open(path) as fh:
        head = fh.readline().split()
        body = fh.read().split()
    return head, body


def test_restart_layout_matches_the_reference_order(lee, orc, tmp_path):
    d = _deck(lee)
    print("[gpu restart] layout", flush=True)
    s = lee.Solver(d, precision=64, receivers=REC)
    s.frame_setup(0)
    s.step(12)
    o = orc.Oracle(d, receivers=REC)
    o.frame_setup(0)
    o.step(12)
    path = str(tmp_path / "ct001_0.dat")
    s.save_restart(path, 12)
    head, body = _tokens(path)
    assert head == ["1", "0", "12"]
    nx, ny, nz = s.nx, s.ny, s.nz
    u, v, w, p = s.state()  # (nz, ny, nx)
    nf = nx * ny * nz
    vals = np.array([float(x) for x in body[: 4 * nf]])
    exp = np.stack([u, v, w, p], axis=-1).transpose(2, 1, 0, 3).reshape(-1)  # i, j, k, field
    assert np.allclose(vals, exp, rtol=1e-9, atol=1e-300)
    U, V, W = o.wind()
    pos = 4 * nf
    wv = np.array([float(x) for x in body[pos: pos + 3 * nf]])
    assert np.allclose(wv, np.stack([U, V, W], axis=-1).transpose(2, 1, 0, 3).reshape(-1), rtol=1e-9, atol=1e-300)
    pos += 3 * nf
    # auxiliaries: x slabs (i, j, k), y slabs (j, k, i), z slabs (k, i, j), each Qu Qv Qw Qp
    for slab in range(6):
        qs = [o.q_get(slab, f) for f in range(4)]
        cnt = qs[0].size
        if cnt == 0:
            continue
        ax = slab // 2
        if ax == 0:
            nn = cnt // (ny * nz)
            arr = np.stack([q.reshape(nz, ny, nn) for q in qs], axis=-1).transpose(2, 1, 0, 3)
        elif ax == 1:
            nn = cnt // (nx * nz)
            arr = np.stack([q.reshape(nz, nn, nx) for q in qs], axis=-1).transpose(1, 0, 2, 3)
        else:
            nn = cnt // (nx * ny)
            arr = np.stack([q.reshape(nn, ny, nx) for q in qs], axis=-1).transpose(0, 2, 1, 3)
        got = np.array([float(x) for x in body[pos: pos + 4 * cnt]])
        assert np.allclose(got, arr.reshape(-1), rtol=1e-9, atol=1e-300), slab
        pos += 4 * cnt
    pu, pv, pw, pp = s.state(1)
    npf = pu.size
    got = np.array([float(x) for x in body[pos: pos + 4 * npf]])
    assert np.allclose(got, np.stack([pu, pv, pw, pp], axis=-1).transpose(2, 1, 0, 3).reshape(-1), rtol=1e-9, atol=1e-300)
    pos += 4 * npf
    assert body[pos] == "LEE-EXT"
    # the first value of a section is padded to width 18 (ofstream::width), every value has
    # an explicit sign and 10 digits after the point
    with open(path) as fh:
        fh.readline()
        line = fh.readline()
    assert line[0] == " " and line[1] in "+-" and len(line.split()[0]) == 17
    assert all(t[0] in "+-" and len(t.split("e")[0]) == 13 for t in line.split()[:50])
    s.close()


@pytest.mark.parametrize("integrator", [0, 1])
def test_restart_round_trip(lee, tmp_path, integrator):
    # the TDBC is an FB_p_v boundary: RK4 runs with an absorbing East face
    d = _deck(lee, bc=(2, 3, 2, 2, 2, 2)) if integrator else _deck(lee)
    d.integrator = integrator
    print(f"[gpu restart] round trip integrator={integrator}", flush=True)
    a = lee.Solver(d, precision=64, receivers=REC)
    a.frame_setup(0)
    a.step(15)
    path = str(tmp_path / "ct.dat")
    a.save_restart(path, 15)
    ra = a.step(20)
    b = lee.Solver(d, precision=64, receivers=REC)
    b.frame_setup(0)
    assert b.load_restart(path) == 15
    assert b.info().steps_done == 15
    rb = b.step(20)
    e = np.linalg.norm(rb - ra) / np.linalg.norm(ra)
    fe = max(np.abs(x - y).max() / max(np.abs(x).max(), 1e-300) for x, y in zip(a.state(), b.state()))
    print(f"  continued run vs uninterrupted: receivers {e:.2e}, fields {fe:.2e}", flush=True)
    a.close(); b.close()
    assert e < 1e-8 and fe < 1e-8


def _contour(path):
    zones = []
    with open(path) as fh:
        lines = fh.read().splitlines()
    i = 0
    head = None
    if lines[0].startswith("VARIABLES"):
        head = lines[0]
        i = 1
    while i < len(lines):
        line = lines[i]
        assert line.startswith('ZONE T = "')
        dims = line.split(",", 1)[1]
        I, J, K = [int(x.split("=")[1]) for x in dims.split(",")[:3]]
        rows = np.array([[float(v) for v in lines[i + 1 + r].split()] for r in range(I * J * K)])
        zones.append((line.split('"')[1], (I, J, K), rows))
        i += 1 + I * J * K
    return head, zones


@pytest.mark.parametrize("porous", [False, True])
def test_contour_is_the_reference_subsample(lee, tmp_path, porous):
    """Every 4th node plus the last on each axis, k outer / i inner, X Y Z P per line; the
    VARIABLES line only over a rigid ground; file ptNNN with NNN = (steps + 1) / out_difft."""
    d = lee.deck.box(18, 23, 14, 0.05, pml=4, bc=(2, 3 if porous else 1, 2, 2, 2, 2),
                     porous_cells=6 if porous else 0, src=(0.0, 0.5, 0.3), gauss_width=0.25, time_domain=40,
                     out_difft=10)
    print(f"[gpu contour] porous={porous}", flush=True)
    s = lee.Solver(d, precision=64)
    s.frame_setup(0)
    s.step(20)
    s.write_contour(str(tmp_path))
    path = tmp_path / "pt002.dat"
    assert path.exists()
    head, zones = _contour(str(path))
    assert (head == 'VARIABLES = "X", "Y", "Z", "P"') == (not porous)
    assert [z[0] for z in zones] == (["Air Zone", "south pore Zone"] if porous else ["Air Zone"])

    def pick(n):
        idx = list(range(0, n, 4))
        if idx[-1] != n - 1:
            idx.append(n - 1)
        return np.array(idx)

    grids = [(s.state()[3], d.air_x, d.air_y, d.air_z)]
    if porous:
        grids.append((s.state(1)[3], d.pore_x, d.pore_y, d.pore_z))
    for (name, (I, J, K), rows), (p, X, Y, Z) in zip(zones, grids):
        ii, jj, kk = pick(p.shape[2]), pick(p.shape[1]), pick(p.shape[0])
        assert (I, J, K) == (len(ii), len(jj), len(kk))
        sub = p[np.ix_(kk, jj, ii)]  # (k, j, i): k outer, i inner
        kg, jg, ig = np.meshgrid(kk, jj, ii, indexing="ij")
        exp = np.stack([X[ig].ravel(), Y[jg].ravel(), Z[kg].ravel(), sub.ravel()], axis=1)
        assert np.allclose(rows, exp, rtol=5e-6, atol=1e-300), name
    s.close()
