"""Decks with an embedded porous obstacle (input lines 24-29, airpore.cpp:270-299) for the
five curve kinds of air::cal_hill_var (air.cpp:1018-1115).  Shared by the oracle-vs-reference
and the GPU parity tests."""
import numpy as np

# canopy / trunk / hill materials: (q, porosity, resistivity)
MATS = ((1.790799820223793, 0.8769089596793096, 5.137565e4),
        (1.4, 0.6, 2.0e4),
        (0.52, 0.5712, 8.0e4))


def coef(*v):
    c = list(v) + [0.0] * (12 - len(v))
    return tuple(float(x) for x in c)


def obstacle_deck(lee, kind, **kw):
    """A deck whose obstacle box overlaps the PML slabs (absorbing faces everywhere)."""
    args = dict(pml=4, bc=(2, 1, 2, 2, 2, 2), src=(0.0, 0.45, 0.45), gauss_width=0.2, time_domain=30, pml_y_time=2)
    args.update(kw)
    if kind == 5:
        # the tree sits at fixed coordinates (canopy about (0, 3) up to z = 3 m, a second
        # trunk at (2, 5)): a 0.1 m grid that holds both
        d = lee.deck.box(32, 34, 33, 0.1, **dict(args, src=(0.0, 3.0, 1.2), gauss_width=0.4))
        d.air_x = -0.8 + 0.1 * np.arange(32)
        d.air_y = 2.0 + 0.1 * np.arange(34)
        d.air_z = 0.1 * np.arange(33)
        d.hill_box = (-0.8 + 0.1 * 2, -0.8 + 0.1 * 30, 2.0 + 0.1 * 3, 2.0 + 0.1 * 32, 0.1, 3.0)
        d.hill_coef = coef()
    else:
        d = lee.deck.box(20, 22, 18, 0.05, **args)
        x, y, z = d.air_x, d.air_y, d.air_z
        # box from node 3 to node n-4 in every direction (bounds land on nodes)
        d.hill_box = (x[3], x[-4], y[3], y[-4], z[1], z[-4])
        xc, yc = 0.5 * (x[0] + x[-1]), 0.5 * (y[0] + y[-1])
        if kind == 1:       # sphere: radius b about (c, d, e)
            d.hill_coef = coef(0.0, 0.22, xc + 0.01, yc - 0.02, 0.3)
        elif kind == 2:     # box [a,b] x [c,d] x [e,f]
            d.hill_coef = coef(xc - 0.12, xc + 0.13, yc - 0.2, yc + 0.1, 0.05, 0.33)
        elif kind == 3:     # three rows of cylinders, radius a, pitch b, `points` per row
            d.hill_coef = coef(0.06, 0.2, xc - 0.2, yc - 0.3, 0.3, xc, yc - 0.25, 0.4, xc + 0.2, yc - 0.28, 0.5)
            d.hill_points = 3
        elif kind == 4:     # g x f array of cylinders: radius a, origin (b, c), pitches d (y), e (x)
            d.hill_coef = coef(0.07, xc - 0.2, yc - 0.2, 0.2, 0.2, 3, 3)
        else:
            raise ValueError(kind)
    d.hill_judge = kind
    d.hill_mat = MATS
    return d
