"""Bounds the lane merging of the 27-cell deposit relies on.

The deterministic deposit sums the rounded terms of up to eight lanes in 32
bits (three butterfly steps).  That is exact only if every non-centre term
stays below 2^32 / 8.  A term is ((m * scale) * Kx) * Ky * Kz with
m * scale < 2^30, so the bound is on the weight products (below 1/2, with
the rounding of each term to an integer):

  * TSC: the weights of a particle at offset t (in cells) from its centre cell
    are 0.5 (0.5 - t)^2, 0.75 - t^2, 0.5 (0.5 + t)^2 (grids.f:1194-1213),
    and |t| <= 1/2 up to rounding for a particle inside the grid or inside a
    patch's BUF box (the index is the cell that contains it); the base grid
    accepts positions up to 2^-10 cells past the box, so |t| <= 1/2 + 2^-10
    there.  A non-centre product has at least one outer weight, so it is at
    most 0.5 * 0.75 * 0.75 = 0.28125.  The kernels evaluate the weights, as
    the reference does, from the three FP32 centres of the stencil, each
    rounded on its own; the plan refuses grids where their spacing may miss
    the cell size by more than a quarter of a cell, and within that every
    weight stays non-negative and every non-centre product below 1/2.
  * Table mode: narrow entries use row 0 of COMPUTE_KERNELS' table
    (grids.f:997-1067); the plan merges only when every non-centre product of
    that row is at most 0.45.

Checked here in FP32 on a dense sample of offsets, including a few ulps past
the ends of the interval, and on the kernel tables the build uses (oracle
restatement, itself pinned on the reference's table).
"""
import itertools

import numpy as np

from oracle import oracle as O

f32 = np.float32


def tsc_weights(t):
    t = np.asarray(t, f32)
    a = f32(0.5) - t
    b = f32(0.5) + t
    return np.stack([f32(0.5) * (a * a), f32(0.75) - t * t, f32(0.5) * (b * b)])


def max_non_centre_product(w_axis_max):
    """Largest product of three per-axis weights (each axis free) other than
    centre x centre x centre; w_axis_max[k] = max over the sample of weight k."""
    best = 0.0
    for k in itertools.product(range(3), repeat=3):
        if k == (1, 1, 1):
            continue
        best = max(best, float(w_axis_max[k[0]]) * float(w_axis_max[k[1]]) * float(w_axis_max[k[2]]))
    return best


def test_tsc_weight_products_fit_the_merge():
    half = f32(0.5)
    ends = []
    for e in (half, -half):
        v = e
        for _ in range(8):   # a few ulps beyond each end
            v = np.nextafter(v, f32(np.sign(e) * 2))
            ends.append(v)
    t = np.concatenate([np.linspace(-0.5, 0.5, 400001, dtype=f32), np.array(ends, f32)])
    w = tsc_weights(t)
    assert np.all(w >= 0)
    bound = max_non_centre_product(w.max(axis=1))
    assert bound <= 0.28125 * (1 + 1e-6), bound
    # the base grid accepts positions up to 2^-10 cells past the box
    # (kTscBaseSlack), i.e. offsets up to 1/2 + 2^-10 from the outer centres
    slack = f32(2.0 ** -10)
    tw = np.concatenate([t, np.linspace(0.5, 0.5 + slack, 1001, dtype=f32),
                         np.linspace(-0.5 - slack, -0.5, 1001, dtype=f32)])
    ww = tsc_weights(tw)
    assert np.all(ww >= 0)
    bound = max(bound, max_non_centre_product(ww.max(axis=1)))
    assert bound <= 0.2822, bound
    # the three centres rounded on their own: each distance off by up to a
    # quarter of a cell (the plan's limit)
    rng = np.random.default_rng(3)
    tt = rng.uniform(-0.5 - 2.0 ** -10, 0.5 + 2.0 ** -10, 200000)
    dd = rng.uniform(-0.25, 0.25, (3, 200000))
    b = np.stack([np.abs(1 + tt + dd[0]), np.abs(tt + dd[1]), np.abs(1 - tt + dd[2])]).astype(f32)
    wc = np.stack([f32(0.5) * (f32(1.5) - b[0]) ** 2, f32(0.75) - b[1] * b[1], f32(0.5) * (f32(1.5) - b[2]) ** 2])
    assert np.all(wc >= 0)
    # below 1/2: eight terms of a mass below 2^30 still sum below 2^32
    dev = max_non_centre_product(wc.max(axis=1))
    assert dev < 0.48
    for b_ in (bound, dev):
        q_max = np.floor(b_ * 2.0 ** 30) + 1   # a term rounded to the nearest integer
        assert 8 * q_max < 2 ** 32
    # the analytic worst case with the centre rounding: three outer weights of
    # a centre a quarter cell off, at the end of the accepted range
    w_out = 0.5 * (1.25 + 2.0 ** -10) ** 2
    assert dev <= w_out ** 3 < 0.48


def test_table_row0_products_pass_the_plan_check():
    for degree in (1, 2):
        K = O.kernels32(3, degree)
        kmax = 1 << 3
        k0 = K[0, kmax - 1:kmax + 2]
        assert np.all(k0 >= 0)
        prod = max(float(f32(f32(k0[a] * k0[b]) * k0[c]))
                   for a, b, c in itertools.product(range(3), repeat=3) if (a, b, c) != (1, 1, 1))
        # the plan's test, with room to spare
        assert prod <= 0.45
        assert prod < 0.08, (degree, prod)


def test_bound_catches_a_wrong_weight():
    # a weight set with a centre-sized outer weight breaks the bound (the
    # check is not vacuous)
    w = np.array([0.75, 0.75, 0.5], f32)
    assert max_non_centre_product(w) > 0.28125
