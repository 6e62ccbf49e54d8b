"""The warm start's extrapolation weights (simkit_amd.learned.EXTRAP): order p
takes sum_j w_j e_{n-j} as the value one step ahead of the degree-p
polynomial through e_n, ..., e_{n-p}. Pinned by what fixes them: every
polynomial of degree <= p is reproduced exactly, and one of degree p + 1 is
not (the weights are then unique, so a wrong sign or index fails)."""

import numpy as np

from simkit_amd.learned import EXTRAP


def test_extrapolation_weights_reproduce_polynomials_of_their_degree():
    rng = np.random.default_rng(0)
    for p, w in EXTRAP.items():
        assert len(w) == p + 1
        t = -np.arange(p + 1, dtype=np.float64)  # e_n at t = 0, e_{n-1} at t = -1, ...
        for deg in range(p + 2):
            c = rng.standard_normal(deg + 1)
            vals = np.polyval(c, t)
            ahead = float(np.dot(w, vals))
            exact = float(np.polyval(c, 1.0))
            if deg <= p:
                assert abs(ahead - exact) <= 1e-12 * max(1.0, np.abs(c).sum()), (p, deg)
            else:
                assert abs(ahead - exact) > 1e-6, (p, deg)


def test_extrapolation_orders_and_weight_sums():
    assert sorted(EXTRAP) == [1, 2, 3]
    for w in EXTRAP.values():
        assert abs(sum(w) - 1.0) < 1e-15  # constants are kept
