"""Lame parameters from Young's modulus and Poisson's ratio (ympr_to_lame.py:12-34)."""


def ympr_to_lame(ym, pr):
    """(mu, lam) = (E / (2 (1 + nu)), E nu / ((1 + nu)(1 - 2 nu)))."""
    mu = ym / (2 * (1 + pr))
    lam = ym * pr / ((1 + pr) * (1 - 2 * pr))
    return mu, lam
