"""Ellipticity conversions (host side), Util/param_util.py:74-100 of the reference."""

import numpy as np


def phi_q2_ellipticity(phi, q):
    """param_util.py:74-84."""
    e1 = (1.0 - q) / (1.0 + q) * np.cos(2 * phi)
    e2 = (1.0 - q) / (1.0 + q) * np.sin(2 * phi)
    return e1, e2


def ellipticity2phi_q(e1, e2):
    """param_util.py:88-100."""
    phi = np.arctan2(e2, e1) / 2
    c = np.minimum(np.sqrt(e1 ** 2 + e2 ** 2), 0.9999)
    q = (1 - c) / (1 + c)
    return phi, q
