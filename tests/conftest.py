import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def lebedev(order):
    """Angular rule from SciPy (an independent source of the Lebedev rules,
    used by the oracle side of the tests)."""
    import numpy as np
    from scipy.integrate import lebedev_rule
    x, w = lebedev_rule(order)
    return np.column_stack([x.T, w])


LEB_ORDERS = [3, 5, 7, 9, 11, 13, 15, 17, 19, 21, 23, 25, 27, 29, 31, 35, 41, 47, 53, 59, 65,
              71, 77, 83, 89, 95, 101, 107, 113, 119, 125, 131]


def next_order(order):
    """Smallest tabulated Lebedev order >= order (IntegratorXX's
    next_algebraic_order, called at grid.hpp:186-187)."""
    for o in LEB_ORDERS:
        if o >= order:
            return o
    raise ValueError(order)
