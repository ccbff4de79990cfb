"""Odd rounding used for the simulation-space dimensions.

Follows fdtdsarsim/utils/rounding.py:4-15: truncate toward zero, then bump an
even result up by one.
"""
import math


def round_to_odd(value):
    """Truncate ``value`` toward zero and return the next odd integer >= it.

    Raises ValueError for negative input, as the reference does.
    """
    if value < 0:
        raise ValueError('Only use this with positive values for now')
    out = int(math.trunc(float(value)))
    if out % 2 == 0:
        out += 1
    return out
