"""Shared reporting for per-path parity tests: how many pixels of a GPU film
differ from the oracle's beyond a relative tolerance (paths whose branch a
last-bit libm difference flips)."""
import inspect

import numpy as np


def diverging(per_px, tol):
    """Print and return the number of pixels whose relative difference is >= tol."""
    n = int((np.asarray(per_px) >= tol).sum())
    caller = inspect.stack()[1].function
    print(f"DIVERGING {caller}: {n} of {np.asarray(per_px).size} pixels differ by >= {tol:g}", flush=True)
    return n
