"""Observables computed from per-image energies (post-processing of the
reference, mepgl/post.py)."""
from __future__ import annotations

import numpy as np


def barriers(fenergy):
    """Energy barriers of a string, post.py:188-189: the entry barrier
    max(F - F_0) and the exit barrier max(F - F_last)."""
    e = np.asarray(fenergy, dtype=np.float64)
    return float(np.max(e - e[0])), float(np.max(e - e[-1]))
