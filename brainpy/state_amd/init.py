"""State initialisers (``braintools.init`` equivalents used by the model specs)."""
from __future__ import annotations

import numpy as np

from . import units as u

__all__ = ['Constant', 'Uniform', 'Normal', 'resolve_init']


class Constant:
    """Every neuron starts at ``value``."""

    def __init__(self, value):
        self.value = value

    def __call__(self, n, dim):
        return np.full(n, float(np.asarray(u.magnitude(self.value, dim)).reshape(())))


class Uniform:
    """Independent uniform draws in ``[low, high)`` from a seeded NumPy PCG64 stream."""

    def __init__(self, low, high, seed=0):
        self.low, self.high, self.seed = low, high, int(seed)

    def __call__(self, n, dim):
        lo = float(u.magnitude(self.low, dim))
        hi = float(u.magnitude(self.high, dim))
        return lo + (hi - lo) * np.random.Generator(np.random.PCG64(self.seed)).random(n)


class Normal:
    """Independent normal draws from a seeded NumPy PCG64 stream."""

    def __init__(self, mean, std, seed=0):
        self.mean, self.std, self.seed = mean, std, int(seed)

    def __call__(self, n, dim):
        m = float(u.magnitude(self.mean, dim))
        s = float(u.magnitude(self.std, dim))
        return m + s * np.random.Generator(np.random.PCG64(self.seed)).standard_normal(n)


def resolve_init(init, n, dim):
    """Initial values of shape (n,) in the NEST unit of ``dim``."""
    if callable(init) and not u.is_quantity(init):
        try:
            out = init(n, dim)
        except TypeError:
            out = u.magnitude(init((n,)), dim)
    else:
        out = u.magnitude(init, dim)
    out = np.broadcast_to(np.asarray(out, dtype=np.float64), (n,)).copy()
    return out
