"""Parameter distributions for per-edge weights and delays (``brainspike.dist``).

``sample(shape, seed)`` draws from a NumPy PCG64 stream; the reference draws
from ``jax.random`` (``_dist.py:22-80``) with the same formulas.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import units as u

__all__ = ['Distribution', 'Normal', 'LogNormal', 'Uniform', 'UniformGrid']


def _split(x):
    if isinstance(x, u.Quantity):
        return float(np.asarray(x.value)), x.unit
    return float(x), None


class Distribution:
    def sample(self, shape, seed):  # pragma: no cover - abstract
        raise NotImplementedError


@dataclass
class Normal(Distribution):
    mean: object
    std: object

    def sample(self, shape, seed):
        m, mu = _split(self.mean)
        s, su = _split(self.std)
        if (mu is None) != (su is None) or (mu is not None and mu.dim != su.dim):
            raise ValueError('mean and std must share units')
        if su is not None and mu is not None:
            s = float(u.Quantity(s, su).to_decimal(mu))
        x = m + s * np.random.Generator(np.random.PCG64(seed)).standard_normal(shape)
        return x if mu is None else u.Quantity(x, mu)


@dataclass
class LogNormal(Distribution):
    mean: float
    std: float

    def sample(self, shape, seed):
        return np.exp(self.mean + self.std *
                      np.random.Generator(np.random.PCG64(seed)).standard_normal(shape))


@dataclass
class Uniform(Distribution):
    low: object
    high: object

    def sample(self, shape, seed):
        lo, lu = _split(self.low)
        hi, hu = _split(self.high)
        if (lu is None) != (hu is None) or (lu is not None and lu.dim != hu.dim):
            raise ValueError('low and high must share units')
        if lu is not None:
            hi = float(u.Quantity(hi, hu).to_decimal(lu))
        x = lo + (hi - lo) * np.random.Generator(np.random.PCG64(seed)).random(shape)
        return x if lu is None else u.Quantity(x, lu)


@dataclass
class UniformGrid(Distribution):
    """Uniform over the grid ``low, low + step, ..., high`` (e.g. delays in whole
    steps of dt)."""
    low: float
    high: float
    step: float

    def sample(self, shape, seed):
        lo, hi, st = (float(u.scalar(x, 'time')) if isinstance(x, u.Quantity) else float(x)
                      for x in (self.low, self.high, self.step))
        k0, k1 = int(round(lo / st)), int(round(hi / st))
        k = np.random.Generator(np.random.PCG64(seed)).integers(k0, k1 + 1, size=shape)
        return k.astype(np.float64) * st
