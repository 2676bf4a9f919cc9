"""Minimal physical units for the NEST-style API (``0.1 * u.ms``, ``20. * u.pA``).

The reference takes ``brainunit`` Quantities everywhere. ``brainunit`` is not
required here: plain numbers are read in NEST's default units (ms, mV, pA, pF,
nS, Hz), this module provides the same spelling for scripts written against the
reference, and genuine ``brainunit`` Quantities are converted when that package
is importable.
"""
from __future__ import annotations

import numpy as np

__all__ = ['Unit', 'Quantity', 'ms', 'second', 'mV', 'volt', 'pA', 'nA', 'pF', 'nF', 'nS',
           'uS', 'Hz', 'kHz', 'UNITLESS', 'magnitude', 'is_quantity']


class Unit:
    """A unit: a dimension name and the factor to that dimension's NEST unit."""
    __array_ufunc__ = None   # ndarray * unit -> Unit.__rmul__ (a Quantity)

    def __init__(self, name: str, dim: str, scale: float):
        self.name, self.dim, self.scale = name, dim, float(scale)

    def __rmul__(self, x):
        return Quantity(x, self)

    def __mul__(self, x):
        return Quantity(x, self)

    def __rtruediv__(self, x):
        if self.dim == 'time':      # 1 / second -> Hz
            return Quantity(np.asarray(x, dtype=float) / (self.scale * 1e-3), Hz)
        raise TypeError(f'cannot divide by unit {self.name}')

    def __repr__(self):
        return self.name


class Quantity:
    """value * unit; arithmetic kept to what network scripts use."""
    __array_ufunc__ = None

    def __init__(self, value, unit: Unit):
        self.value = value
        self.unit = unit

    @property
    def mantissa(self):
        return self.value

    def to_decimal(self, unit: Unit):
        if unit.dim != self.unit.dim:
            raise TypeError(f'cannot convert {self.unit.name} to {unit.name}')
        return np.asarray(self.value, dtype=float) * (self.unit.scale / unit.scale)

    def _same(self, other):
        if not isinstance(other, Quantity) or other.unit.dim != self.unit.dim:
            raise TypeError('unit mismatch')
        return other.to_decimal(self.unit)

    def __add__(self, other):
        return Quantity(np.asarray(self.value) + self._same(other), self.unit)

    def __sub__(self, other):
        return Quantity(np.asarray(self.value) - self._same(other), self.unit)

    def __neg__(self):
        return Quantity(-np.asarray(self.value), self.unit)

    def __mul__(self, x):
        if isinstance(x, (Quantity, Unit)):
            raise TypeError('products of quantities are not supported')
        return Quantity(np.asarray(self.value) * x, self.unit)

    __rmul__ = __mul__

    def __truediv__(self, x):
        if isinstance(x, Quantity):
            return self.to_decimal(self.unit) / x.to_decimal(self.unit) if x.unit.dim == self.unit.dim \
                else NotImplemented
        if isinstance(x, Unit):
            return self.to_decimal(x)
        return Quantity(np.asarray(self.value) / x, self.unit)

    def __getitem__(self, item):
        return Quantity(np.asarray(self.value)[item], self.unit)

    def __len__(self):
        return len(np.asarray(self.value))

    @property
    def shape(self):
        return np.shape(self.value)

    def __repr__(self):
        return f'{self.value} * {self.unit.name}'


UNITLESS = Unit('1', 'one', 1.0)
ms = Unit('ms', 'time', 1.0)
second = Unit('second', 'time', 1e3)
mV = Unit('mV', 'voltage', 1.0)
volt = Unit('volt', 'voltage', 1e3)
pA = Unit('pA', 'current', 1.0)
nA = Unit('nA', 'current', 1e3)
pF = Unit('pF', 'capacitance', 1.0)
nF = Unit('nF', 'capacitance', 1e3)
nS = Unit('nS', 'conductance', 1.0)
uS = Unit('uS', 'conductance', 1e3)
Hz = Unit('Hz', 'frequency', 1.0)
kHz = Unit('kHz', 'frequency', 1e3)

_NEST = {'time': ms, 'voltage': mV, 'current': pA, 'capacitance': pF, 'conductance': nS,
         'frequency': Hz, 'one': UNITLESS}
_BRAINUNIT_NAMES = {'time': 'ms', 'voltage': 'mV', 'current': 'pA', 'capacitance': 'pF',
                    'conductance': 'nS', 'frequency': 'Hz'}


def is_quantity(x) -> bool:
    return isinstance(x, Quantity) or (hasattr(x, 'to_decimal') and hasattr(x, 'unit'))


def magnitude(x, dim: str):
    """Magnitude of ``x`` in the NEST unit of ``dim`` (plain numbers pass through)."""
    if x is None:
        return None
    if isinstance(x, Quantity):
        return x.to_decimal(_NEST[dim])
    if hasattr(x, 'to_decimal') and hasattr(x, 'unit'):   # a brainunit Quantity
        import brainunit as bu  # noqa: F401  (only reachable when it is installed)
        return np.asarray(x.to_decimal(getattr(bu, _BRAINUNIT_NAMES[dim])), dtype=float)
    return np.asarray(x, dtype=float)


def scalar(x, dim: str) -> float:
    m = magnitude(x, dim)
    m = np.asarray(m, dtype=float)
    if m.size != 1:
        raise ValueError(f'expected a scalar {dim}, got shape {m.shape}')
    return float(m.reshape(()))
