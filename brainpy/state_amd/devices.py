"""NEST devices on the network-step path: Poisson drive and recorders.

* ``poisson_generator`` (``_nest_device/poisson_generator.py:223-576``): a
  deferred spec realised per target segment at ``connect``; each element is an
  independent Poisson train of mean ``rate * dt / 1000`` per step, active iff
  ``t_min_step < step <= t_max_step``. Draws come from the engine's
  counter-based stream (Threefry-2x32 keyed by the derived seed and the step).
* ``dc_generator`` (``_nest_device/dc_generator.py:31-262``): a constant
  current (pA) inside the half-open window ``[origin + start, origin + stop)``,
  injected into the neuron's current-input seam (``simulator.py:1612-1745``).
* ``spike_recorder`` / ``voltmeter`` / ``multimeter``: tapped by the Simulator
  every step after the update (``simulator.py:1858-1867``).
"""
from __future__ import annotations

import math

import numpy as np

from . import units as u

__all__ = ['poisson_generator', 'dc_generator', 'spike_recorder', 'voltmeter', 'multimeter']


class poisson_generator:
    """Poisson spike-train generator (NEST ``poisson_generator``)."""
    __module__ = 'brainpy.state_amd'
    _is_generator = True

    def __init__(self, in_size=1, rate=0.0, start=0.0, stop=None, origin=0.0, rng_seed=0,
                 name=None):
        self.in_size = int(in_size)
        self.rate = u.scalar(rate, 'frequency')
        self.start = u.scalar(start, 'time')
        self.stop = np.inf if stop is None else u.scalar(stop, 'time')
        self.origin = u.scalar(origin, 'time')
        self.rng_seed = int(rng_seed)
        self.name = name
        if self.rate < 0.0:
            raise ValueError('The rate cannot be negative.')
        if self.stop < self.start:
            raise ValueError('stop >= start required.')

    @staticmethod
    def _time_to_step(time_ms, dt_ms):
        return int(np.rint(time_ms / dt_ms))

    @staticmethod
    def _assert_grid_time(name, time_ms, dt_ms):
        if not np.isfinite(time_ms):
            return
        ratio = time_ms / dt_ms
        if not math.isclose(ratio, np.rint(ratio), rel_tol=0.0, abs_tol=1e-12):
            raise ValueError(f'{name} must be a multiple of the simulation resolution.')

    def step_window(self, dt_ms):
        """(t_min_step, t_max_step): active iff t_min_step < step <= t_max_step."""
        for nm, v in (('origin', self.origin), ('start', self.start), ('stop', self.stop)):
            self._assert_grid_time(nm, v, dt_ms)
        t_min = self._time_to_step(self.origin + self.start, dt_ms)
        t_max = (self._time_to_step(self.origin + self.stop, dt_ms) if np.isfinite(self.stop)
                 else np.iinfo(np.int64).max)
        return t_min, t_max

    def lam(self, dt_ms):
        return self.rate * dt_ms / 1000.0


class dc_generator:
    """DC current generator (NEST ``dc_generator``): ``amplitude`` (pA) while
    ``origin + start <= t < origin + stop`` (``stop=None``: no end), else 0,
    with ``t`` the step's time ``n * dt``."""
    __module__ = 'brainpy.state_amd'
    _is_generator = True
    _injects_current = True

    def __init__(self, in_size=1, amplitude=0.0, start=0.0, stop=None, origin=0.0, name=None):
        self.in_size = int(in_size)
        self.amplitude = u.scalar(amplitude, 'current')
        self.start = u.scalar(start, 'time')
        self.stop = None if stop is None else u.scalar(stop, 'time')
        self.origin = u.scalar(origin, 'time')
        self.name = name

    def step_window(self, dt_ms):
        """(n_start, n_stop): active at steps n_start <= n < n_stop, i.e. where
        ``n * dt >= origin + start`` and ``n * dt < origin + stop``, both
        compared in float64 as the reference compares ``t`` (dc_generator.py:252-262)."""
        def first_at_or_after(t):
            n = max(0, int(math.floor(t / dt_ms)) - 2)
            while float(n) * dt_ms < t:
                n += 1
            return n
        t_start = self.origin + self.start
        n_start = first_at_or_after(t_start)
        if self.stop is None:
            return n_start, np.iinfo(np.int64).max
        n_stop = first_at_or_after(self.origin + self.stop)
        return n_start, max(n_stop, n_start)     # stop <= start: never active


class spike_recorder:
    """Records the binarised spikes of a population segment every step."""
    __module__ = 'brainpy.state_amd'

    def __init__(self, in_size=1, start=0.0, stop=None, origin=0.0, time_in_steps=False,
                 frozen=False, name=None):
        if frozen:
            raise ValueError('spike_recorder cannot be frozen.')
        self.in_size = 1
        self.start, self.stop, self.origin = start, stop, origin
        self.name = name


class multimeter:
    """Records named state variables of a population segment every step."""
    __module__ = 'brainpy.state_amd'

    def __init__(self, in_size=1, record_from=('V_m',), interval=1.0, offset=0.0, start=0.0,
                 stop=None, origin=0.0, time_in_steps=False, frozen=False, name=None):
        self.in_size = 1
        self.record_from = tuple(record_from)
        self.interval = interval
        self.name = name


class voltmeter(multimeter):
    """A multimeter preset to ``record_from=('V_m',)``."""
    __module__ = 'brainpy.state_amd'

    def __init__(self, in_size=1, interval=1.0, offset=0.0, start=0.0, stop=None, origin=0.0,
                 time_in_steps=False, frozen=False, name=None):
        super().__init__(in_size, record_from=('V_m',), interval=interval, offset=offset,
                         start=start, stop=stop, origin=origin, time_in_steps=time_in_steps,
                         frozen=frozen, name=name)
