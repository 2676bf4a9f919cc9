"""NEST devices on the network-step path: Poisson drive and recorders.

* ``poisson_generator`` (``_nest_device/poisson_generator.py:223-576``): a
  deferred spec realised per target segment at ``connect``; each element is an
  independent Poisson train of mean ``rate * dt / 1000`` per step, active iff
  ``t_min_step < step <= t_max_step``. Draws come from the engine's
  counter-based stream (Threefry-2x32 keyed by the derived seed and the step).
* ``dc_generator`` (``_nest_device/dc_generator.py:31-262``): a constant
  current (pA) inside the half-open window ``[origin + start, origin + stop)``,
  injected into the neuron's current-input seam (``simulator.py:1612-1745``).
* ``step_current_generator`` (``_nest_device/step_current_generator.py:231-335``):
  a piecewise-constant current, the latest plateau whose change time is
  ``<= t``, gated by the same half-open window; realised as one dc plateau
  per schedule entry (at most one is non-zero at any step).
* ``ac_generator`` (``_nest_device/ac_generator.py:236-325``): ``amplitude *
  sin(2 pi frequency t + phase) + offset`` inside the same window, evaluated
  on the device each step.
* ``spike_recorder`` / ``voltmeter`` / ``multimeter``: tapped by the Simulator
  every step after the update (``simulator.py:1858-1867``).
"""
from __future__ import annotations

import math

import numpy as np

from . import units as u

__all__ = ['poisson_generator', 'dc_generator', 'step_current_generator', 'ac_generator', 'spike_recorder',
           'voltmeter', 'multimeter']


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
        t_start = self.origin + self.start
        n_start = _first_step_at_or_after(t_start, dt_ms)
        if self.stop is None:
            return n_start, np.iinfo(np.int64).max
        n_stop = _first_step_at_or_after(self.origin + self.stop, dt_ms)
        return n_start, max(n_stop, n_start)     # stop <= start: never active

    def plateaus(self, dt_ms):
        """[(amplitude, n_start, n_stop)]: the current as dc plateaus."""
        return [(self.amplitude,) + self.step_window(dt_ms)]


def _first_step_at_or_after(t, dt_ms):
    """The first step n with ``n * dt >= t`` (float64 product, as the step's time)."""
    if t == math.inf:
        return np.iinfo(np.int64).max
    n = max(0, int(math.floor(t / dt_ms)) - 2)
    while float(n) * dt_ms < t:
        n += 1
    return n


def _schedule(x, dim):
    if isinstance(x, u.Quantity) or isinstance(x, np.ndarray):
        return [float(v) for v in np.ravel(u.magnitude(x, dim))]
    return [u.scalar(v, dim) for v in x]


class step_current_generator:
    """Piecewise-constant current generator (NEST ``step_current_generator``):
    ``amplitude_values[k]`` (pA) from ``amplitude_times[k]`` (ms) until the next
    change time, 0 before the first, gated by ``origin + start <= t < origin +
    stop`` (``step_current_generator.py:269-335``: the plateau is the last
    ``k`` with ``amplitude_times[k] <= t``, compared in float64)."""
    __module__ = 'brainpy.state_amd'
    _is_generator = True
    _injects_current = True

    def __init__(self, in_size=1, amplitude_times=(), amplitude_values=(), start=0.0, stop=None,
                 origin=0.0, name=None):
        self.in_size = int(in_size)
        times = _schedule(amplitude_times, 'time')
        values = _schedule(amplitude_values, 'current')
        if len(times) != len(values):
            raise ValueError('amplitude_times and amplitude_values must have the same length. '
                             f'Got {len(times)} and {len(values)}.')
        if len(times) == 0:
            raise AssertionError('At least one schedule entry is required. Got len(amplitude_times) = 0.')
        for i in range(1, len(times)):
            if times[i] <= times[i - 1]:
                raise ValueError('amplitude_times must be strictly increasing. '
                                 f'Got {times[i - 1]} >= {times[i]} at index {i}.')
        self.amplitude_times = times
        self.amplitude_values = values
        self.start = u.scalar(start, 'time')
        self.stop = None if stop is None else u.scalar(stop, 'time')
        self.origin = u.scalar(origin, 'time')
        self.name = name

    def plateaus(self, dt_ms):
        """[(amplitude, n_start, n_stop)]: plateau k is on at steps n_start <= n
        < n_stop, the steps where ``amplitude_times[k] <= n * dt <
        amplitude_times[k + 1]`` inside the activity window. The plateaus are
        disjoint, so their sum is the scheduled current bit for bit; empty
        ones are dropped."""
        w0 = _first_step_at_or_after(self.origin + self.start, dt_ms)
        w1 = (np.iinfo(np.int64).max if self.stop is None
              else _first_step_at_or_after(self.origin + self.stop, dt_ms))
        edges = [_first_step_at_or_after(t, dt_ms) for t in self.amplitude_times]
        edges.append(np.iinfo(np.int64).max)
        out = []
        for k, amp in enumerate(self.amplitude_values):
            n0, n1 = max(edges[k], w0), min(edges[k + 1], w1)
            if n0 < n1:
                out.append((amp, n0, n1))
        return out


class ac_generator:
    """Sinusoidal current generator (NEST ``ac_generator``): ``amplitude *
    sin(omega t + phi) + offset`` (pA) while ``origin + start <= t < origin +
    stop``, else 0, with ``omega = 2 pi frequency / 1000`` rad/ms, ``phi`` the
    phase in degrees converted to radians and ``t = n * dt``
    (``ac_generator.py:264-325``)."""
    __module__ = 'brainpy.state_amd'
    _is_generator = True
    _injects_current = True

    def __init__(self, in_size=1, amplitude=0.0, offset=0.0, frequency=0.0, phase=0.0, start=0.0,
                 stop=None, origin=0.0, name=None):
        self.in_size = int(in_size)
        self.amplitude = u.scalar(amplitude, 'current')
        self.offset = u.scalar(offset, 'current')
        self.frequency = u.scalar(frequency, 'frequency')
        self.phase = float(phase)
        self.start = u.scalar(start, 'time')
        self.stop = None if stop is None else u.scalar(stop, 'time')
        self.origin = u.scalar(origin, 'time')
        self.name = name

    def step_window(self, dt_ms):
        """(n_start, n_stop), as for dc_generator."""
        return dc_generator.step_window(self, dt_ms)

    def ac_params(self):
        """(amplitude, omega in rad/ms, phase in rad, offset)."""
        return (self.amplitude, 2.0 * math.pi * self.frequency / 1000.0,
                self.phase * 2.0 * math.pi / 360.0, self.offset)

    def plateaus(self, dt_ms):
        """One channel: the window, with the sinusoid's parameters."""
        return [('ac',) + self.step_window(dt_ms)]


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
