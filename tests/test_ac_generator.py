"""ac_generator through the current-input seam, checked with the oracle:
the engine's sine against the C library's (a few ulp over a wide argument
range), the reference's ac tests (ac_generator_test.py:85-254: zero/peak/zero
at t = 0, 1, 2 ms for 250 Hz, a 90 degree phase, the windowed 550 pA + 1000 pA
trace against a manual sine) through the exact iaf_psc_delta recurrence, and
equivalences with dc_generator (frequency 0 at 90 degrees, offset only) bit
for bit."""
import math

import numpy as np
import pytest

import oracle
from brainpy.state_amd import (Simulator, ac_generator, dc_generator, iaf_psc_delta, units as u,
                               voltmeter)

NEST_DELTA = dict(E_L=-70.0, C_m=250.0, tau_m=10.0, V_th=1e9, V_reset=-70.0, t_ref=2.0)


def test_sine_matches_the_c_library():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-1e4, 1e4, 20000), rng.uniform(-4, 4, 20000),
                         rng.uniform(-1e-6, 1e-6, 200), np.arange(-64, 65) * (math.pi / 4),
                         [1e5 + 0.3, -7.5e5, 1.5e6]])
    worst = 0.0
    for x in xs:
        got, want = oracle.sin(x), math.sin(x)
        worst = max(worst, abs(got - want) / np.spacing(max(abs(want), 1e-300)))
    assert worst <= 4.0, f'{worst} ulp'
    assert oracle.sin(0.0) == 0.0 and oracle.sin(-0.0) == 0.0   # the sign of a zero is not kept
    assert oracle.sin(1e-20) == 1e-20
    assert math.isnan(oracle.sin(math.inf)) and math.isnan(oracle.sin(math.nan))


def _trace(n_steps, gens, **params):
    sim = Simulator(dt=0.1 * u.ms)
    pop = sim.create(iaf_psc_delta, 1, **(params or NEST_DELTA))
    for cls, kw in gens:
        sim.connect(sim.create(cls, **kw), pop)
    vm = sim.create(voltmeter)
    sim.connect(vm, pop)
    v, _, _ = oracle.OracleNet(sim.build_desc()).run(n_steps)
    return np.asarray(v)[:, 0]


def _analytical(current, E_L=-70.0, C_m=250.0, tau_m=10.0, dt=0.1):
    """iaf_psc_delta without threshold under the current I[n] of step n."""
    P33 = math.exp(-dt / tau_m)
    P30 = (tau_m / C_m) * (1.0 - P33)
    V, out = E_L, np.empty(len(current))
    for n, I in enumerate(current):
        V = E_L + (V - E_L) * P33 + I * P30
        out[n] = V
    return out


def _manual(n_steps, amp, offset, freq, phase_deg, start, stop, dt=0.1):
    t = np.arange(n_steps) * dt
    active = (t >= start) & (t < stop)
    return np.where(active, np.sin(t * (freq / 1000.0 * 2 * math.pi) + phase_deg * 2 * math.pi / 360.0)
                    * amp + offset, 0.0)


@pytest.mark.parametrize('amp,offset,freq,phase,start,stop', [
    (550.0, 1000.0, 100.0, 0.0, 5.0, 6.0),          # ac_generator_test.py:207-254
    (100.0, 0.0, 250.0, 0.0, 0.0, math.inf),        # ac_generator_test.py:85-109
    (100.0, 0.0, 250.0, 90.0, 0.0, math.inf),       # ac_generator_test.py:111-123
    (300.0, 50.0, 37.0, 33.0, 12.5, 71.2),
])
def test_trace_matches_manual_sine(amp, offset, freq, phase, start, stop):
    n = 1000
    kw = dict(amplitude=amp * u.pA, offset=offset * u.pA, frequency=freq * u.Hz, phase=phase,
              start=start * u.ms)
    if math.isfinite(stop):
        kw['stop'] = stop * u.ms
    v = _trace(n, [(ac_generator, kw)])
    I = _manual(n, amp, offset, freq, phase, start, stop)
    np.testing.assert_allclose(v, _analytical(I), atol=1e-10, rtol=0)


def test_known_times_through_the_membrane():
    """Recover the current from the exact recurrence: 250 Hz, 100 pA gives 0,
    100 and 0 pA at t = 0, 1 and 2 ms."""
    v = _trace(30, [(ac_generator, dict(amplitude=100.0, frequency=250.0))])
    P33 = math.exp(-0.1 / 10.0)
    P30 = (10.0 / 250.0) * (1.0 - P33)
    prev = np.concatenate([[-70.0], v[:-1]])
    I = (v - (-70.0) - (prev + 70.0) * P33) / P30
    assert abs(I[0]) < 1e-9 and abs(I[10] - 100.0) < 1e-9 and abs(I[20]) < 1e-9


def test_zero_frequency_at_ninety_degrees_is_a_dc_generator():
    a = _trace(800, [(ac_generator, dict(amplitude=300.0, offset=100.0, phase=90.0, start=10.0,
                                         stop=55.5))])
    b = _trace(800, [(dc_generator, dict(amplitude=400.0, start=10.0, stop=55.5))])
    assert np.array_equal(a, b)
    a = _trace(800, [(ac_generator, dict(offset=250.0, frequency=40.0, origin=3.0))])
    b = _trace(800, [(dc_generator, dict(amplitude=250.0, origin=3.0))])
    assert np.array_equal(a, b)
