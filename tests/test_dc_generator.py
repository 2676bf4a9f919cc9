"""dc_generator through the current-input seam, checked with the oracle against
the reference's own dc_generator tests (dc_generator_test.py:296-560): the
same trace as I_e, the exact iaf_psc_delta recurrence inside start/stop/origin
windows, the steady state, additivity; and the one-step latch of the models
that buffer their current input (iaf_psc_exp i_0, iaf_psc_alpha y0)."""
import math

import numpy as np
import pytest

import oracle
from brainpy.state_amd import (Simulator, dc_generator, iaf_psc_alpha, iaf_psc_delta, iaf_psc_exp,
                               units as u, voltmeter)

NEST_DELTA = dict(E_L=-70.0, C_m=250.0, tau_m=10.0, V_th=-55.0, V_reset=-70.0, t_ref=2.0)


def _trace(model, n_steps, dcs=(), **params):
    sim = Simulator(dt=0.1 * u.ms)
    pop = sim.create(model, 1, **params)
    for kw in dcs:
        sim.connect(sim.create(dc_generator, **kw), pop)
    vm = sim.create(voltmeter)
    sim.connect(vm, pop)
    v, _, _ = oracle.OracleNet(sim.build_desc()).run(n_steps)
    return np.asarray(v)[:, 0]


def _analytical(n_steps, windows, E_L=-70.0, C_m=250.0, tau_m=10.0, V_th=-55.0, V_reset=-70.0,
                t_ref=2.0, dt=0.1):
    """dc_generator_test.py:_analytical_vm_trace with several (amp, t0, t1) windows."""
    P33 = math.exp(-dt / tau_m)
    P30 = (tau_m / C_m) * (1.0 - P33)
    refr = math.ceil(t_ref / dt)
    V, ref, out = E_L, 0, np.empty(n_steps)
    for step in range(n_steps):
        t = step * dt
        I = sum(a for a, t0, t1 in windows if t0 <= t < t1)
        if ref == 0:
            V = E_L + (V - E_L) * P33 + I * P30
        else:
            ref -= 1
        if V >= V_th and ref == 0:
            ref, V = refr, V_reset
        out[step] = V
    return out


@pytest.mark.parametrize('amp', [300.0, 500.0])
def test_dc_equals_I_e(amp):
    a = _trace(iaf_psc_delta, 2000, [dict(amplitude=amp * u.pA)], **NEST_DELTA)
    b = _trace(iaf_psc_delta, 2000, I_e=amp, **NEST_DELTA)
    assert np.array_equal(a, b)


@pytest.mark.parametrize('amp,start,stop,origin', [(300.0, 0.0, None, 0.0), (500.0, 0.0, None, 0.0),
                                                  (300.0, 20.0, 60.0, 0.0), (300.0, 10.0, 50.0, 30.0),
                                                  (300.0, 60.0, 20.0, 0.0)])
def test_delta_trace_matches_closed_form(amp, start, stop, origin):
    n = 1000
    kw = dict(amplitude=amp, start=start, origin=origin)
    if stop is not None:
        kw['stop'] = stop
    v = _trace(iaf_psc_delta, n, [kw], **NEST_DELTA)
    t1 = origin + stop if stop is not None else math.inf
    ref = _analytical(n, [(amp, origin + start, t1)])
    np.testing.assert_allclose(v, ref, atol=1e-10, rtol=0)


def test_steady_state_and_return_to_rest():
    v = _trace(iaf_psc_delta, 2000, [dict(amplitude=300.0)], **NEST_DELTA)
    assert v[-1] == pytest.approx(-70.0 + 300.0 * 10.0 / 250.0, abs=1e-6)
    v = _trace(iaf_psc_delta, 2000, [dict(amplitude=300.0, stop=50.0)], **NEST_DELTA)
    assert v[-1] == pytest.approx(-70.0, abs=1e-5)


def test_two_generators_add():
    v = _trace(iaf_psc_delta, 1000, [dict(amplitude=200.0, start=0.0, stop=50.0),
                                     dict(amplitude=100.0, start=25.0, stop=75.0)], **NEST_DELTA)
    ref = _analytical(1000, [(200.0, 0.0, 50.0), (100.0, 25.0, 75.0)])
    np.testing.assert_allclose(v, ref, atol=1e-10, rtol=0)


@pytest.mark.parametrize('model', [iaf_psc_exp, iaf_psc_alpha])
def test_latching_models_see_the_current_one_step_later(model):
    """i_0 / y0 hold this step's current for the next step (iaf_psc_exp.py:603-672,
    iaf_psc_alpha.py:597-677): from rest, the dc trace is the I_e trace one
    step later, bit for bit."""
    a = _trace(model, 1500, [dict(amplitude=450.0)])
    b = _trace(model, 1500, I_e=450.0)
    assert a[0] == _trace(model, 1)[0]                  # no current in the first step
    assert np.array_equal(a[1:], b[:-1])
    assert np.any(a != a[0])


def test_step_window_semantics():
    g = dc_generator(amplitude=1.0, start=2.0, stop=5.0, origin=1.0)
    assert g.step_window(0.1) == (30, 60)                # [3.0 ms, 6.0 ms)
    assert dc_generator(start=0.3).step_window(0.1)[0] == 3
    n0, n1 = dc_generator(start=5.0, stop=2.0).step_window(0.1)
    assert n1 <= n0                                      # stop <= start: never active
    assert dc_generator().step_window(0.1)[1] == np.iinfo(np.int64).max
