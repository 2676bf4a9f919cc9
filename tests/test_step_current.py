"""step_current_generator through the current-input seam, checked with the
oracle against the reference's own step_current_generator tests
(step_current_generator_test.py:50-240, simulator_current_test.py:115-128): the
schedule look-up (zero before the first change, the latest plateau after,
start/stop/origin gating), the constructor's errors, the exact iaf_psc_delta
recurrence under a two-step schedule, the iaf_psc_alpha charge-up; plus
equivalences that pin the plateau lowering: a one-entry schedule is a
dc_generator bit for bit, and a schedule equals the sum of dc windows."""
import math

import numpy as np
import pytest

import oracle
from brainpy.state_amd import (Simulator, dc_generator, iaf_psc_alpha, iaf_psc_delta,
                               step_current_generator, units as u, voltmeter)
from brainpy.state_amd.init import Constant

NEST_DELTA = dict(E_L=-70.0, C_m=250.0, tau_m=10.0, V_th=-55.0, V_reset=-70.0, t_ref=2.0)


def _current_at(gen, t_ms, dt=0.1):
    n = int(round(t_ms / dt))
    return sum(a for a, n0, n1 in gen.plateaus(dt) if n0 <= n < n1)


def _trace(model, n_steps, gens=(), **params):
    sim = Simulator(dt=0.1 * u.ms)
    pop = sim.create(model, 1, **params)
    for cls, kw in gens:
        sim.connect(sim.create(cls, **kw), pop)
    vm = sim.create(voltmeter)
    sim.connect(vm, pop)
    v, _, _ = oracle.OracleNet(sim.build_desc()).run(n_steps)
    return np.asarray(v)[:, 0]


def test_single_step():
    g = step_current_generator(amplitude_times=[10. * u.ms], amplitude_values=[200. * u.pA])
    assert _current_at(g, 5.0) == 0.0
    assert _current_at(g, 10.0) == 200.0
    assert _current_at(g, 50.0) == 200.0


def test_multiple_steps():
    g = step_current_generator(amplitude_times=[10., 30., 50.] * u.ms,
                               amplitude_values=[100., -200., 300.] * u.pA)
    assert [_current_at(g, t) for t in (5., 20., 40., 60.)] == [0.0, 100.0, -200.0, 300.0]


def test_start_stop_gating_and_origin():
    g = step_current_generator(amplitude_times=[5.], amplitude_values=[200.], start=10., stop=30.)
    assert [_current_at(g, t) for t in (8., 15., 30.)] == [0.0, 200.0, 0.0]
    g = step_current_generator(amplitude_times=[5.], amplitude_values=[200.], start=10., stop=30.,
                               origin=100.)
    assert [_current_at(g, t) for t in (105., 115., 135.)] == [0.0, 200.0, 0.0]
    g = step_current_generator(amplitude_times=[5.], amplitude_values=[200.], start=30., stop=10.)
    assert g.plateaus(0.1) == []                         # stop <= start: never active


def test_plateaus_are_disjoint_and_cover_the_schedule():
    g = step_current_generator(amplitude_times=[0.3, 1.0, 2.5, 7.0], amplitude_values=[1., 2., 3., 4.],
                               start=0.7, stop=6.0)
    p = g.plateaus(0.1)
    assert [a for a, _, _ in p] == [1., 2., 3.]          # the 7 ms entry lies past stop
    assert [(n0, n1) for _, n0, n1 in p] == [(7, 10), (10, 25), (25, 60)]


def test_constructor_errors():
    with pytest.raises(AssertionError):
        step_current_generator()
    with pytest.raises(ValueError):
        step_current_generator(amplitude_times=[1., 2.], amplitude_values=[100.])
    with pytest.raises(ValueError):
        step_current_generator(amplitude_times=[1., 2., 2.], amplitude_values=[1., 2., 3.])


def test_delta_trace_matches_closed_form():
    """step_current_generator_test.py:178-241: 300 pA from 10 ms, -100 pA from
    30 ms, into a NEST-default iaf_psc_delta, against the exact recurrence."""
    n_steps, dt = 500, 0.1
    amp_times, amp_values = [10.0, 30.0], [300.0, -100.0]
    v = _trace(iaf_psc_delta, n_steps, [(step_current_generator,
                                         dict(amplitude_times=amp_times, amplitude_values=amp_values))],
               **NEST_DELTA)
    P33 = math.exp(-dt / 10.0)
    P30 = (10.0 / 250.0) * (1.0 - P33)
    V, ref, out = -70.0, 0, np.empty(n_steps)
    for step in range(n_steps):
        t = step * dt
        I = 0.0
        for tk, ak in zip(amp_times, amp_values):
            if t >= tk:
                I = ak
        if ref == 0:
            V = -70.0 + (V + 70.0) * P33 + I * P30
        else:
            ref -= 1
        if V >= -55.0 and ref == 0:
            ref, V = math.ceil(2.0 / dt), -70.0
        out[step] = V
    np.testing.assert_allclose(v, out, atol=1e-10, rtol=0)
    assert v[150] > -70.0 + 1.0 and v[-1] < -70.0      # charged by 300 pA, pulled below rest by -100 pA


def test_alpha_neuron_charges_after_the_change():
    """simulator_current_test.py:115-128: 0 pA until 10 ms, then 250 pA."""
    npar = dict(C_m=250., tau_m=10., tau_syn_ex=2., tau_syn_in=2., t_ref=2., E_L=-70., V_reset=-70.,
                V_th=-55., V_initializer=Constant(-70. * u.mV))
    v = _trace(iaf_psc_alpha, 600, [(step_current_generator,
                                     dict(amplitude_times=[10.] * u.ms, amplitude_values=[250.] * u.pA))],
               **npar)
    assert v[50] == pytest.approx(-70.0, abs=1e-6)
    assert v[-1] > -66.0


def test_one_entry_schedule_is_a_dc_generator():
    a = _trace(iaf_psc_delta, 1500, [(step_current_generator,
                                      dict(amplitude_times=[0.0], amplitude_values=[420.0], start=12.0,
                                           stop=97.3))], **NEST_DELTA)
    b = _trace(iaf_psc_delta, 1500, [(dc_generator, dict(amplitude=420.0, start=12.0, stop=97.3))],
               **NEST_DELTA)
    assert np.array_equal(a, b)


def test_schedule_equals_its_dc_windows():
    """A schedule [A0 from t0, A1 from t1, 0 from t2] drives the neuron exactly
    as dc windows [t0, t1) of A0 and [t1, t2) of A1."""
    a = _trace(iaf_psc_delta, 1200, [(step_current_generator,
                                      dict(amplitude_times=[5.0, 40.0, 80.0],
                                           amplitude_values=[330.0, 180.0, 0.0]))], **NEST_DELTA)
    b = _trace(iaf_psc_delta, 1200, [(dc_generator, dict(amplitude=330.0, start=5.0, stop=40.0)),
                                     (dc_generator, dict(amplitude=180.0, start=40.0, stop=80.0))],
               **NEST_DELTA)
    assert np.array_equal(a, b)
    assert np.any(a[50:400] != -70.0)


def test_long_schedule_closed_form():
    """A schedule of 1500 changes (one every 0.2 ms, values of both signs)
    into a NEST-default iaf_psc_delta: the oracle's one schedule channel
    against the exact recurrence with the reference's look-up (the plateau of
    the last amplitude_times[k] <= t, step_current_generator.py:269-335)."""
    n_steps, dt = 3200, 0.1
    rng = np.random.default_rng(4)
    amp_times = np.round(np.arange(1500) * 0.2 + 1.0, 1)
    amp_values = rng.uniform(-150.0, 450.0, 1500)
    gen = (step_current_generator, dict(amplitude_times=amp_times * u.ms, amplitude_values=amp_values * u.pA))
    v = _trace(iaf_psc_delta, n_steps, [gen], **NEST_DELTA)
    P33 = math.exp(-dt / 10.0)
    P30 = (10.0 / 250.0) * (1.0 - P33)
    V, ref, out = -70.0, 0, np.empty(n_steps)
    for step in range(n_steps):
        t = step * dt
        k = int(np.searchsorted(amp_times, t, side='right')) - 1
        I = amp_values[k] if k >= 0 else 0.0
        if ref == 0:
            V = -70.0 + (V + 70.0) * P33 + I * P30
        else:
            ref -= 1
        if V >= -55.0 and ref == 0:
            ref, V = math.ceil(2.0 / dt), -70.0
        out[step] = V
    np.testing.assert_allclose(v, out, atol=1e-9, rtol=0)
    assert np.ptp(v) > 5.0
