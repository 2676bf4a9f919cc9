"""GPU: dc_generator, step_current_generator and ac_generator current inputs
(full populations, subsets, start/stop/origin windows, schedules, sinusoids,
scaled by connection weights) on all four models, mixed with Poisson drive and recurrent
projections: spikes and V_m bit-exact against the oracle from the same
descriptor."""
import numpy as np
import pytest

import oracle
from brainpy.state_amd import (Simulator, ac_generator, aeif_cond_alpha, dc_generator, fixed_indegree,
                               iaf_psc_alpha, iaf_psc_delta, iaf_psc_exp, poisson_generator,
                               spike_recorder, step_current_generator, units as u, voltmeter)
from brainpy.state_amd.init import Uniform

pytestmark = pytest.mark.gpu


def _net(model, dtype='float64', schedules=False):
    sim = Simulator(dt=0.1 * u.ms, dtype=dtype)
    cond = model is aeif_cond_alpha
    kw = {} if cond else dict(V_initializer=Uniform(-70., -60., seed=5))
    if model is iaf_psc_delta:
        kw.update(E_L=-70.0, V_th=-55.0, V_reset=-70.0, tau_m=10.0)
    a = sim.create(model, 300, **kw)
    b = sim.create(model, 80, **kw)
    rt = dict(receptor_type=1) if cond else {}
    rti = dict(receptor_type=2) if cond else {}
    wE, wI = (2.0, 4.0) if cond else ((0.3, -1.5) if model is iaf_psc_delta else (40.0, -160.0))
    sim.connect(sim.create(poisson_generator, rate=6000. * u.Hz), a + b, weight=wE, delay=1.0, **rt)
    # dc drive: a whole population, a subset with a window, a scaled one with an origin
    sim.connect(sim.create(dc_generator, amplitude=320.0 * u.pA), a)
    sim.connect(sim.create(dc_generator, amplitude=450.0, start=20.0, stop=80.0), b[:40])
    sim.connect(sim.create(dc_generator, amplitude=100.0, start=10.0, stop=50.0, origin=40.0), a[100:],
                weight=1.5)
    if schedules:
        # step drive: a schedule on a whole population (negative plateau too),
        # a gated one on a subset, a scaled one with an origin
        sim.connect(sim.create(step_current_generator, amplitude_times=[0.0, 30.0, 60.0, 95.0],
                               amplitude_values=[150.0, -200.0, 400.0, 0.0]), a)
        sim.connect(sim.create(step_current_generator, amplitude_times=[12.3, 47.0],
                               amplitude_values=[500.0, 250.0], start=15.0, stop=120.0), b[10:70])
        sim.connect(sim.create(step_current_generator, amplitude_times=[5.0, 25.0],
                               amplitude_values=[300.0, 120.0], origin=20.0), a[:150], weight=0.8)
        # sinusoids: a whole population, and a gated, phase-shifted, scaled subset
        sim.connect(sim.create(ac_generator, amplitude=180.0, offset=60.0, frequency=35.0, phase=20.0), b)
        sim.connect(sim.create(ac_generator, amplitude=250.0, frequency=110.0, phase=-45.0, start=8.0,
                               stop=130.0), a[50:250], weight=1.3)
    sim.connect(a, a + b, weight=wE, delay=1.0, rule=fixed_indegree(20), seed=4, **rt)
    sim.connect(b, a + b, weight=wI, delay=2.0, rule=fixed_indegree(8), seed=5, **rti)
    for pop in (a, b):
        sim.connect(pop, sim.create(spike_recorder))
    vm = sim.create(voltmeter)
    sim.connect(vm, a[:6])
    vm2 = sim.create(voltmeter)
    sim.connect(vm2, b[30:36])
    return sim


@pytest.mark.parametrize('model', [iaf_psc_delta, iaf_psc_exp, iaf_psc_alpha, aeif_cond_alpha])
def test_dc_current_network_bit_exact(model):
    sim = _net(model)
    desc = sim.build_desc()
    eng = sim._ensure_engine()
    eng.reset()
    n = 1500
    vm_g, (g_g, s_g) = eng.run(n)
    vm_o, g_o, s_o = oracle.OracleNet(desc).run(n)
    assert len(g_o) > 20, 'network too quiet for a useful check'
    assert np.array_equal(g_g, g_o) and np.array_equal(s_g, s_o)
    np.testing.assert_allclose(vm_g, vm_o, rtol=1e-12, atol=0)


def test_dc_current_float32_runs_against_float32_oracle():
    sim = _net(iaf_psc_exp, dtype='float32')
    desc = sim.build_desc()
    eng = sim._ensure_engine()
    eng.reset()
    vm_g, (g_g, s_g) = eng.run(1000)
    vm_o, g_o, s_o = oracle.OracleNet(desc).run(1000)
    assert np.array_equal(g_g, g_o) and np.array_equal(s_g, s_o)


@pytest.mark.parametrize('model', [iaf_psc_delta, iaf_psc_exp, iaf_psc_alpha, aeif_cond_alpha])
def test_step_current_network_bit_exact(model):
    sim = _net(model, schedules=True)
    desc = sim.build_desc()
    eng = sim._ensure_engine()
    eng.reset()
    n = 1500
    vm_g, (g_g, s_g) = eng.run(n)
    vm_o, g_o, s_o = oracle.OracleNet(desc).run(n)
    assert len(g_o) > 20, 'network too quiet for a useful check'
    assert np.array_equal(g_g, g_o) and np.array_equal(s_g, s_o)
    np.testing.assert_allclose(vm_g, vm_o, rtol=1e-12, atol=0)


def test_step_and_ac_currents_float32_against_float32_oracle():
    sim = _net(iaf_psc_alpha, dtype='float32', schedules=True)
    desc = sim.build_desc()
    eng = sim._ensure_engine()
    eng.reset()
    vm_g, (g_g, s_g) = eng.run(1000)
    vm_o, g_o, s_o = oracle.OracleNet(desc).run(1000)
    assert len(g_o) > 20
    assert np.array_equal(g_g, g_o) and np.array_equal(s_g, s_o)
    np.testing.assert_allclose(vm_g, vm_o, rtol=1e-6, atol=1e-4)


@pytest.mark.parametrize('model', [iaf_psc_delta, iaf_psc_exp, aeif_cond_alpha])
def test_long_schedules_bit_exact(model):
    """step_current_generator schedules of 2000 changes (one channel each, any
    length: the device looks the plateau up once per window step), on a whole
    population and on a weighted subset."""
    from brainpy.state_amd import NodeView
    sim = _net(model)
    pops = [NodeView.of(p) for p in sim._pops]
    rng = np.random.default_rng(9)
    t = np.round(np.arange(2000) * 0.1 + 0.3, 1)
    sim.connect(sim.create(step_current_generator, amplitude_times=t,
                           amplitude_values=rng.uniform(-100.0, 400.0, 2000)), pops[0])
    sim.connect(sim.create(step_current_generator, amplitude_times=t[::2] + 0.05,
                           amplitude_values=rng.uniform(0.0, 600.0, 1000), start=3.0, stop=150.0),
                pops[1], weight=0.7)
    desc = sim.build_desc()
    assert sum(desc.c.channels[k].kind == 6 for k in range(desc.c.n_channels)) == 2
    eng = sim._ensure_engine()
    eng.reset()
    n = 2500
    vm_g, (g_g, s_g) = eng.run(n)
    vm_o, g_o, s_o = oracle.OracleNet(desc).run(n)
    assert len(g_o) > 20, 'network too quiet for a useful check'
    assert np.array_equal(g_g, g_o) and np.array_equal(s_g, s_o)
    np.testing.assert_allclose(vm_g, vm_o, rtol=1e-12, atol=0)


def test_ac_run_beyond_the_exact_sine_range_is_refused():
    """The engine's sine is exact for |argument| < 1.6e6 rad; a run whose
    ac_generator argument would pass that is refused before any launch."""
    from brainpy.state_amd._lib import EngineError
    sim = Simulator(dt=0.1 * u.ms)
    p = sim.create(iaf_psc_delta, 4)
    sim.connect(sim.create(ac_generator, amplitude=10.0, frequency=1000.0), p)
    eng = sim._ensure_engine()
    eng.reset()
    eng.run(100)                          # 2 pi rad/ms * 10 ms: fine
    with pytest.raises(EngineError, match='1.6e6'):
        eng.run(3_000_000)                # 300 s at 2 pi rad/ms: 1.9e6 rad
    eng.close()
    sim._engine = None
