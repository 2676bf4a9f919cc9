"""GPU parity of fractional InputDelay reads (delay.py:114-128) against the oracle.

A homogeneous or axonal delay whose ratio to dt has a fraction in the network
dtype delivers from two frames: both as events under comm='sparse', as the
graded parts w (1 - f) and w f under comm='dense' and receptor scatter; a
Poisson generator's counts are mixed the same way. The cases include ratios
that are whole in exact arithmetic but not in floating point (1.2 / 0.1 =
11.999999999999998 in float64, 1.3 / 0.1 = 12.999999 in float32).
"""
import numpy as np
import pytest

import oracle
from brainpy.state_amd import (Simulator, aeif_cond_alpha, dist, fixed_indegree, iaf_psc_alpha,
                               iaf_psc_delta, iaf_psc_exp, poisson_generator, spike_recorder,
                               voltmeter)
from brainpy.state_amd import units as u
from brainpy.state_amd.init import Uniform

pytestmark = pytest.mark.gpu


def _net(model, comm, delay, dtype='float64', hetero=False, axonal=False, n=300):
    sim = Simulator(dt=0.1 * u.ms, dtype=dtype)
    cond = model is aeif_cond_alpha
    kw = {} if cond else dict(V_initializer=Uniform(-70., -56., seed=5))
    a = sim.create(model, n, **kw)
    b = sim.create(model, n // 4, **kw)
    noise = sim.create(poisson_generator, rate=12000. * u.Hz)
    rt = dict(receptor_type=1) if cond else {}
    rti = dict(receptor_type=2) if cond else {}
    wE, wI = (3.0, 4.0) if cond else (60.0, -240.0)
    if model is iaf_psc_delta:
        wE, wI = 0.5, -2.0
    sim.connect(noise, a, weight=wE, delay=delay, **rt)
    sim.connect(noise, b, weight=wE, delay=0.1 + delay, **rt)
    w = dist.Normal(wE, 0.2 * wE) if hetero else wE
    d_ee = delay
    if axonal:   # one delay per pre neuron, whole and fractional ratios mixed
        rng = np.random.default_rng(2)
        d_ee = rng.choice([1.0, 1.2, 1.5, 0.7, 2.3, 0.15], n) * u.ms
    sim.connect(a, a + b, weight=w, delay=d_ee, rule=fixed_indegree(30), seed=4, comm=comm, **rt)
    sim.connect(b, a + b, weight=wI, delay=2 * delay, rule=fixed_indegree(10), seed=5, comm=comm, **rti)
    sim.connect(a, sim.create(spike_recorder))
    sim.connect(b, sim.create(spike_recorder))
    vm = sim.create(voltmeter)
    sim.connect(vm, a[:5])
    return sim


def _check(sim, n_steps, rtol=1e-12):
    desc = sim.build_desc()
    eng = sim._ensure_engine()
    eng.reset()
    vm_g, (g_g, s_g) = eng.run(n_steps)
    vm_o, g_o, s_o = oracle.OracleNet(desc).run(n_steps)
    assert len(g_o) > 300, 'network too quiet for a useful check'
    assert len(g_g) == len(g_o), f'spike count differs: gpu {len(g_g)} oracle {len(g_o)}'
    assert np.array_equal(s_g, s_o) and np.array_equal(g_g, g_o), 'raster differs'
    np.testing.assert_allclose(vm_g, vm_o, rtol=rtol, atol=0)
    sim._engine.close()
    sim._engine = None


@pytest.mark.parametrize('model', [iaf_psc_delta, iaf_psc_exp, iaf_psc_alpha, aeif_cond_alpha])
@pytest.mark.parametrize('comm', ['sparse', 'dense'])
def test_near_integral_delay_lerps_fp64(model, comm):
    # 1.2 / 0.1 = 11.999999999999998: frames 11 and 12; 2.4 / 0.1 and
    # 1.3 / 0.1 likewise fractional for the inhibitory and generator reads
    _check(_net(model, comm, 1.2), 2000)


@pytest.mark.parametrize('comm', ['sparse', 'dense'])
def test_half_step_delay_and_heterogeneous_weights(comm):
    # 0.15 / 0.1 = 1.4999999999999998: frames 1 and 2 at about half each
    _check(_net(iaf_psc_exp, comm, 0.15, hetero=True), 2000)


def test_fractional_delays_in_float32():
    # in float32 1.3 / 0.1 = 12.999999 (fractional) and 1.2 / 0.1 = 12 (whole)
    _check(_net(iaf_psc_exp, 'sparse', 1.3, dtype='float32'), 2000, rtol=0)
    _check(_net(iaf_psc_alpha, 'dense', 1.3, dtype='float32'), 2000, rtol=0)


@pytest.mark.parametrize('comm', ['sparse', 'dense'])
def test_axonal_delays_mixing_whole_and_fractional(comm):
    _check(_net(iaf_psc_delta, comm, 1.0, axonal=True), 2000)
