"""GPU parity on edge cases of the network step, each against the oracle
(bit-exact raster, V_m to 1e-12 relative in FP64):

* ragged and one-neuron populations (tiles partly filled, a tile of one lane);
* projections with no edges (p = 0, K = 0) next to populated ones;
* a count table too long for LDS (in-degree 20,000 from one projection: the
  inputs phase reads the table from global memory);
* delays far beyond the window (up to 40 ms): arrivals past the LDS horizon
  go through the arrival ring, with 16-bit and with 32-bit edge words;
* only some populations recorded, runs of 0 and 1 steps, and a silent network;
* aeif_cond_alpha's instability guard: the run reports it as the oracle does,
  and a run below the guard's threshold equals the oracle.
"""
import numpy as np
import pytest

import oracle
from brainpy.state_amd import Simulator, units as u
from brainpy.state_amd.devices import poisson_generator, spike_recorder, voltmeter
from brainpy.state_amd.dist import UniformGrid
from brainpy.state_amd.init import Constant
from brainpy.state_amd.models import iaf_psc_alpha, iaf_psc_delta, iaf_psc_exp
from brainpy.state_amd.rules import all_to_all, fixed_indegree, pairwise_bernoulli

pytestmark = pytest.mark.gpu


def _both(sim, n_steps, split=None):
    desc = sim.build_desc()
    eng = sim._ensure_engine()
    eng.reset()
    parts = [eng.run(n) for n in (split or [n_steps])]
    vm_g = np.concatenate([p[0] for p in parts])
    g_g = np.concatenate([p[1][0] for p in parts])
    s_g = np.concatenate([p[1][1] for p in parts])
    vm_o, g_o, s_o = oracle.OracleNet(desc).run(n_steps)
    assert len(g_g) == len(g_o), f'spike count differs: gpu {len(g_g)} oracle {len(g_o)}'
    assert np.array_equal(s_g, s_o) and np.array_equal(g_g, g_o), 'raster differs'
    np.testing.assert_allclose(vm_g, vm_o, rtol=1e-12, atol=0)
    return g_o


def _delta(**kw):
    p = dict(C_m=250. * u.pF, tau_m=20. * u.ms, t_ref=2. * u.ms, E_L=0. * u.mV, V_reset=10. * u.mV,
             V_th=20. * u.mV, V_initializer=Constant(0. * u.mV))
    p.update(kw)
    return p


def test_ragged_and_single_neuron_populations():
    sim = Simulator(dt=0.1 * u.ms)
    pops = [sim.create(iaf_psc_delta, n, params=_delta()) for n in (1, 63, 65, 257)]
    noise = sim.create(poisson_generator, rate=18000. * u.Hz)
    for p in pops:
        sim.connect(noise, p, weight=0.1 * u.mV, delay=0.1 * u.ms)
    for k, a in enumerate(pops):
        for b in pops:
            sim.connect(a, b, rule=fixed_indegree(min(5, a.size)), weight=(0.4 if k % 2 == 0 else -1.0) * u.mV,
                        delay=(0.5 + 0.1 * k) * u.ms, seed=k + 1)
    recs = [sim.create(spike_recorder) for _ in pops]
    for p, r in zip(pops, recs):
        sim.connect(p, r)
    vm = sim.create(voltmeter)
    sim.connect(vm, pops[0])
    g = _both(sim, 3000)
    assert np.isin(0, g) and len(g) > 1000       # the one-neuron population fires


def test_empty_projections_beside_populated_ones():
    sim = Simulator(dt=0.1 * u.ms)
    a = sim.create(iaf_psc_exp, 200, params=dict(I_e=380.0))
    b = sim.create(iaf_psc_exp, 150)
    sim.connect(a, b, rule=pairwise_bernoulli(0.0), weight=50. * u.pA, delay=1. * u.ms, seed=1)
    sim.connect(a, b, rule=fixed_indegree(0), weight=50. * u.pA, delay=1. * u.ms, seed=2)
    sim.connect(a, b, rule=fixed_indegree(30), weight=400. * u.pA, delay=1. * u.ms, seed=3)
    sim.connect(b, a, rule=pairwise_bernoulli(0.0), weight=-50. * u.pA, delay=2. * u.ms, seed=4)
    ra, rb = sim.create(spike_recorder), sim.create(spike_recorder)
    sim.connect(a, ra)
    sim.connect(b, rb)
    g = _both(sim, 2000)
    assert (g >= 200).any()                      # b is driven through the one populated projection


def test_count_table_beyond_lds_uses_global_table():
    """In-degree 20,000 from one projection: the S[k] table (20,001 doubles)
    does not fit the LDS budget and is read from global memory."""
    sim = Simulator(dt=0.1 * u.ms)
    src = sim.create(iaf_psc_delta, 2000, params=_delta())
    dst = sim.create(iaf_psc_delta, 40, params=_delta())
    noise = sim.create(poisson_generator, rate=25000. * u.Hz)
    sim.connect(noise, src, weight=0.1 * u.mV, delay=0.1 * u.ms)
    sim.connect(src, dst, rule=fixed_indegree(20000), weight=0.013 * u.mV, delay=1.5 * u.ms, seed=5)
    rs, rd = sim.create(spike_recorder), sim.create(spike_recorder)
    sim.connect(src, rs)
    sim.connect(dst, rd)
    vm = sim.create(voltmeter)
    sim.connect(vm, dst[:5])
    g = _both(sim, 2000)
    assert (g >= 2000).any()


@pytest.mark.parametrize('tile_size,k_ab', [(0, 40), (0, 200), (256, 40)])
def test_long_delays_through_the_arrival_ring(tile_size, k_ab):
    """Delays up to 40 ms, through both gather paths. Automatic tiles are 32
    neurons here, so the edge words fit 16 bits (delay * stride + lane <=
    400 * 32 + 31); they are 16-bit when the (pre, tile) segments hold six
    edges or more on average (over both projections: 11 with in-degree 200;
    3 with 40, 32-bit). With
    256-neuron tiles 400 * 256 + 255 needs 32-bit words."""
    sim = Simulator(dt=0.1 * u.ms, tile_size=tile_size)
    a = sim.create(iaf_psc_alpha, 300, params=dict(I_e=400.0))
    b = sim.create(iaf_psc_alpha, 300)
    w_ab = 80. * 40 / k_ab
    sim.connect(a, b, rule=fixed_indegree(k_ab), weight=w_ab * u.pA, delay=UniformGrid(0.2, 40.0, 0.1), seed=6)
    sim.connect(b, a, rule=fixed_indegree(20), weight=-40. * u.pA, delay=UniformGrid(5.0, 35.0, 0.1), seed=7)
    ra, rb = sim.create(spike_recorder), sim.create(spike_recorder)
    sim.connect(a, ra)
    sim.connect(b, rb)
    vm = sim.create(voltmeter)
    sim.connect(vm, b[:4])
    g = _both(sim, 4000, split=[1777, 2223])
    assert (g >= 300).any()


def test_partially_recorded_network_and_short_runs():
    """Only population b is recorded (the rest still delivers); runs of 0 and
    1 steps between longer ones."""
    sim = Simulator(dt=0.1 * u.ms)
    a = sim.create(iaf_psc_delta, 300, params=_delta())
    b = sim.create(iaf_psc_delta, 100, params=_delta())
    noise = sim.create(poisson_generator, rate=20000. * u.Hz)
    sim.connect(noise, a + b, weight=0.1 * u.mV, delay=0.1 * u.ms)
    sim.connect(a, b, rule=fixed_indegree(30), weight=0.2 * u.mV, delay=1.0 * u.ms, seed=8)
    sim.connect(b, a, rule=all_to_all, weight=-0.05 * u.mV, delay=2.0 * u.ms)
    rb = sim.create(spike_recorder)
    sim.connect(b, rb)
    vm = sim.create(voltmeter)
    sim.connect(vm, a[:3])
    g = _both(sim, 2500, split=[0, 1, 1200, 0, 1, 1298])
    assert len(g) > 100 and (g >= 300).all()     # only b's spikes are kept


def test_silent_network():
    sim = Simulator(dt=0.1 * u.ms)
    a = sim.create(iaf_psc_exp, 128)
    sim.connect(a, a, rule=fixed_indegree(10), weight=20. * u.pA, delay=1. * u.ms, seed=9)
    r = sim.create(spike_recorder)
    sim.connect(a, r)
    vm = sim.create(voltmeter)
    sim.connect(vm, a[:2])
    g = _both(sim, 1000)
    assert len(g) == 0


def test_aeif_instability_is_reported_like_the_oracle():
    """A -1e7 pA current drives one aeif_cond_alpha neuron past -1e3 mV in its
    first step with the current: the run fails with the reference's message
    (aeif_cond_alpha.py:703-706); at -500 pA it runs and equals the oracle."""
    from brainpy.state_amd._lib import EngineError
    from test_aeif_unstable import unstable_net
    sim = unstable_net(-1.0e7)   # (the simulator owns the engine: keep it)
    eng = sim._ensure_engine()
    eng.reset()
    with pytest.raises(EngineError, match='Numerical instability in aeif_cond_alpha dynamics'):
        eng.run(40)
    _both(unstable_net(-500.0), 40)
