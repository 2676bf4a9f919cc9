"""GPU: edits through SynapseCollection.set reach the engine. The edited
network runs bit-exact against the oracle built from the same descriptor, and
matches a network built with those weights and delays from the start."""
import numpy as np
import pytest

import oracle
from brainpy.state_amd import (Simulator, explicit_edges, fixed_indegree, iaf_psc_exp,
                               poisson_generator, spike_recorder, stdp_synapse, iaf_psc_delta,
                               units as u)
from brainpy.state_amd.init import Uniform

pytestmark = pytest.mark.gpu


def _net(edges=None, w_ie=-240.0, d_ie=2.0):
    sim = Simulator(dt=0.1 * u.ms)
    a = sim.create(iaf_psc_exp, 240, V_initializer=Uniform(-70., -56., seed=5))
    b = sim.create(iaf_psc_exp, 60, V_initializer=Uniform(-70., -56., seed=6))
    noise = sim.create(poisson_generator, rate=12000. * u.Hz)
    sim.connect(noise, a + b, weight=60.0, delay=1.0)
    if edges is None:
        pe = sim.connect(a, a, weight=60.0, delay=1.0, rule=fixed_indegree(30), seed=4)
    else:
        pre, post, w = edges
        pe = sim.connect(a, a, weight=w, delay=1.0, rule=explicit_edges(pre, post))
    sim.connect(a, b, weight=60.0, delay=1.0, rule=fixed_indegree(30), seed=7)
    sim.connect(b, a + b, weight=w_ie, delay=d_ie, rule=fixed_indegree(10), seed=5)
    sr = [sim.create(spike_recorder), sim.create(spike_recorder)]
    sim.connect(a, sr[0])
    sim.connect(b, sr[1])
    return sim, a, b, pe, sr


def _run(sim, n):
    desc = sim.build_desc()
    eng = sim._ensure_engine()
    eng.reset()
    vm, (g, s) = eng.run(n)
    vm_o, g_o, s_o = oracle.OracleNet(desc).run(n)
    assert np.array_equal(g, g_o) and np.array_equal(s, s_o), 'raster differs from the oracle'
    return g, s


def test_edited_network_runs_bit_exact_and_equals_a_fresh_build():
    rng = np.random.default_rng(3)
    sim0, a0, b0, pe0, _ = _net()
    pre, post = np.asarray(pe0.pre_idx), np.asarray(pe0.post_idx)
    # a heterogeneous twin of the same edges, edited through the collection
    sim, a, b, pe, sr = _net(edges=(pre, post, np.full(pre.shape[0], 60.0)))
    sim.simulate(50.0 * u.ms)
    c = sim.get_connections(source=a, target=a)
    w_new = rng.uniform(20.0, 100.0, len(c))
    c.set('weight', w_new)
    sim.get_connections(source=b).set('weight', -180.0)
    sim.get_connections(source=b).set('delay', 1.5)
    with pytest.raises(RuntimeError):
        sim.cont(10.0 * u.ms)
    g1, s1 = _run(sim, 1500)
    # the same network built with the edited values from the start
    order = np.argsort(pre, kind='stable')
    w_sampled = np.empty_like(w_new)
    w_sampled[order] = w_new
    sim2, a2, b2, pe2, sr2 = _net(edges=(pre, post, w_sampled), w_ie=-180.0, d_ie=1.5)
    g2, s2 = _run(sim2, 1500)
    assert len(g1) > 100
    assert np.array_equal(g1, g2) and np.array_equal(s1, s2)
    assert np.array_equal(sim.get_connections(source=a, target=a).get('weight'), w_new)


def test_plastic_weights_read_live_after_simulate():
    sim = Simulator(dt=0.1 * u.ms)
    A = sim.create(iaf_psc_delta, 80, E_L=0., V_th=20., V_reset=10., tau_m=20., t_ref=2.,
                   V_initializer=Uniform(0., 19., seed=2))
    noise = sim.create(poisson_generator, rate=30000. * u.Hz)
    sim.connect(noise, A, weight=0.6, delay=1.0)
    p = sim.connect(A, A, rule=fixed_indegree(8), seed=2,
                    synapse=stdp_synapse(weight=1.0, delay=1.0, Wmax=3.0, lambda_=0.05))
    c = sim.get_connections(synapse='stdp_synapse')
    before = c.get('weight').copy()
    sim.simulate(300.0 * u.ms)
    after = c.get('weight')
    assert np.array_equal(after, sim.get_weights(p))
    assert not np.array_equal(after, before)
