"""GPU checks of sharded runs: a network split into shards that exchange each
window's spike list (bs_net_inject_spikes) gives the raster and V_m of the
unsharded network, bit for bit against the oracle."""
import numpy as np
import pytest

import oracle
from brainpy.state_amd import Shards, Simulator, units as u
from brainpy.state_amd.devices import poisson_generator, spike_recorder, voltmeter
from brainpy.state_amd.dist import UniformGrid
from brainpy.state_amd.models import iaf_psc_alpha, iaf_psc_exp
from brainpy.state_amd.networks import brunel_delta
from brainpy.state_amd.rules import fixed_indegree, pairwise_bernoulli

pytestmark = pytest.mark.gpu


def test_brunel_two_shards_equal_the_unsharded_network():
    sim, h = brunel_delta(order=60, n_vm=4)
    n = 3000
    sh = Shards(sim, [[h['ne']], [h['ni']]])
    assert sh.W == 16
    g, s, vms = sh.run(n)
    sh.close()
    vm_o, g_o, s_o = oracle.OracleNet(sim.build_desc()).run(n)
    assert len(g_o) > 1000 and len(g) == len(g_o)
    assert np.array_equal(g, g_o) and np.array_equal(s, s_o)
    np.testing.assert_allclose(vms[0], vm_o, rtol=1e-12, atol=0)     # the E shard holds the taps
    eng = sim._ensure_engine()                                         # and the plain engine agrees
    eng.reset()
    _, (g1, s1) = eng.run(n)
    assert np.array_equal(g1, g) and np.array_equal(s1, s)


def test_three_shards_per_edge_delays_and_split_runs():
    sim = Simulator(dt=0.1 * u.ms, max_window=11)
    a = sim.create(iaf_psc_exp, 300, params=dict(I_e=200.0))
    b = sim.create(iaf_psc_alpha, 200)
    c = sim.create(iaf_psc_exp, 150)
    noise = sim.create(poisson_generator, rate=12000. * u.Hz)
    sim.connect(noise, a + b + c, weight=40. * u.pA, delay=0.1 * u.ms)
    sim.connect(a, b + c, rule=fixed_indegree(30), weight=60. * u.pA, delay=UniformGrid(1.0, 3.0, 0.1), seed=1)
    sim.connect(b, a, rule=pairwise_bernoulli(0.1), weight=-90. * u.pA, delay=UniformGrid(1.0, 2.0, 0.1), seed=2)
    sim.connect(c, a + c, rule=fixed_indegree(15), weight=70. * u.pA, delay=1.5 * u.ms, seed=3)
    recs = [sim.create(spike_recorder) for _ in range(3)]
    for p, r in zip((a, b, c), recs):
        sim.connect(p, r)
    vm = sim.create(voltmeter)
    sim.connect(vm, c[:3])
    sh = Shards(sim, [[a], [b], [c]])
    parts = [sh.run(n) for n in (1000, 777, 1)]
    sh.close()
    g = np.concatenate([p[0] for p in parts])
    s = np.concatenate([p[1] for p in parts])
    vm_c = np.concatenate([p[2][2] for p in parts])
    vm_o, g_o, s_o = oracle.OracleNet(sim.build_desc()).run(1778)
    assert len(g_o) > 500 and len(g) == len(g_o)
    assert np.array_equal(g, g_o) and np.array_equal(s, s_o)
    np.testing.assert_allclose(vm_c, vm_o, rtol=1e-12, atol=0)


def test_inject_rejects_bad_spike_lists():
    from brainpy.state_amd._lib import EngineError
    from brainpy.state_amd.simulator import _Engine, _NetBuilder
    sim, h = brunel_delta(order=10, n_vm=0)            # E: gids 0..39, I: 40..49
    eng = _Engine(_NetBuilder(sim, remote=[1]).build())
    W = eng.window()
    eng.run(W)
    eng.inject(np.array([40, 41], np.int32), np.array([0, W - 1], np.int32))   # fine
    with pytest.raises(EngineError, match='local to this shard'):
        eng.inject(np.array([3], np.int32), np.array([0], np.int32))
    with pytest.raises(EngineError, match='outside the window'):
        eng.inject(np.array([42], np.int32), np.array([W], np.int32))
    with pytest.raises(EngineError, match='gid out of range'):
        eng.inject(np.array([50], np.int32), np.array([0], np.int32))
    with pytest.raises(EngineError, match='repeated'):
        eng.inject(np.full(11, 45, np.int32), np.zeros(11, np.int32))
    eng.close()
