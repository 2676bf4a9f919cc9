"""CPU tests of Simulator.get_connections / SynapseCollection (the reference's
connection_introspection.py:362-571): canonical edge order, filters, live
reads, guarded writes, and how a write reaches the next engine build."""
import numpy as np
import pytest

from brainpy.state_amd import (Simulator, SynapseCollection, explicit_edges, fixed_indegree,
                               iaf_psc_delta, iaf_psc_exp, stdp_synapse, units as u)


def _net():
    sim = Simulator(dt=0.1 * u.ms)
    E = sim.create(iaf_psc_exp, 6)
    I = sim.create(iaf_psc_exp, 3)
    pre = np.array([4, 1, 1, 0, 5, 1, 2])
    post = np.array([0, 2, 1, 3, 1, 0, 2])
    w = np.array([10., 11., 12., 13., 14., 15., 16.])
    p1 = sim.connect(E, I, rule=explicit_edges(pre, post % 3), weight=w, delay=1.0)
    p2 = sim.connect(I, E, rule=fixed_indegree(2), weight=-30.0, delay=2.0, seed=3)
    return sim, E, I, p1, p2, pre, post % 3, w


def test_canonical_order_is_stable_sort_by_source():
    sim, E, I, p1, p2, pre, post, w = _net()
    c = sim.get_connections(source=E, target=I)
    assert isinstance(c, SynapseCollection) and len(c) == 7
    o = np.argsort(pre, kind='stable')
    assert np.array_equal(c.source, pre[o]) and np.array_equal(c.target, post[o])
    assert np.array_equal(c.get('weight'), w[o])
    assert np.allclose(c.get('delay'), 1.0)
    assert list(c)[:2] == [(int(pre[o][0]), int(post[o][0])), (int(pre[o][1]), int(post[o][1]))]
    d = c.get(['source', 'weight'])
    assert set(d) == {'source', 'weight'}


def test_filters_by_population_index_and_synapse():
    sim, E, I, p1, p2, pre, post, w = _net()
    assert len(sim.get_connections()) == 7 + 6 * 2
    assert len(sim.get_connections(source=I)) == 12
    sub = sim.get_connections(source=E[1:2])
    assert np.all(sub.source == 1) and len(sub) == int((pre == 1).sum())
    assert len(sim.get_connections(synapse='stdp_synapse')) == 0
    assert len(sim.get_connections(synapse='static_synapse')) == 19
    h = sim.get_connections(source=I)
    assert np.allclose(h.get('weight'), -30.0) and np.allclose(h.get('delay'), 2.0)


def test_set_weight_scalar_and_per_edge_write_the_sampled_positions():
    sim, E, I, p1, p2, pre, post, w = _net()
    c = sim.get_connections(source=E[1:2])               # the three edges from pre 1
    c.set('weight', [1.0, 2.0, 3.0])
    sel = np.nonzero(pre == 1)[0]                        # sampled positions, in order
    assert np.array_equal(np.asarray(p1.weight)[sel], [1.0, 2.0, 3.0])
    others = np.setdiff1d(np.arange(7), sel)
    assert np.array_equal(np.asarray(p1.weight)[others], w[others])
    sim.get_connections(source=E, target=I).set('weight', 5.0 * u.pA)
    assert np.all(np.asarray(p1.weight) == 5.0)
    h = sim.get_connections(source=I)
    h.set('weight', -12.5)                               # scalar on a homogeneous projection
    assert p2.homogeneous and p2.weight == -12.5
    with pytest.raises(ValueError, match='homogeneous'):
        h.set('weight', np.ones(len(h)))
    with pytest.raises(ValueError, match='per-edge'):
        c.set('weight', [1.0, 2.0])
    with pytest.raises(KeyError):
        c.set('source', 1)
    with pytest.raises(KeyError):
        c.get('bogus')


def test_set_delay_rounds_to_grid_and_guards_plastic():
    sim, E, I, p1, p2, pre, post, w = _net()
    sim.get_connections(source=I).set('delay', 0.73 * u.ms)
    assert p2.delay == 7
    assert np.allclose(sim.get_connections(source=I).get('delay'), 0.7)
    with pytest.raises(ValueError, match='scalar'):
        sim.get_connections(source=I).set('delay', [1.0, 2.0])
    sim2 = Simulator(dt=0.1 * u.ms)
    A = sim2.create(iaf_psc_delta, 4)
    pl = sim2.connect(A, A, rule=fixed_indegree(2), synapse=stdp_synapse(weight=1.0, delay=1.0),
                      seed=1)
    c = sim2.get_connections(synapse='stdp_synapse')
    assert len(c) == 8 and np.allclose(c.get('weight'), 1.0)
    with pytest.raises(ValueError, match='plastic rule'):
        c.set('weight', 2.0)
    with pytest.raises(ValueError, match='plastic delay'):
        c.set('delay', 2.0)
    assert pl.plastic


def test_write_invalidates_the_engine_and_blocks_cont():
    sim, E, I, p1, p2, pre, post, w = _net()
    sim._rollout_ready = True                 # as after a simulate()
    sim.get_connections(source=E).set('weight', 1.0)
    assert sim._engine is None and sim._edited_mid_rollout
    with pytest.raises(RuntimeError, match='edited'):
        sim.cont(1.0 * u.ms)
    d = sim.build_desc()                      # the next engine sees the edited weights
    assert d.c.n_projs >= 2


def test_edited_network_equals_a_fresh_build_in_the_oracle():
    """Per-edge and scalar weight writes plus a delay write give exactly the
    network built with those values from the start (oracle rasters)."""
    import oracle
    from brainpy.state_amd import poisson_generator, spike_recorder
    from brainpy.state_amd.init import Uniform

    def net(edges=None, w_ie=-240.0, d_ie=2.0):
        sim = Simulator(dt=0.1 * u.ms)
        a = sim.create(iaf_psc_exp, 240, V_initializer=Uniform(-70., -56., seed=5))
        b = sim.create(iaf_psc_exp, 60, V_initializer=Uniform(-70., -56., seed=6))
        sim.connect(sim.create(poisson_generator, rate=12000. * u.Hz), a + b, weight=60.0, delay=1.0)
        if edges is None:
            pe = sim.connect(a, a, weight=60.0, delay=1.0, rule=fixed_indegree(30), seed=4)
        else:
            pe = sim.connect(a, a, weight=edges[2], delay=1.0, rule=explicit_edges(edges[0], edges[1]))
        sim.connect(a, b, weight=60.0, delay=1.0, rule=fixed_indegree(30), seed=7)
        sim.connect(b, a + b, weight=w_ie, delay=d_ie, rule=fixed_indegree(10), seed=5)
        for pop in (a, b):
            sim.connect(pop, sim.create(spike_recorder))
        return sim, a, b, pe

    sim0, _, _, pe0 = net()
    pre, post = np.asarray(pe0.pre_idx), np.asarray(pe0.post_idx)
    sim, a, b, _ = net(edges=(pre, post, np.full(pre.shape[0], 60.0)))
    w_new = np.random.default_rng(3).uniform(20.0, 100.0, pre.shape[0])
    sim.get_connections(source=a, target=a).set('weight', w_new)
    sim.get_connections(source=b).set('weight', -180.0)
    sim.get_connections(source=b).set('delay', 1.5)
    ws = np.empty_like(w_new)
    ws[np.argsort(pre, kind='stable')] = w_new
    sim2, _, _, _ = net(edges=(pre, post, ws), w_ie=-180.0, d_ie=1.5)
    _, g1, s1 = oracle.OracleNet(sim.build_desc()).run(1500)
    _, g2, s2 = oracle.OracleNet(sim2.build_desc()).run(1500)
    _, g0, _ = oracle.OracleNet(sim0.build_desc()).run(1500)
    assert len(g1) > 50 and len(g1) != len(g0)
    assert np.array_equal(g1, g2) and np.array_equal(s1, s2)
