"""GPU checks of the device connectivity path (bs_conn_* and bs_net_create's
device segment build).

* the device samplers' edges equal the oracle's restatement bit for bit
  (indices are integers: exact), with and without index maps;
* an engine built from device-resident edges gives the raster and V_m of the
  oracle run on the same edges (downloaded) and of an engine built from the
  host copies, for COUNT channels (homogeneous weights), FIXED channels
  (per-edge weights) and per-edge delays;
* invalid device edges are rejected with the host build's messages.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from brainpy.state_amd import Simulator, units as u
from brainpy.state_amd import device_conn
from brainpy.state_amd._lib import EngineError
from brainpy.state_amd.dist import Normal, UniformGrid
from brainpy.state_amd.devices import poisson_generator, spike_recorder, voltmeter
from brainpy.state_amd.models import iaf_psc_exp
from brainpy.state_amd.networks import brunel_delta
from brainpy.state_amd.rules import fixed_indegree, pairwise_bernoulli
from brainpy.state_amd.simulator import _Engine

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('seed,n_pre,n_post,K,excl', [
    (1, 800, 1000, 80, False), (2, 1000, 1000, 100, True), (2 ** 40 + 3, 3, 5000, 7, True),
    (7, 1, 300, 4, False), (8, 20000, 64, 2000, False)])
def test_fixed_indegree_equals_oracle(seed, n_pre, n_post, K, excl):
    de = device_conn.sample_fixed_indegree(seed, n_pre, n_post, K, exclude_self=excl)
    pre, post = de.host()
    rp, rq = oracle.conn_fixed_indegree(seed, n_pre, n_post, K, excl)
    assert de.n_edges == n_post * K and de.max_indegree == K
    assert np.array_equal(pre, rp) and np.array_equal(post, rq)


def test_fixed_indegree_index_maps():
    pre_map = np.arange(500, 800, dtype=np.int32)[::-1].copy()
    post_map = np.arange(0, 400, 2, dtype=np.int32)
    de = device_conn.sample_fixed_indegree(5, 300, 200, 30, exclude_self=False, pre_map=pre_map,
                                           post_map=post_map)
    rp, rq = oracle.conn_fixed_indegree(5, 300, 200, 30, False)
    pre, post = de.host()
    assert np.array_equal(pre, pre_map[rp]) and np.array_equal(post, post_map[rq])


@pytest.mark.parametrize('seed,n_pre,n_post,p,excl', [
    (3, 400, 1000, 0.1, False), (4, 777, 777, 0.05, True), (5, 64, 3000, 0.5, False),
    (6, 50, 50, 1.0, True), (9, 30, 257, 0.0, False), (10, 2000, 300, 0.01, True)])
def test_bernoulli_equals_oracle(seed, n_pre, n_post, p, excl):
    de = device_conn.sample_pairwise_bernoulli(seed, n_pre, n_post, p, exclude_self=excl)
    pre, post = de.host()
    rp, rq = oracle.conn_bernoulli(seed, n_pre, n_post, p, excl)
    assert np.array_equal(pre, rp) and np.array_equal(post, rq)
    kmax = int(np.bincount(rq, minlength=n_post).max()) if len(rq) else 0
    assert de.max_indegree == kmax


def test_bernoulli_index_maps():
    pre_map = (np.arange(120, dtype=np.int32) * 3 + 1)
    post_map = np.arange(900, 1400, dtype=np.int32)
    de = device_conn.sample_pairwise_bernoulli(11, 120, 500, 0.2, exclude_self=True, pre_map=pre_map,
                                               post_map=post_map)
    rp, rq = oracle.conn_bernoulli(11, 120, 500, 0.2, True)
    pre, post = de.host()
    assert np.array_equal(pre, pre_map[rp]) and np.array_equal(post, post_map[rq])


def _run_three_ways(sim, n_steps):
    """(device-built engine, host-built engine, oracle) on the same edges."""
    eng = sim._ensure_engine()
    assert any(bool(p.on_device) for p in eng.desc.c.projs[:eng.desc.c.n_projs])
    eng.reset()
    vm_d, (g_d, s_d) = eng.run(n_steps)
    desc_h = sim.build_desc(host_edges=True)
    eng_h = _Engine(desc_h)
    vm_h, (g_h, s_h) = eng_h.run(n_steps)
    eng_h.close()
    vm_o, g_o, s_o = oracle.OracleNet(desc_h).run(n_steps)
    return (vm_d, g_d, s_d), (vm_h, g_h, s_h), (vm_o, g_o, s_o)


def _same(a, b, rtol=1e-12):
    assert len(a[1]) == len(b[1]), f'spike count {len(a[1])} vs {len(b[1])}'
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    np.testing.assert_allclose(a[0], b[0], rtol=rtol, atol=0)


def test_device_brunel_delta_equals_host_build_and_oracle():
    sim, h = brunel_delta(order=100, n_vm=5, connectivity='device')
    d, hh, o = _run_three_ways(sim, 3000)
    assert len(o[1]) > 1000
    _same(d, o)
    _same(d, hh, rtol=0)


def test_device_brunel_per_edge_delays():
    sim, h = brunel_delta(order=60, delay=('uniform', 1.0, 3.0), n_vm=4, connectivity='device')
    d, hh, o = _run_three_ways(sim, 2500)
    assert len(o[1]) > 500
    _same(d, o)
    _same(d, hh, rtol=0)


def test_device_bernoulli_subsets_fixed_weights_and_delays():
    """pairwise_bernoulli between population slices (index maps), per-edge
    normal weights (FIXED channel, fixed-point sums) and per-edge delays."""
    sim = Simulator(dt=0.1 * u.ms, connectivity='device')
    a = sim.create(iaf_psc_exp, 600, params=dict(I_e=300.0))
    b = sim.create(iaf_psc_exp, 400)
    noise = sim.create(poisson_generator, rate=15000. * u.Hz)
    sim.connect(noise, a + b, weight=30. * u.pA, delay=0.1 * u.ms)
    sim.connect(a[100:500], b[50:350], rule=pairwise_bernoulli(0.1), weight=Normal(40. * u.pA, 10. * u.pA),
                delay=UniformGrid(0.5, 2.0, 0.1), seed=4)
    sim.connect(b, a, rule=pairwise_bernoulli(0.05), weight=-80. * u.pA, delay=1.0 * u.ms, seed=5)
    sim.connect(a, a, rule=fixed_indegree(20), weight=Normal(25. * u.pA, 5. * u.pA),
                delay=UniformGrid(1.0, 3.0, 0.1), allow_autapses=False, seed=6)
    ra, rb = sim.create(spike_recorder), sim.create(spike_recorder)
    sim.connect(a, ra)
    sim.connect(b, rb)
    vm = sim.create(voltmeter)
    sim.connect(vm, b[:6])
    d, hh, o = _run_three_ways(sim, 2000)
    assert len(o[1]) > 500
    _same(d, o)
    _same(d, hh, rtol=0)


def test_invalid_device_edges_are_rejected():
    import torch
    sim, h = brunel_delta(order=20, n_vm=0, connectivity='device')
    desc = sim.build_desc()
    pj = desc.c.projs[0]
    assert pj.on_device == 1
    bad = torch.full((int(pj.n_edges),), 5, dtype=torch.int32, device='cuda')
    bad[17] = 10 ** 6                        # out of the pre population
    torch.cuda.synchronize()
    pj.pre = C.cast(C.c_void_p(bad.data_ptr()), C.POINTER(C.c_int32))
    with pytest.raises(EngineError, match='edge index out of range'):
        _Engine(desc)
    del bad
