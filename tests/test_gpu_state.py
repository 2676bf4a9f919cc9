"""GPU checks of the boundary's state and device-run entry points.

* bs_net_save_state / bs_net_load_state: save -> run equals load -> run, in
  the same handle and in a fresh handle of the same network, for a network
  with per-edge delays, Poisson drive and an stdp_synapse projection (raster,
  V_m and the plastic weights); a snapshot of another network is rejected.
* bs_net_set_state: writing the initial V_m of a network built from rest
  gives the raster of the network built with those values.
* bs_net_run_device: the raster bits and V_m rows written on a torch stream
  equal bs_net_run's spike list and traces, over split runs.
"""
import numpy as np
import pytest

import oracle
from brainpy.state_amd import (Simulator, fixed_indegree, iaf_psc_exp, poisson_generator,
                               spike_recorder, stdp_synapse, voltmeter)
from brainpy.state_amd import units as u
from brainpy.state_amd._lib import EngineError
from brainpy.state_amd.init import Constant, Uniform
from brainpy.state_amd.networks import brunel_delta

pytestmark = pytest.mark.gpu


def _plastic(V_init=None):
    sim = Simulator(dt=0.1 * u.ms)
    init = V_init if V_init is not None else Uniform(-70., -56., seed=5)
    e = sim.create(iaf_psc_exp, 300, V_initializer=init)
    i = sim.create(iaf_psc_exp, 75, V_initializer=init)
    sim.connect(sim.create(poisson_generator, rate=14000. * u.Hz), e + i, weight=40.0, delay=1.0)
    syn = stdp_synapse(weight=40.0, lambda_=0.05, alpha=1.1, mu_plus=0.0, mu_minus=1.0, Wmax=120.0)
    d = np.random.default_rng(3).choice([1.0, 1.5, 2.0, 2.5], 300) * u.ms
    proj = sim.connect(e, e, rule=fixed_indegree(20), synapse=syn, delay=d, seed=3)
    sim.connect(e, i, weight=40.0, delay=1.5, rule=fixed_indegree(20), seed=4)
    sim.connect(i, e + i, weight=-160.0, delay=1.0, rule=fixed_indegree(5), seed=5)
    sim.connect(e, sim.create(spike_recorder))
    sim.connect(i, sim.create(spike_recorder))
    vm = sim.create(voltmeter)
    sim.connect(vm, e[:4])
    return sim, proj


def test_save_load_continues_bit_for_bit():
    sim, proj = _plastic()
    eng = sim._ensure_engine()
    eng.reset()
    eng.run(1237)
    blob = eng.save_state()
    vm_a, (g_a, s_a) = eng.run(1500)
    w_a = sim.get_weights(proj)
    assert len(g_a) > 500
    eng.load_state(blob)
    vm_b, (g_b, s_b) = eng.run(1500)
    assert np.array_equal(g_a, g_b) and np.array_equal(s_a, s_b)
    assert np.array_equal(vm_a, vm_b) and np.array_equal(sim.get_weights(proj), w_a)
    # a fresh handle of the same network continues the same way
    sim2, proj2 = _plastic()
    eng2 = sim2._ensure_engine()
    eng2.load_state(blob)
    vm_c, (g_c, s_c) = eng2.run(1500)
    assert np.array_equal(g_a, g_c) and np.array_equal(s_a, s_c) and np.array_equal(vm_a, vm_c)
    assert np.array_equal(sim2.get_weights(proj2), w_a)
    # and the whole run equals the oracle's uninterrupted one
    vm_o, g_o, s_o = oracle.OracleNet(sim.build_desc()).run(1237 + 1500)
    late = s_o >= 1237
    assert np.array_equal(g_o[late], g_c) and np.array_equal(s_o[late], s_c)
    # another network's snapshot is refused
    other, _ = brunel_delta(order=20, n_vm=2)
    with pytest.raises(EngineError):
        other._ensure_engine().load_state(blob)
    for s_ in (sim, sim2, other):
        s_._engine.close()
        s_._engine = None


def test_simulator_save_load_and_cont():
    sim, h = brunel_delta(order=60, delay=('uniform', 1.0, 3.0), n_vm=3)
    sim.simulate(50. * u.ms)
    blob = sim.save_state()
    r1 = sim.cont(80. * u.ms)
    sim.load_state(blob)
    r2 = sim.cont(80. * u.ms)
    assert np.array_equal(r1.spikes(h['esr']), r2.spikes(h['esr']))
    assert np.array_equal(np.asarray(r1.trace(h['vm']).value), np.asarray(r2.trace(h['vm']).value))
    assert float(r2.times.to_decimal(u.ms)[0]) == pytest.approx(50.0)
    sim._engine.close()
    sim._engine = None


def test_set_state_equals_building_with_the_values():
    sim_a, _ = _plastic()                          # V from Uniform(-70, -56)
    sim_b, _ = _plastic(V_init=Constant(-70.))     # from rest, then V written
    eng_a = sim_a._ensure_engine()
    eng_b = sim_b._ensure_engine()
    eng_a.reset()
    eng_b.reset()
    v0, v1 = eng_a.state(0, 0, 300), eng_a.state(1, 0, 75)
    assert np.ptp(v0) > 10.0 and np.all(eng_b.state(0, 0, 300) == -70.0)
    eng_b.set_state(0, 0, v0)
    eng_b.set_state(1, 0, v1)
    np.testing.assert_array_equal(eng_b.state(0, 0, 300), v0)
    _, (g_a, s_a) = eng_a.run(1000)
    _, (g_b, s_b) = eng_b.run(1000)
    assert len(g_a) > 100 and np.array_equal(g_a, g_b) and np.array_equal(s_a, s_b)
    for s_ in (sim_a, sim_b):
        s_._engine.close()
        s_._engine = None


def test_run_device_raster_equals_host_run():
    import torch
    sim, h = brunel_delta(order=100, delay=('uniform', 1.0, 3.0), n_vm=5)
    eng = sim._ensure_engine()
    N = int(eng.desc.ntot)
    eng.reset()
    ref = [eng.run(n) for n in (2345, 6000)]
    eng.reset()
    st = torch.cuda.Stream()
    outs = [eng.run_device(n, stream=st) for n in (2345, 6000)]
    st.synchronize()
    start = 0
    for n, (vm_h, (g_h, s_h)), (bits, vm_d, status) in zip((2345, 6000), ref, outs):
        assert int(status.item()) == 0 and bits.shape[0] == n
        b = np.ascontiguousarray(bits.cpu().numpy()).view(np.uint32)
        flags = np.unpackbits(b.view(np.uint8), bitorder='little').reshape(n, -1)[:, :N]
        s_d, g_d = np.nonzero(flags)                 # (step, gid), sorted by step then gid
        assert np.array_equal(g_d, g_h) and np.array_equal(s_d + start, s_h)
        np.testing.assert_array_equal(vm_d.cpu().numpy(), vm_h)
        start += n
    assert len(ref[1][1][0]) > 1000
    sim._engine.close()
    sim._engine = None
