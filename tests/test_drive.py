"""Injected Poisson drive (bs_net_set_drive / Simulator.set_poisson_drive) in
the oracle: counts fed in for a range of generator steps stand for the
generator's draws (poisson_generator.py:501-576: count of the step, kept
inside the activity window), and arrive after the holder lag plus the
connection's delay like drawn ones. Pinned here by (1) feeding the oracle the
counts its own stream draws, which must change nothing, including across a
start/stop window and a fractional delay, and (2) a one-neuron known answer."""
import numpy as np
import pytest

import oracle
from brainpy.state_amd import Simulator, units as u
from brainpy.state_amd.devices import poisson_generator, spike_recorder, voltmeter
from brainpy.state_amd.init import Constant
from brainpy.state_amd.models import iaf_psc_delta
from brainpy.state_amd.rules import fixed_indegree


def _net():
    sim = Simulator(dt=0.1 * u.ms)
    par = dict(C_m=250. * u.pF, tau_m=20. * u.ms, t_ref=2. * u.ms, E_L=0. * u.mV, V_reset=10. * u.mV,
               V_th=20. * u.mV, V_initializer=Constant(0. * u.mV))
    e = sim.create(iaf_psc_delta, 80, params=par)
    i = sim.create(iaf_psc_delta, 20, params=par)
    noise = sim.create(poisson_generator, rate=20000. * u.Hz)
    burst = sim.create(poisson_generator, rate=30000. * u.Hz, start=10. * u.ms, stop=25. * u.ms)
    c1 = sim.connect(noise, e, weight=0.1 * u.mV, delay=1.5 * u.ms)
    c2 = sim.connect(noise, i, weight=0.1 * u.mV, delay=1.5 * u.ms)
    c3 = sim.connect(burst, e[10:50], weight=0.2 * u.mV, delay=0.15 * u.ms)   # fractional: 1.5 steps
    extra = sim.create(poisson_generator, rate=5000. * u.Hz)
    c4 = sim.connect(extra, e[:30], weight=0.1 * u.mV, delay=0.5 * u.ms)      # E's second Poisson channel
    sim.connect(e, e + i, weight=0.1 * u.mV, delay=1.5 * u.ms, rule=fixed_indegree(8), seed=1)
    sim.connect(i, e + i, weight=-0.5 * u.mV, delay=1.5 * u.ms, rule=fixed_indegree(2), seed=2)
    sim.connect(e, sim.create(spike_recorder))
    sim.connect(i, sim.create(spike_recorder))
    vm = sim.create(voltmeter)
    sim.connect(vm, e[:5])
    return sim, (c1, c2, c3, c4)


def _stream(conn, dt, step0, n):
    lam = conn.gen.lam(dt)
    return np.array([[oracle.poisson(int(conn.seed) & 0xFFFFFFFF, step0 + t, e, lam)
                      for e in range(len(conn.post_idx))] for t in range(n)], np.int32)


def test_injecting_the_streams_own_counts_changes_nothing():
    sim, conns = _net()
    desc = sim.build_desc()
    vm0, g0, s0 = oracle.OracleNet(desc).run(400)
    assert len(g0) > 50
    o = oracle.OracleNet(desc)
    # channel 3 is active over generator steps 101..250: inject 60..319, so
    # the injected rows before and after the window must be gated to zero
    for conn, (a, n) in zip(conns, ((40, 200), (0, 400), (60, 260), (5, 300))):
        o.set_drive(desc.chan_index[id(conn)], a, _stream(conn, sim.dt, a, n))
    vm1, g1, s1 = o.run(400)
    assert np.array_equal(g0, g1) and np.array_equal(s0, s1)
    assert np.array_equal(vm0, vm1)


def test_changed_counts_change_the_run():
    sim, conns = _net()
    desc = sim.build_desc()
    _, g0, _ = oracle.OracleNet(desc).run(400)
    o = oracle.OracleNet(desc)
    o.set_drive(desc.chan_index[id(conns[0])], 0, np.zeros((400, 80), np.int32))
    _, g1, _ = o.run(400)
    assert (g1 < 80).sum() < (g0 < 80).sum()   # E lost its main drive


def test_injected_count_arrives_after_holder_lag_and_delay():
    """A count k of generator step n moves the target at step n + 1 + d by
    k * weight (iaf_psc_delta from rest, E_L = 0: V = k J exactly)."""
    sim = Simulator(dt=0.1 * u.ms)
    nrn = sim.create(iaf_psc_delta, 1, E_L=0., V_th=1e9, V_reset=0., t_ref=2.,
                     V_initializer=Constant(0.0))
    gen = sim.create(poisson_generator, rate=5000. * u.Hz)
    conn = sim.connect(gen, nrn, weight=0.25 * u.mV, delay=1.0 * u.ms)   # 10 steps
    vm = sim.create(voltmeter)
    sim.connect(vm, nrn)
    desc = sim.build_desc()
    o = oracle.OracleNet(desc)
    counts = np.zeros((60, 1), np.int32)
    counts[20, 0] = 3
    o.set_drive(desc.chan_index[id(conn)], 0, counts)
    v, _, _ = o.run(40)
    assert np.all(v[:31, 0] == 0.0)
    assert v[31, 0] == 0.75


def test_set_poisson_drive_validates_its_arguments():
    sim, conns = _net()
    with pytest.raises(ValueError):
        sim.set_poisson_drive(conns[0], np.zeros((10, 79)))      # one column per target neuron
    with pytest.raises(ValueError):
        sim.set_poisson_drive(conns[0], -np.ones((10, 80)))
    with pytest.raises(TypeError):
        sim.set_poisson_drive(object(), np.zeros((10, 80)))
    sim.set_poisson_drive(conns[2], np.ones((5, 40)), start_step=3)   # kept for the engine
    assert len(sim._drives) == 1
    sim.set_poisson_drive(conns[2], None)
    assert not sim._drives
