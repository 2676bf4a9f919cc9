"""GPU parity of stdp_synapse projections: raster, V_m and final weights
against the oracle's literal per-step, per-edge restatement."""
import numpy as np
import pytest

import oracle
from brainpy.state_amd import (Simulator, aeif_cond_alpha, fixed_indegree, iaf_psc_exp,
                               poisson_generator, spike_recorder, stdp_synapse, voltmeter)
from brainpy.state_amd import units as u
from brainpy.state_amd.init import Uniform

pytestmark = pytest.mark.gpu


def _plastic_net(model=iaf_psc_exp, n=400, delay=('uniform', 1.0, 3.0), dtype='float64',
                 mu=(0.0, 1.0)):
    sim = Simulator(dt=0.1 * u.ms, dtype=dtype)
    cond = model is aeif_cond_alpha
    kw = {} if cond else dict(V_initializer=Uniform(-70., -56., seed=5))
    e = sim.create(model, n, **kw)
    i = sim.create(model, n // 4, **kw)
    noise = sim.create(poisson_generator, rate=14000. * u.Hz)
    rt = dict(receptor_type=1) if cond else {}
    rti = dict(receptor_type=2) if cond else {}
    wE, wI = (3.0, 6.0) if cond else (40.0, -160.0)
    sim.connect(noise, e + i, weight=wE, delay=1.0, **rt)
    if isinstance(delay, tuple):
        # axonal delays: one per pre neuron in [lo, hi] ms, among the values
        # whose ratio to dt is whole in float64 (a plastic projection needs
        # whole-step delays; 1.2 / 0.1 = 11.999999999999998 would lerp)
        rng = np.random.default_rng(11)
        lo, hi = int(round(delay[1] * 10)), int(round(delay[2] * 10))
        whole = [k for k in range(lo, hi + 1) if (k / 10) / 0.1 == k]
        delay = rng.choice(whole, n) / 10 * u.ms
    syn = stdp_synapse(weight=wE, tau_plus=20., tau_minus=20., lambda_=0.05,
                       alpha=1.1, mu_plus=mu[0], mu_minus=mu[1], Wmax=3 * wE)
    proj = sim.connect(e, e, rule=fixed_indegree(20), synapse=syn, delay=delay, seed=3, **rt)
    sim.connect(e, i, weight=wE, delay=1.5, rule=fixed_indegree(20), seed=4, **rt)
    sim.connect(i, e + i, weight=wI, delay=1.0, rule=fixed_indegree(5), seed=5, **rti)
    sim.connect(e, sim.create(spike_recorder))
    sim.connect(i, sim.create(spike_recorder))
    vm = sim.create(voltmeter)
    sim.connect(vm, e[:4])
    return sim, proj


def _compare(sim, proj, n_steps, rtol=1e-12):
    desc = sim.build_desc()
    eng = sim._ensure_engine()
    eng.reset()
    vm_g, (g_g, s_g) = eng.run(n_steps)
    w_g = sim.get_weights(proj)
    orc = oracle.OracleNet(desc)
    vm_o, g_o, s_o = orc.run(n_steps)
    w_o = orc.weights(desc.proj_index[id(proj)], proj.n_edges)
    assert len(g_o) > 200
    assert len(g_g) == len(g_o), f'spike count differs: gpu {len(g_g)} oracle {len(g_o)}'
    assert np.array_equal(s_g, s_o) and np.array_equal(g_g, g_o)
    np.testing.assert_allclose(vm_g, vm_o, rtol=rtol, atol=0)
    w0 = np.asarray(proj.weight, np.float64)
    assert np.any(w_o != w0), 'no plasticity happened'
    np.testing.assert_allclose(w_g, w_o, rtol=rtol, atol=0)
    return w_g


def test_stdp_iaf_psc_exp_per_edge_delays():
    sim, proj = _plastic_net()
    _compare(sim, proj, 2000)


@pytest.mark.parametrize('mu', [(1.0, 1.0), (0.4, 0.4)])
def test_stdp_multiplicative_and_fractional_mu(mu):
    sim, proj = _plastic_net(n=200, delay=1.0, mu=mu)
    _compare(sim, proj, 1500)


def test_stdp_aeif_cond_alpha():
    sim, proj = _plastic_net(model=aeif_cond_alpha, n=200, delay=('uniform', 1.0, 5.0))
    _compare(sim, proj, 1000)


def test_stdp_chunks_and_continue():
    """Two runs back to back equal one long run (traces and weights carry over)."""
    sim, proj = _plastic_net(n=200, delay=1.0)
    desc = sim.build_desc()
    eng = sim._ensure_engine()
    eng.reset()
    eng.run(700)
    vm_b, (g_b, s_b) = eng.run(800)
    orc = oracle.OracleNet(desc)
    orc.run(700)
    vm_o, g_o, s_o = orc.run(800)
    assert np.array_equal(g_b, g_o) and np.array_equal(s_b, s_o)
    np.testing.assert_allclose(sim.get_weights(proj),
                               orc.weights(desc.proj_index[id(proj)], proj.n_edges), rtol=1e-12)


def _two_plastic_net():
    sim = Simulator(dt=0.1 * u.ms)
    kw = dict(V_initializer=Uniform(-70., -56., seed=5))
    e = sim.create(iaf_psc_exp, 300, **kw)
    i = sim.create(iaf_psc_exp, 75, **kw)
    noise = sim.create(poisson_generator, rate=14000. * u.Hz)
    sim.connect(noise, e + i, weight=40., delay=1.0)
    see = stdp_synapse(weight=40., tau_plus=20., tau_minus=20., lambda_=0.05, alpha=1.1,
                       mu_plus=0.0, mu_minus=1.0, Wmax=120.)
    sie = stdp_synapse(weight=-160., tau_plus=15., tau_minus=30., lambda_=0.03, alpha=0.9,
                       mu_plus=1.0, mu_minus=1.0, Wmax=-480.)
    pee = sim.connect(e, e, rule=fixed_indegree(20), synapse=see, delay=2.0, seed=3)
    pie = sim.connect(i, e, rule=fixed_indegree(5), synapse=sie, delay=1.0, seed=4)
    sim.connect(e, i, weight=40., delay=1.5, rule=fixed_indegree(20), seed=5)
    sim.connect(i, i, weight=-160., delay=1.0, rule=fixed_indegree(5), seed=6)
    sim.connect(e, e, weight=10., delay=2.0, rule=fixed_indegree(10), seed=7)   # static, after both
    sim.connect(e, sim.create(spike_recorder))
    sim.connect(i, sim.create(spike_recorder))
    vm = sim.create(voltmeter)
    sim.connect(vm, e[:4])
    return sim, (pee, pie)


def test_two_plastic_projections_into_one_population_and_a_static_tail():
    """Excitatory and inhibitory stdp_synapse projections both onto E, each
    with its own traces and time constants (the inhibitory one with negative
    weights and Wmax); a static projection registered after them, whose input
    joins the sum after both plastic channels, in channel order; E both pre
    and post of a plastic projection. Raster, V_m and both weight sets equal
    the oracle."""
    sim, projs = _two_plastic_net()
    desc = sim.build_desc()
    eng = sim._ensure_engine()
    assert eng.window() == 11
    eng.reset()
    vm_g, (g_g, s_g) = eng.run(2000)
    orc = oracle.OracleNet(desc)
    vm_o, g_o, s_o = orc.run(2000)
    assert len(g_o) > 200
    assert np.array_equal(g_g, g_o) and np.array_equal(s_g, s_o)
    np.testing.assert_allclose(vm_g, vm_o, rtol=1e-12, atol=0)
    for proj in projs:
        w_o = orc.weights(desc.proj_index[id(proj)], proj.n_edges)
        assert np.any(w_o != np.asarray(proj.weight, np.float64)), 'no plasticity happened'
        np.testing.assert_allclose(sim.get_weights(proj), w_o, rtol=1e-12, atol=0)


def test_stdp_runs_in_windows_of_d_min_plus_one_steps():
    """Plastic projections no longer shorten the window: with a 1 ms minimum
    delay every launch advances d_min + 1 = 11 steps."""
    sim, _ = _plastic_net(n=200, delay=1.0)
    assert sim._ensure_engine().window() == 11


@pytest.mark.parametrize('records', [64, 128])
def test_stdp_sub_windows_equal_the_oracle(records):
    """A block holding only ``records`` plastic arrivals at once runs a busier
    window in sub-windows (each one's arrival lists gathered in turn); the
    raster, V_m and weights still equal the oracle's step-by-step run. (This
    network's 32-neuron tiles see up to 47 plastic arrivals in one step and
    about 125 per 11-step window.)"""
    sim, proj = _plastic_net(n=300)
    sim.plastic_records = records
    _compare(sim, proj, 1500)


@pytest.mark.parametrize('tile,records', [(32, 64), (64, 128)])
def test_stdp_aeif_sub_windows_equal_the_oracle(tile, records):
    """aeif_cond_alpha lanes move through a window independently (each lane
    takes its next step once its own substeps for the current one are done);
    with the window's plastic arrivals split into sub-windows, the lanes meet
    again at every sub-window's end. Raster, V_m and weights equal the
    oracle. (This network's E tiles see up to 37 / 68 plastic arrivals in one
    step and 170 / 320 in one 11-step window at 32 / 64 neurons per tile.)"""
    sim, proj = _plastic_net(model=aeif_cond_alpha, n=300, delay=('uniform', 1.0, 5.0))
    sim.tile_size = tile
    sim.plastic_records = records
    _compare(sim, proj, 1500)


def test_stdp_record_capacity_error_is_reported():
    """More plastic arrivals in one step than a block holds is an error the
    run reports (not a silent drop)."""
    from brainpy.state_amd._lib import EngineError
    sim, _ = _plastic_net(n=300)
    sim.plastic_records = 1
    eng = sim._ensure_engine()
    eng.reset()
    with pytest.raises(EngineError, match='record capacity'):
        eng.run(1500)


@pytest.mark.parametrize('model,unit', [(iaf_psc_exp, 'pA'), (aeif_cond_alpha, 'nS')])
def test_record_weight_trajectory_equals_the_oracle_step_by_step(model, unit):
    """record_weight (simulator.py:884-910): the (n_steps, n_edges) weight
    trajectory of a plastic projection, post-update every step, in the CSR
    edge order, in the weight unit of the target model; every row equals the
    oracle's weights after that step."""
    sim, proj = _plastic_net(model=model, n=200, delay=1.0)
    sim.record_weight(proj)
    n = 400
    res = sim.simulate(n * 0.1 * u.ms)
    tr = res.weight_trace(proj)
    assert tr.unit == getattr(u, unit)
    w = np.asarray(tr.value)
    assert w.shape == (n, proj.n_edges)
    np.testing.assert_array_equal(w[-1], sim.get_weights(proj))
    desc = sim.build_desc()
    orc = oracle.OracleNet(desc)
    k = desc.proj_index[id(proj)]
    rows = []
    for _ in range(n):
        orc.run(1)
        rows.append(orc.weights(k, proj.n_edges))
    np.testing.assert_allclose(w, np.array(rows), rtol=1e-12, atol=0)
    assert np.any(w[-1] != w[0]), 'no plasticity happened'
