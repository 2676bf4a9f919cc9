"""GPU checks of the batched ensemble: every network advanced inside one engine
gives exactly the raster of its own plain run; the launch shape (tile size,
window bound) changes nothing; the autotuner returns a measured candidate;
current generators (dc, step, ac) survive the merge."""
import numpy as np
import pytest

from brainpy.state_amd import Ensemble
from brainpy.state_amd.networks import brunel_exp_ensemble

pytestmark = pytest.mark.gpu


def _plain(sim, n):
    eng = sim._ensure_engine()
    eng.reset()
    vm, (g, s) = eng.run(n)
    return g, s


def test_ensemble_equals_sequential_plain_runs():
    sims, p = brunel_exp_ensemble(6, n_min=200, n_max=1500, seed=5)
    ens = Ensemble(sims)
    parts = ens.run(2000)
    total = 0
    for sim, (vm, g, s) in zip(sims, parts):
        g0, s0 = _plain(sim, 2000)
        assert np.array_equal(g, g0) and np.array_equal(s, s0)
        total += len(g)
    assert total > 2000
    ens.close()


def test_ensemble_launch_shape_does_not_change_results_and_simulate_api():
    from brainpy.state_amd import units as u
    sims, _ = brunel_exp_ensemble(4, n_min=300, n_max=900, seed=7)
    a = Ensemble(sims, tile_size=64)
    ra = a.simulate(100. * u.ms)
    a.close()
    b = Ensemble(sims, tile_size=256, max_window=4)
    rb = b.simulate(100. * u.ms)
    rb2 = b.cont(50. * u.ms)
    b.close()
    for x, y, sim in zip(ra, rb, sims):
        rid = next(iter(sim._taps))
        assert np.array_equal(x._rec[rid], y._rec[rid])
    assert rb2[0].times.to_decimal(u.ms)[0] == pytest.approx(100.0)


def test_autotune_returns_a_timed_candidate():
    sims, _ = brunel_exp_ensemble(3, n_min=300, n_max=600, seed=9)
    best, table = Ensemble.autotune(sims, candidates=[(64, 0), (256, 0), (128, 4)], steps=100)
    assert best in table and len(table) == 3
    assert all(v > 0 for v in table.values())


def test_device_connectivity_ensemble_equals_plain_runs():
    """Networks whose edges were sampled on the device: one ensemble engine
    built from the device arrays equals each network's own plain run."""
    sims, _ = brunel_exp_ensemble(5, n_min=300, n_max=1200, seed=11, connectivity='device')
    ens = Ensemble(sims)
    parts = ens.run(1500)
    ens.close()
    total = 0
    for sim, (vm, g, s) in zip(sims, parts):
        g0, s0 = _plain(sim, 1500)
        assert np.array_equal(g, g0) and np.array_equal(s, s0)
        total += len(g)
    assert total > 1000


def _current_driven(k):
    from brainpy.state_amd import (Simulator, ac_generator, dc_generator, fixed_indegree, iaf_psc_exp,
                                   spike_recorder, step_current_generator)
    sim = Simulator(dt=0.1)
    a = sim.create(iaf_psc_exp, 200 + 60 * k)
    b = sim.create(iaf_psc_exp, 90)
    sim.connect(sim.create(dc_generator, amplitude=300.0 + 20 * k), a)
    sim.connect(sim.create(step_current_generator, amplitude_times=[0.0, 40.0 + k],
                           amplitude_values=[350.0, 150.0]), b)
    sim.connect(sim.create(ac_generator, amplitude=200.0, offset=50.0, frequency=20.0 + 7 * k,
                           phase=15.0 * k, start=5.0), a[: 100], weight=1.1)
    sim.connect(a, b, weight=60.0, delay=1.0, rule=fixed_indegree(20), seed=k + 1)
    sim.connect(b, a, weight=-120.0, delay=1.5, rule=fixed_indegree(10), seed=k + 11)
    for pop in (a, b):
        sim.connect(pop, sim.create(spike_recorder))
    return sim


def test_current_generators_in_an_ensemble_equal_plain_runs():
    """dc, step and ac current channels keep their meaning when the networks'
    descriptors are merged into one engine."""
    sims = [_current_driven(k) for k in range(3)]
    ens = Ensemble(sims)
    parts = ens.run(1500)
    total = 0
    for sim, (vm, g, s) in zip(sims, parts):
        g0, s0 = _plain(sim, 1500)
        assert np.array_equal(g, g0) and np.array_equal(s, s0)
        total += len(g)
    assert total > 500
    ens.close()
