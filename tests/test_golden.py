"""The committed golden fixtures (tests/golden, regenerated by make_golden.py):
the oracle still reproduces them on the CPU, and the engine matches them on
the GPU without running the oracle."""
import json
import os

import numpy as np
import pytest

import oracle
from brainpy.state_amd.networks import brunel_delta, brunel_exp

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_threefry_fixture_matches_oracle():
    kat = json.load(open(os.path.join(G, 'threefry_kat.json')))['threefry2x32_20']
    for k0, k1, c0, c1, o0, o1 in kat:
        assert oracle.threefry2x32(k0, k1, c0, c1) == (o0, o1)


def test_oracle_reproduces_brunel_fixtures():
    f = _load('brunel1000_delta_f64.npz')
    sim, _ = brunel_delta(order=200, n_vm=10, dtype='float64')
    vm, g, s = oracle.OracleNet(sim.build_desc()).run(2000)
    assert np.array_equal(g, f['gid']) and np.array_equal(s, f['step'])
    assert np.array_equal(vm, f['vm'])
    f2 = _load('brunel_exp_mixed_f32.npz')
    sim, _ = brunel_exp(375, seed=3)
    _, g, s = oracle.OracleNet(sim.build_desc()).run(1500)
    assert np.array_equal(g, f2['gid']) and np.array_equal(s, f2['step'])


@pytest.mark.gpu
def test_engine_matches_golden_fixtures():
    f = _load('brunel1000_delta_f64.npz')
    sim, _ = brunel_delta(order=200, n_vm=10, dtype='float64')
    eng = sim._ensure_engine()
    eng.reset()
    vm, (g, s) = eng.run(2000)
    assert len(f['gid']) > 1000
    assert np.array_equal(g, f['gid']) and np.array_equal(s, f['step'])
    np.testing.assert_allclose(vm, f['vm'], rtol=1e-12, atol=0)
    f2 = _load('brunel_exp_mixed_f32.npz')
    sim, _ = brunel_exp(375, seed=3)
    eng = sim._ensure_engine()
    eng.reset()
    _, (g, s) = eng.run(1500)
    assert len(f2['gid']) > 100
    assert np.array_equal(g, f2['gid']) and np.array_equal(s, f2['step'])
