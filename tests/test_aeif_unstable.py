"""aeif_cond_alpha's instability guard in the oracle: after an accepted
substep, V below -1e3 mV or |w| above 1e6 pA makes the step raise
'Numerical instability in aeif_cond_alpha dynamics.' (aeif_cond_alpha.py:
703-706, the guard in the integrator's event function). A strong enough
negative current drives one neuron past -1e3 mV in its first step with the
current; a weaker one leaves the run quiet. The engine reports the same
condition (tests/test_gpu_edge_cases.py)."""
import pytest

import oracle
from brainpy.state_amd import Simulator, aeif_cond_alpha, dc_generator, units as u, voltmeter


def unstable_net(amplitude):
    sim = Simulator(dt=0.1 * u.ms)
    pop = sim.create(aeif_cond_alpha, 20)
    sim.connect(sim.create(dc_generator, amplitude=amplitude * u.pA), pop[5:6])
    vm = sim.create(voltmeter)
    sim.connect(vm, pop[4:7])
    return sim


def test_strong_negative_current_raises():
    # dV/dt = I / C_m = -1e7 pA / 281 pF ~ -3.6e4 mV/ms: past -1e3 mV in 0.03 ms
    o = oracle.OracleNet(unstable_net(-1.0e7).build_desc())
    with pytest.raises(RuntimeError, match='Numerical instability in aeif_cond_alpha dynamics'):
        o.run(40)


def test_moderate_current_stays_quiet():
    vm, _, _ = oracle.OracleNet(unstable_net(-500.0).build_desc()).run(40)
    # the driven neuron (column 1) hyperpolarises, its neighbours stay at rest
    assert -1e3 < vm[-1, 1] < -75.0
    assert vm[-1, 0] == vm[-1, 2] and abs(vm[-1, 0] + 70.6) < 1e-3
