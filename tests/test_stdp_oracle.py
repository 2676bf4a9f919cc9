"""stdp_synapse on the CPU oracle against a closed-form pair-timing trajectory.

Two neurons fire on fixed schedules (a supra-threshold I_e makes each fire on
its first non-refractory step: periods ref_count + 1). One plastic edge with a
5-step delay connects them. The expected weight follows the reference rule
(stdp_synapse.py:163-179) with traces as closed-form sums of exponentials:
K+ at step n sums P+^(n - a) over the edge's arrivals a <= n, K- sums
P-^(n - 1 - s) over post spikes s <= n - 1; the same step's own spike is
excluded on each side.
"""
import math

import pytest

import oracle
from brainpy.state_amd import Simulator, iaf_psc_delta, one_to_one, spike_recorder, stdp_synapse
from brainpy.state_amd.init import Constant

DT = 0.1


def build(mu_plus, mu_minus, lam=0.1, alpha=1.05, w0=10.0, Wmax=100.0, delay=0.5):
    sim = Simulator(dt=DT)
    pre = sim.create(iaf_psc_delta, 1, E_L=0., V_th=1., V_reset=0., I_e=1e6, t_ref=2.0,
                     V_initializer=Constant(0.))
    post = sim.create(iaf_psc_delta, 1, E_L=0., V_th=1., V_reset=0., I_e=1e6, t_ref=1.3,
                      V_initializer=Constant(0.))
    syn = stdp_synapse(weight=w0, delay=delay, tau_plus=20., tau_minus=15., lambda_=lam,
                       alpha=alpha, mu_plus=mu_plus, mu_minus=mu_minus, Wmax=Wmax)
    proj = sim.connect(pre, post, rule=one_to_one, synapse=syn)
    for pop in (pre, post):
        sim.connect(pop, sim.create(spike_recorder))
    return sim, proj


def expected(T, mu_plus, mu_minus, lam=0.1, alpha=1.05, w0=10.0, Wmax=100.0, d=5):
    pre_sp = set(range(0, T, 21))     # ceil(2.0 / 0.1) + 1
    post_sp = set(range(0, T, 14))    # ceil(1.3 / 0.1) + 1
    arr = sorted(s + 1 + d for s in pre_sp)
    Pp, Pm = math.exp(-DT / 20.), math.exp(-DT / 15.)
    w = w0
    for n in range(T):
        x = 1.0 if n in arr else 0.0
        kp = sum(Pp ** (n - a) for a in arr if a <= n) - x
        pf = 1.0 if (n - 1) in post_sp else 0.0
        km = sum(Pm ** (n - 1 - s) for s in post_sp if s <= n - 1) - pf
        if pf:
            nw = w / Wmax + lam * (1.0 - w / Wmax) ** mu_plus * kp
            w = nw * Wmax if nw < 1.0 else Wmax
        if x:
            nw = w / Wmax - alpha * lam * (w / Wmax) ** mu_minus * km
            w = nw * Wmax if nw > 0.0 else 0.0
    return w


@pytest.mark.parametrize('mu', [(0.0, 0.0), (1.0, 1.0), (0.4, 0.7)])
def test_stdp_pair_trajectory_matches_closed_form(mu):
    T = 300
    sim, proj = build(*mu)
    o = oracle.OracleNet(sim.build_desc())
    _, g, s = o.run(T)
    assert sorted(set(s[g == 0].tolist())) == list(range(0, T, 21))
    assert sorted(set(s[g == 1].tolist())) == list(range(0, T, 14))
    w = o.weights(0, 1)[0]
    assert w == pytest.approx(expected(T, *mu), rel=1e-11)
    assert w != 10.0


def test_stdp_clamps():
    sim, proj = build(0.0, 0.0, lam=5.0)
    o = oracle.OracleNet(sim.build_desc())
    o.run(300)
    w = o.weights(0, 1)[0]
    assert w == pytest.approx(expected(300, 0.0, 0.0, lam=5.0), abs=1e-12)
    assert w in (0.0, 100.0) or 0.0 < w < 100.0


# ---- the reference's single-edge known answers (stdp_synapse_rule_test.py:65-131),
# with the expected weights worked by hand from the rule's closed form.
def test_edge_potentiation_on_post_spike():
    # nw = 0.1 + 0.1 * 0.9 * 0.5 = 0.145
    assert oracle.stdp_edge(10., 0., 1., 0.5, 0., lambda_=0.1) == pytest.approx(14.5, rel=1e-14)


def test_edge_depression_on_pre_spike():
    # nw = 0.1 - 0.1 * 0.1 * 0.5 = 0.095
    assert oracle.stdp_edge(10., 1., 0., 0., 0.5, lambda_=0.1) == pytest.approx(9.5, rel=1e-14)


def test_edge_no_spike_no_change():
    assert oracle.stdp_edge(10., 0., 0., 0.7, 0.7) == 10.0


def test_edge_simultaneous_excludes_own_spike():
    assert oracle.stdp_edge(10., 1., 1., 1., 1., lambda_=0.1) == pytest.approx(10.0, rel=1e-14)
    # prior post trace 1.0: depression sees K- = 1 -> 0.1 - 0.01 = 0.09
    assert oracle.stdp_edge(10., 1., 1., 1., 2., lambda_=0.1) == pytest.approx(9.0, rel=1e-14)


def test_edge_clamps():
    assert oracle.stdp_edge(99., 0., 1., 10., 0., lambda_=5.) == 100.0
    assert oracle.stdp_edge(1., 1., 0., 0., 10., lambda_=5.) == 0.0


def test_edge_mu_zero_additive():
    # nw = 0.1 + 0.1 * 1 * 0.5 = 0.15 (no weight dependence)
    w = oracle.stdp_edge(10., 0., 1., 0.5, 0., lambda_=0.1, mu_plus=0., mu_minus=0.)
    assert w == pytest.approx(15.0, rel=1e-14)
    # multiplicative: depression scales with w/Wmax
    w1 = oracle.stdp_edge(50., 1., 0., 0., 1., lambda_=0.1, alpha=2.)
    assert w1 == pytest.approx(100. * (0.5 - 2 * 0.1 * 0.5 * 1.0), rel=1e-14)
