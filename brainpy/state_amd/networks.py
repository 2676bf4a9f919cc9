"""Standard network builders on the Simulator API.

``brunel_delta`` is Brunel's (2000) balanced random network of
``iaf_psc_delta`` neurons as ``examples/nest_like/brunel_delta.py:28-61`` wires
it (Poisson drive at ``eta * nu_th``, fixed in-degree ``epsilon * N`` with
multapses and autapses, inhibition ``-g J``), with the parameters of the
benchmark workloads: tau_m = 20 ms, V_th = 20 mV, V_reset = 10 mV,
t_ref = 2 ms, C_m = 250 pF, J = 0.1 mV, g = 5, eta = 2, dt = 0.1 ms.

``delay`` is either a scalar (ms) or ``('uniform', lo_ms, hi_ms)``: per-edge
delays drawn uniformly from the integer multiples of dt in [lo, hi] with a
seeded NumPy PCG64 stream.
"""
from __future__ import annotations

import numpy as np

from . import units as u
from .devices import poisson_generator, spike_recorder, voltmeter
from .dist import UniformGrid
from .init import Constant
from .models import aeif_cond_alpha, iaf_psc_delta, iaf_psc_exp
from .plasticity import stdp_synapse
from .rules import fixed_indegree, pairwise_bernoulli
from .simulator import Simulator

__all__ = ['brunel_delta', 'brunel_exp']


def _edge_delays(sim, rule_K, n_post, lo, hi, seed):
    dt = sim.dt
    lo_s, hi_s = int(round(lo / dt)), int(round(hi / dt))
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.integers(lo_s, hi_s + 1, size=rule_K * n_post).astype(np.float64) * dt


def brunel_delta(order=200, *, NE=None, NI=None, epsilon=0.1, g=5.0, eta=2.0, J=0.1,
                 delay=1.5, dt=0.1, tau_m=20.0, V_th=20.0, V_reset=10.0, t_ref=2.0, C_m=250.0,
                 seed_conn=(1, 2), seed_delay=3, n_vm=10, record_all=True, dtype='float64',
                 max_window=0, connectivity='host'):
    """Brunel-A network of iaf_psc_delta neurons; returns ``(sim, handles)``.
    ``connectivity='device'`` samples the fixed in-degree edges on the device."""
    NE = 4 * order if NE is None else int(NE)
    NI = order if NI is None else int(NI)
    CE, CI = int(epsilon * NE), int(epsilon * NI)
    nu_th = V_th / (J * CE * tau_m)
    p_rate = 1000.0 * (eta * nu_th) * CE
    npar = dict(C_m=C_m * u.pF, tau_m=tau_m * u.ms, t_ref=t_ref * u.ms, E_L=0. * u.mV,
                V_reset=V_reset * u.mV, V_th=V_th * u.mV, V_initializer=Constant(0. * u.mV))
    sim = Simulator(dt=dt * u.ms, dtype=dtype, max_window=max_window, connectivity=connectivity)
    ne = sim.create(iaf_psc_delta, NE, params=npar)
    ni = sim.create(iaf_psc_delta, NI, params=npar)
    noise = sim.create(poisson_generator, rate=p_rate * u.Hz)
    if isinstance(delay, tuple):
        _, lo, hi = delay
        dE = [_edge_delays(sim, CE, NE, lo, hi, seed_delay),
              _edge_delays(sim, CE, NI, lo, hi, seed_delay + 1)]
        dI = [_edge_delays(sim, CI, NE, lo, hi, seed_delay + 2),
              _edge_delays(sim, CI, NI, lo, hi, seed_delay + 3)]
        d_noise = lo * u.ms
    else:
        dE = dI = None
        d_noise = delay * u.ms
    sim.connect(noise, ne, weight=J * u.mV, delay=d_noise)
    sim.connect(noise, ni, weight=J * u.mV, delay=d_noise)
    if dE is None:
        pe = sim.connect(ne, ne + ni, weight=J * u.mV, delay=delay * u.ms,
                         rule=fixed_indegree(CE), comm='sparse', allow_multapses=True,
                         seed=seed_conn[0])
        pi = sim.connect(ni, ne + ni, weight=-g * J * u.mV, delay=delay * u.ms,
                         rule=fixed_indegree(CI), comm='sparse', allow_multapses=True,
                         seed=seed_conn[1])
    else:
        pe = [sim.connect(ne, post, weight=J * u.mV, delay=dd * u.ms, rule=fixed_indegree(CE),
                          comm='sparse', allow_multapses=True, seed=seed_conn[0])
              for post, dd in zip((ne, ni), dE)]
        pi = [sim.connect(ni, post, weight=-g * J * u.mV, delay=dd * u.ms,
                          rule=fixed_indegree(CI), comm='sparse', allow_multapses=True,
                          seed=seed_conn[1])
              for post, dd in zip((ne, ni), dI)]
    esr = sim.create(spike_recorder)
    isr = sim.create(spike_recorder)
    if record_all:
        sim.connect(ne, esr)
        sim.connect(ni, isr)
    vm = None
    if n_vm:
        vm = sim.create(voltmeter)
        sim.connect(vm, ne[:n_vm])
    return sim, dict(ne=ne, ni=ni, noise=noise, esr=esr, isr=isr, vm=vm, pe=pe, pi=pi,
                     CE=CE, CI=CI, p_rate=p_rate)


def brunel_exp(N, *, frac_e=0.8, epsilon=0.1, g=5.0, eta=2.0, J=20.0, delay=1.5, dt=0.1,
               tau_m=20.0, tau_syn=0.5, V_th=-50.0, E_L=-70.0, V_reset=-70.0, t_ref=2.0,
               C_m=250.0, seed=1, dtype='float32', record=True, sim=None, connectivity='host'):
    """Brunel network of iaf_psc_exp neurons (PSC amplitudes in pA).

    The external rate is ``eta * nu_th`` with ``nu_th`` the rate that drives the
    mean free membrane potential to threshold: ``(V_th - E_L) / (J * CE * tau_syn / C_m * tau_m)``
    (the PSC charge ``J * tau_syn`` deposited through ``C_m``).
    """
    NE = int(round(frac_e * N))
    NI = int(N) - NE
    CE, CI = max(1, int(epsilon * NE)), max(1, int(epsilon * NI))
    theta = V_th - E_L
    psp_area = J * tau_syn / C_m * tau_m          # mV ms per input spike
    nu_th = theta / (psp_area * CE)                # spikes / ms per input
    p_rate = 1000.0 * (eta * nu_th) * CE
    npar = dict(C_m=C_m, tau_m=tau_m, t_ref=t_ref, E_L=E_L, V_reset=V_reset, V_th=V_th,
                tau_syn_ex=tau_syn, tau_syn_in=tau_syn, V_initializer=Constant(E_L))
    if sim is None:
        sim = Simulator(dt=dt * u.ms, dtype=dtype, connectivity=connectivity)
    ne = sim.create(iaf_psc_exp, NE, params=npar)
    ni = sim.create(iaf_psc_exp, NI, params=npar)
    noise = sim.create(poisson_generator, rate=p_rate * u.Hz)
    sim.connect(noise, ne, weight=J * u.pA, delay=delay * u.ms)
    sim.connect(noise, ni, weight=J * u.pA, delay=delay * u.ms)
    sim.connect(ne, ne + ni, weight=J * u.pA, delay=delay * u.ms, rule=fixed_indegree(CE),
                comm='sparse', seed=seed)
    sim.connect(ni, ne + ni, weight=-g * J * u.pA, delay=delay * u.ms, rule=fixed_indegree(CI),
                comm='sparse', seed=seed + 1)
    esr = isr = None
    if record:
        esr = sim.create(spike_recorder)
        isr = sim.create(spike_recorder)
        sim.connect(ne, esr)
        sim.connect(ni, isr)
    return sim, dict(ne=ne, ni=ni, esr=esr, isr=isr, CE=CE, CI=CI, p_rate=p_rate)


def brunel_exp_ensemble(n_networks=256, *, n_min=500, n_max=20000, g_range=(3.0, 7.0),
                        eta_range=(0.9, 4.0), seed=0, dtype='float32', **kw):
    """Independent Brunel iaf_psc_exp networks of uneven size: N log-uniform in
    [n_min, n_max] (80 % E), 10 % fixed in-degree, g and nu_ext/nu_thr uniform
    per network, tau_syn = 0.5 ms. Returns ``(sims, params)``; advance them
    together with :class:`~brainpy.state_amd.Ensemble`."""
    rng = np.random.default_rng(seed)
    N = np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_networks)).astype(np.int64)
    g = rng.uniform(*g_range, n_networks)
    eta = rng.uniform(*eta_range, n_networks)
    sims = []
    for k in range(n_networks):
        sim, _ = brunel_exp(int(N[k]), g=float(g[k]), eta=float(eta[k]), seed=2 * (seed * n_networks + k) + 1,
                            dtype=dtype, **kw)
        sims.append(sim)
    return sims, dict(N=N, g=g, eta=eta)


def aeif_stdp(n=50000, *, frac_e=0.8, p=0.01, g_E=1.0, g_I=4.0, rate=8000.0, w_noise=1.0,
              delay=(1.0, 5.0), dt=0.1, seed=1, n_vm=10, plastic=True, tau_plus=20.0,
              lambda_=0.01, alpha=1.05, Wmax=3.0, mu=0.0, dtype='float64', record=True,
              connectivity='host'):
    """Conductance-based network of aeif_cond_alpha neurons (NEST defaults:
    a = 4 nS, b = 80.5 pA, tau_w = 144 ms, Delta_T = 2 mV; adaptive RKF45),
    random connectivity with probability ``p``, g_E / g_I in nS, per-edge delays
    uniform over the dt grid in ``delay`` ms, stdp_synapse (additive for
    ``mu = 0``) on E->E, Poisson drive ``rate`` Hz at ``w_noise`` nS on every
    neuron. Returns ``(sim, handles)``."""
    NE = int(round(frac_e * n))
    NI = int(n) - NE
    sim = Simulator(dt=dt * u.ms, dtype=dtype, connectivity=connectivity)
    ne = sim.create(aeif_cond_alpha, NE)
    ni = sim.create(aeif_cond_alpha, NI)
    noise = sim.create(poisson_generator, rate=rate * u.Hz)
    d = UniformGrid(delay[0], delay[1], dt)
    sim.connect(noise, ne + ni, weight=w_noise, delay=dt * u.ms, receptor_type=1)
    if plastic:
        syn = stdp_synapse(weight=g_E, tau_plus=tau_plus, lambda_=lambda_, alpha=alpha,
                           mu_plus=mu, mu_minus=mu, Wmax=Wmax)
        pee = sim.connect(ne, ne, rule=pairwise_bernoulli(p), synapse=syn, delay=d, seed=seed,
                          receptor_type=1, allow_autapses=False)
    else:
        pee = sim.connect(ne, ne, weight=g_E, rule=pairwise_bernoulli(p), delay=d, seed=seed,
                          receptor_type=1, allow_autapses=False)
    sim.connect(ne, ni, weight=g_E, rule=pairwise_bernoulli(p), delay=d, seed=seed + 1,
                receptor_type=1)
    sim.connect(ni, ne + ni, weight=g_I, rule=pairwise_bernoulli(p), delay=d, seed=seed + 2,
                receptor_type=2)
    esr = isr = vm = None
    if record:
        esr = sim.create(spike_recorder)
        isr = sim.create(spike_recorder)
        sim.connect(ne, esr)
        sim.connect(ni, isr)
    if n_vm:
        vm = sim.create(voltmeter)
        sim.connect(vm, ne[:n_vm])
    return sim, dict(ne=ne, ni=ni, noise=noise, pee=pee, esr=esr, isr=isr, vm=vm)
