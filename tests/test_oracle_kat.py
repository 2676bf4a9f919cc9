"""Pinning the CPU oracle against the reference's own known answers (CPU only).

Each test cites the reference test it restates. The oracle's neuron updates are
driven one neuron at a time with explicit per-step inputs, as the reference's
unit tests drive ``model.update()``.
"""
import math

import numpy as np
import pytest

import oracle
from brainpy.state_amd import aeif_cond_alpha, iaf_psc_alpha, iaf_psc_delta, iaf_psc_exp
from brainpy.state_amd.init import Constant
from brainpy.state_amd.models import alpha_propagator_p31_p32, propagator_exp

DT = 0.1


# ------------------------------------------------------------- primitives
def test_threefry2x32_random123_known_answers():
    """Threefry-2x32-20 KATs published with Random123 (Salmon et al., SC'11)."""
    assert oracle.threefry2x32(0, 0, 0, 0) == (0x6b200159, 0x99ba4efe)
    assert oracle.threefry2x32(0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff) == (0x1cb996fc, 0xbb002be7)
    assert oracle.threefry2x32(0x13198a2e, 0x03707344, 0x243f6a88, 0x85a308d3) == (0xc4923a9c, 0x483df7a0)


def test_exp_log_pow_within_one_ulp_of_libm():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-700, 700, 3000), rng.uniform(-1, 1, 3000), [0.0, 1.0, -1.0]])
    for x in xs:
        e, r = oracle.exp(x), math.exp(x)
        assert abs(e - r) <= math.ulp(r), x
    ys = np.concatenate([10 ** rng.uniform(-300, 300, 3000), rng.uniform(0.5, 2.0, 2000)])
    for y in ys:
        lo, r = oracle.log(y), math.log(y)
        assert abs(lo - r) <= 2 * math.ulp(r) + 1e-300, y
    for y in ys[::7]:
        # exp(y log x): the relative error grows with |y log x| (one ulp of the
        # argument); the controller's arguments stay within |y log x| < 10
        p, r = oracle.pow(y, 0.2), y ** 0.2
        assert abs(p - r) <= 4e-16 * max(1.0, abs(0.2 * math.log(y))) * abs(r), y
    assert oracle.pow(0.0, 0.2) == 0.0 and oracle.pow(3.0, 0.0) == 1.0


@pytest.mark.parametrize('lam', [0.8, 2.0, 6.5, 40.0])
def test_poisson_sampler_moments(lam):
    """Mean and variance of the inversion sampler over 60k draws (Fano factor 1)."""
    x = np.array([oracle.poisson(7, s, e, lam) for s in range(60) for e in range(1000)])
    n = len(x)
    assert abs(x.mean() - lam) < 5 * math.sqrt(lam / n)
    assert abs(x.var() / lam - 1.0) < 0.05


def test_poisson_sampler_pmf_matches_scipy():
    """Empirical P(k) of the sampler vs scipy.stats.poisson.pmf (lam = 2)."""
    scipy_stats = pytest.importorskip('scipy.stats')
    x = np.array([oracle.poisson(11, s, e, 2.0) for s in range(200) for e in range(500)])
    n = len(x)
    for k in range(7):
        p = scipy_stats.poisson.pmf(k, 2.0)
        assert abs(np.mean(x == k) - p) < 5 * math.sqrt(p * (1 - p) / n), k


# --------------------------------------------------------------- propagators
def test_propagator_exp_closed_form_and_singular_limit():
    """P21 = tau_s tau_m / (C (tau_m - tau_s)) (e^{-h/tau_m} - e^{-h/tau_s})."""
    ts, tm, c, h = 2.0, 10.0, 250.0, 0.1
    ref = ts * tm / (c * (tm - ts)) * (math.exp(-h / tm) - math.exp(-h / ts))
    assert float(propagator_exp(ts, tm, c, h)) == pytest.approx(ref, rel=1e-12)
    # tau_s == tau_m: limit h/C e^{-h/tau}
    assert float(propagator_exp(10.0, 10.0, c, h)) == pytest.approx(h / c * math.exp(-h / 10.0), rel=1e-12)


def test_alpha_propagators_against_quadrature():
    """P31, P32 as integrals of the alpha kernel through the membrane filter."""
    from scipy.integrate import quad
    ts, tm, c, h = 2.0, 10.0, 250.0, 0.1
    p31, p32 = alpha_propagator_p31_p32(ts, tm, c, h)
    # y3(h) response to dI(0)=1 (pA/ms) and I(0)=1 pA with dI=0
    I_of_dI = lambda t: t * math.exp(-t / ts)          # current from unit dI
    I_of_I = lambda t: math.exp(-t / ts)               # current from unit I
    v31 = quad(lambda t: math.exp(-(h - t) / tm) * I_of_dI(t), 0, h, epsabs=1e-16)[0] / c
    v32 = quad(lambda t: math.exp(-(h - t) / tm) * I_of_I(t), 0, h, epsabs=1e-16)[0] / c
    assert float(p31) == pytest.approx(v31, rel=1e-9)
    assert float(p32) == pytest.approx(v32, rel=1e-9)


# ------------------------------------------------------------- iaf_psc_delta
def test_iaf_psc_delta_matches_nest_step_recurrence():
    """iaf_psc_delta_test.py:161-224 (refractory_input, deltas 2, 20, 0, 0, 0)."""
    m = iaf_psc_delta(1, E_L=-70., C_m=250., tau_m=10., t_ref=0.25, V_th=-55., V_reset=-70.,
                      I_e=0., refractory_input=True, V_initializer=Constant(-56.))
    nrn = oracle.Neuron(m, DT)
    vs, ss = [], []
    for dv in [2., 20., 0., 0., 0.]:
        ss.append(1.0 if nrn.step(dv) else 0.0)
        vs.append(nrn.V)
    # NEST recurrence restated from the reference test
    E_L, C, tau = -70.0, 250.0, 10.0
    V_th, V_reset = -55.0 - E_L, -70.0 - E_L
    P33 = math.exp(-0.1 / tau)
    P30 = (1.0 - P33) * tau / C
    refr = math.ceil(0.25 / 0.1)
    y3, y0, r, buf = -56.0 - E_L, 0.0, 0, 0.0
    rv, rs = [], []
    for dv in [2., 20., 0., 0., 0.]:
        if r == 0:
            y3 = P30 * y0 + P33 * y3 + dv + buf
            buf = 0.0
        else:
            buf += dv * math.exp(-r * 0.1 / tau)
            r -= 1
        sp = 1.0 if y3 >= V_th else 0.0
        if sp:
            r, y3 = refr, V_reset
        rv.append(y3 + E_L)
        rs.append(sp)
    assert ss == rs
    assert all(abs(a - b) < 1e-9 for a, b in zip(vs, rv))


def test_iaf_psc_delta_refractory_discards_inputs():
    """iaf_psc_delta_test.py:88-120: inputs during the 3 refractory steps are lost."""
    m = iaf_psc_delta(1, E_L=-70., V_th=-55., V_reset=-70., tau_m=10., t_ref=0.25,
                      V_initializer=Constant(-56.))
    nrn = oracle.Neuron(m, DT)
    assert nrn.step(2.0)
    assert nrn.V == -70.0
    for _ in range(3):
        assert not nrn.step(20.0)
    nrn.step(0.0)
    assert nrn.V == pytest.approx(-70.0, abs=1e-12)


def test_iaf_psc_delta_passive_decay_closed_form():
    """No input: V_n - E_L = (V_0 - E_L) decay^n (each step one multiply)."""
    m = iaf_psc_delta(1, E_L=-65., tau_m=20., V_th=100., V_initializer=Constant(-50.))
    nrn = oracle.Neuron(m, DT)
    for _ in range(200):
        nrn.step(0.0)
    assert nrn.V - (-65.0) == pytest.approx(15.0 * math.exp(-200 * DT / 20.0), rel=1e-12)


# --------------------------------------------------------------- iaf_psc_exp
def test_iaf_psc_exp_dc_first_spike_4_8_ms():
    """iaf_psc_exp_test.py:84-103: I_e = 1000 pA from -70 mV fires at 4.8 ms."""
    nrn = oracle.Neuron(iaf_psc_exp(1, I_e=1000., V_initializer=Constant(-70.)), DT)
    first = None
    for k in range(80):
        if nrn.step(0.0) and first is None:
            first = (k + 1) * DT
    assert first == pytest.approx(4.8, abs=1e-12)


def test_iaf_psc_exp_psc_kernel_shape():
    """A 100 pA event decays as exp(-t/tau_syn) and charges V through P21."""
    m = iaf_psc_exp(1, E_L=0., V_th=1e6, V_reset=-1., tau_syn_ex=2., V_initializer=Constant(0.))
    nrn = oracle.Neuron(m, DT)
    nrn.step(100.0)
    assert nrn.s[1] == 100.0 and nrn.V == 0.0      # V sees the PSC one step later
    nrn.step(0.0)
    p21 = float(propagator_exp(2.0, 10.0, 250.0, DT))
    assert nrn.V == pytest.approx(100.0 * p21, rel=1e-15)
    for _ in range(9):
        nrn.step(0.0)
    assert nrn.s[1] == pytest.approx(100.0 * math.exp(-1.0 / 2.0), rel=1e-13)


# ------------------------------------------------------------- iaf_psc_alpha
NEST_ALPHA_VM = {0.1: -69.6020, 0.2: -69.2079, 0.3: -68.8178, 0.4: -68.4316, 0.5: -68.0492,
                 4.3: -56.0204, 4.4: -55.7615, 4.5: -55.5051, 4.6: -55.2513, 4.7: -55.0001,
                 4.8: -70.0000, 4.9: -70.0000, 5.0: -70.0000}


def test_iaf_psc_alpha_matches_nest_testsuite_trace():
    """iaf_psc_alpha_test.py:200-245: NEST V_m values (+-8e-4) and first spike at 4.8 ms."""
    nrn = oracle.Neuron(iaf_psc_alpha(1, I_e=1000., V_initializer=Constant(-70.)), DT)
    first = None
    for k in range(80):
        sp = nrn.step(0.0)
        t = round((k + 1) * DT, 1)
        if sp and first is None:
            first = t
        if t in NEST_ALPHA_VM:
            assert nrn.V == pytest.approx(NEST_ALPHA_VM[t], abs=8e-4), t
    assert first == pytest.approx(4.8, abs=1e-12)


# ----------------------------------------------------------- aeif_cond_alpha
def test_aeif_current_input_one_step_delay():
    """aeif_cond_alpha_test.py:259-279: I_stim = 100 pA gives V = 0.0355871886... after a step."""
    m = aeif_cond_alpha(1, V_th=1e6, V_peak=1e6, Delta_T=0.0, g_L=0.0, a=0.0, b=0.0, I_e=0.0,
                        V_initializer=Constant(0.0), V_reset=0.0)
    nrn = oracle.Neuron(m, DT)
    nrn.s[7] = 100.0          # the current latched at the previous step
    nrn.step()
    assert nrn.V == pytest.approx(0.03558718861209964, abs=1e-12)


def _aeif_ref_rhs(y, refr, i_stim, p):
    v, dgex, gex, dgin, gin, w = y
    veff = p['V_reset'] if refr else min(v, p['V_peak'])
    dts = 1.0 if p['Delta_T'] == 0.0 else p['Delta_T']
    isp = 0.0 if p['Delta_T'] == 0.0 else p['g_L'] * p['Delta_T'] * math.exp((veff - p['V_th']) / dts)
    dv = 0.0 if refr else (-p['g_L'] * (veff - p['E_L']) + isp - gex * (veff - p['E_ex'])
                           - gin * (veff - p['E_in']) - w + p['I_e'] + i_stim) / p['C_m']
    return np.array([dv, -dgex / p['tau_ex'], dgex - gex / p['tau_ex'], -dgin / p['tau_in'],
                     dgin - gin / p['tau_in'], (p['a'] * (veff - p['E_L']) - w) / p['tau_w']])


def _aeif_ref_step(y, r, h, i_stim, p, w_step):
    """The reference test's independent RKF45 integrator (aeif_cond_alpha_test.py:66-190)."""
    t, min_h, atol = 0.0, 1e-8, p['atol']
    h = max(h, min_h)
    spikes = 0
    for _ in range(100000):
        if not t < DT:
            break
        h = max(min_h, min(h, DT - t))
        refr = r > 0
        k1 = _aeif_ref_rhs(y, refr, i_stim, p)
        k2 = _aeif_ref_rhs(y + h * (1 / 4) * k1, refr, i_stim, p)
        k3 = _aeif_ref_rhs(y + h * (3 * k1 / 32 + 9 * k2 / 32), refr, i_stim, p)
        k4 = _aeif_ref_rhs(y + h * (1932 * k1 / 2197 - 7200 * k2 / 2197 + 7296 * k3 / 2197), refr, i_stim, p)
        k5 = _aeif_ref_rhs(y + h * (439 * k1 / 216 - 8 * k2 + 3680 * k3 / 513 - 845 * k4 / 4104), refr, i_stim, p)
        k6 = _aeif_ref_rhs(y + h * (-8 * k1 / 27 + 2 * k2 - 3544 * k3 / 2565 + 1859 * k4 / 4104
                                    - 11 * k5 / 40), refr, i_stim, p)
        y4 = y + h * (25 * k1 / 216 + 1408 * k3 / 2565 + 2197 * k4 / 4104 - k5 / 5)
        y5 = y + h * (16 * k1 / 135 + 6656 * k3 / 12825 + 28561 * k4 / 56430 - 9 * k5 / 50 + 2 * k6 / 55)
        err = float(np.max(np.abs(y5 - y4)))
        if err <= atol or h <= min_h:
            ya = y5.copy()
            if r > 0:
                ya[0] = p['V_reset']
            if r <= 0 and ya[0] >= p['V_peak']:
                ya[0] = p['V_reset']
                ya[5] += p['b']
                spikes += 1
                if p['refn'] > 0:
                    r = p['refn'] + 1
            t += h
            y = ya
            fac = 5.0 if err == 0.0 else min(5.0, max(0.2, 0.9 * (atol / max(err, 1e-30)) ** 0.2))
            h = max(min_h, h * fac)
        else:
            h = max(min_h, h * min(1.0, max(0.2, 0.9 * (atol / max(err, 1e-30)) ** 0.25)))
    if r > 0:
        r -= 1
    if w_step >= 0:
        y[1] += math.e / p['tau_ex'] * w_step
    else:
        y[3] += math.e / p['tau_in'] * (-w_step)
    return y, r, h, spikes


def test_aeif_rkf45_matches_independent_reference_integrator():
    """aeif_cond_alpha_test.py:281-395: 60 steps with currents and conductance jumps, atol 2e-6."""
    prm = dict(V_peak=0.0, V_reset=-58.0, t_ref=0.25, g_L=11.0, C_m=200.0, E_ex=0.0, E_in=-85.0,
               E_L=-70.0, Delta_T=2.0, tau_w=300.0, a=3.0, b=40.0, V_th=-50.0, tau_syn_ex=0.2,
               tau_syn_in=2.0, I_e=1200.0, gsl_error_tol=1e-6)
    m = aeif_cond_alpha(1, **prm, V_initializer=Constant(-68.0), w_initializer=Constant(5.0))
    nrn = oracle.Neuron(m, DT)
    x_seq = [0.0, 20.0, 0.0, -30.0, 0.0, 40.0, 0.0, 0.0, -10.0, 0.0, 0.0, 0.0] + [0.0] * 48
    w_seq = [0.0, 5.0, -2.0, 0.0, 4.0, -3.0, 0.0, 0.0, 1.0, 0.0, 0.0, -2.5] + [0.0] * 48
    p = dict(V_peak=0.0, V_reset=-58.0, g_L=11.0, C_m=200.0, E_ex=0.0, E_in=-85.0, E_L=-70.0,
             Delta_T=2.0, tau_w=300.0, a=3.0, b=40.0, V_th=-50.0, tau_ex=0.2, tau_in=2.0,
             I_e=1200.0, atol=1e-6, refn=math.ceil(0.25 / 0.1))
    y = np.array([-68.0, 0.0, 0.0, 0.0, 0.0, 5.0])
    r, h, i_stim = 0, 0.1, 0.0
    any_spike = False
    for x, w in zip(x_seq, w_seq):
        y, r, h, ns = _aeif_ref_step(y, r, h, i_stim, p, w)
        i_stim = x
        sp = nrn.step(max(w, 0.0), max(-w, 0.0))
        nrn.s[7] = x                  # the current enters at the next step
        assert (ns > 0) == sp
        any_spike |= sp
        np.testing.assert_allclose(nrn.s[:6], y, atol=2e-6, rtol=0)
        assert nrn.r.value == r
        assert nrn.s[6] == pytest.approx(h, abs=2e-6)
    assert any_spike


# ------------------------------------------------------------ network level
def test_input_delay_latency_is_delay_plus_holder_lag():
    """delay_test.py (out[t] = in[t-d]) plus the Simulator's one-step holder lag:
    a spike at step s moves the target at step s + 1 + d."""
    from brainpy.state_amd import Simulator, one_to_one, spike_recorder
    sim = Simulator(dt=DT)
    src = sim.create(iaf_psc_delta, 1, E_L=0., V_th=1.0, V_reset=0., I_e=0.,
                     V_initializer=Constant(5.0), t_ref=100.)   # fires at step 0 only
    dst = sim.create(iaf_psc_delta, 1, E_L=0., V_th=1.0, V_reset=0., t_ref=100.,
                     V_initializer=Constant(0.0))
    sim.connect(src, dst, rule=one_to_one, weight=2.0, delay=0.7)   # 7 steps
    r1 = sim.create(spike_recorder)
    r2 = sim.create(spike_recorder)
    sim.connect(src, r1)
    sim.connect(dst, r2)
    o = oracle.OracleNet(sim.build_desc())
    _, g, s = o.run(30)
    assert list(zip(g.tolist(), s.tolist())) == [(0, 0), (1, 8)]


def test_event_conservation_delivered_equals_spikes_times_outdegree():
    """Sum of delivered events equals spikes x out-degree (homogeneous weight 1 mV)."""
    from brainpy.state_amd import Simulator, fixed_indegree, voltmeter
    sim = Simulator(dt=DT)
    src = sim.create(iaf_psc_delta, 50, E_L=0., V_th=1.0, V_reset=0., t_ref=1000.,
                     V_initializer=Constant(5.0))   # every source fires once, at step 0
    dst = sim.create(iaf_psc_delta, 40, E_L=0., V_th=1e9, tau_m=1e12,
                     V_initializer=Constant(0.0))
    sim.connect(src, dst, rule=fixed_indegree(7), weight=1.0, delay=0.3, seed=3)
    vm = sim.create(voltmeter)
    sim.connect(vm, dst)
    o = oracle.OracleNet(sim.build_desc())
    v, g, s = o.run(10)
    assert np.sum(v[-1]) == pytest.approx(50 * 7 * 40 / 50 * 1.0 * 1.0, rel=1e-12)


def test_fixed_point_delivery_is_the_exact_sum_within_quantisation():
    """Per-edge weights are summed as 2^-F fixed-point quanta and converted
    once. Pinned against math.fsum of the arriving weights: every post neuron
    of a passive population (decay exactly 1, E_L = 0, threshold out of
    reach) holds, after the single arrival step, exactly the delivered sum y;
    |y - fsum| must stay within k 2^-(F+1) (quantising k weights) plus half
    an ulp (the one conversion), for weights spread over 12 decades and both
    signs. A dropped, doubled or sign-flipped term breaks the bound."""
    from brainpy.state_amd import Simulator, all_to_all, spike_recorder
    from brainpy.state_amd import units as u
    sim = Simulator(dt=0.1 * u.ms, dtype='float64')
    n_pre, n_post = 60, 40
    pre = sim.create(iaf_psc_delta, n_pre, E_L=0., V_th=20., V_reset=0., tau_m=10., t_ref=1e6,
                     V_initializer=Constant(21.))
    post = sim.create(iaf_psc_delta, n_post, E_L=0., V_th=1e300, V_reset=0., tau_m=1e30, t_ref=2.,
                      V_initializer=Constant(0.))
    rng = np.random.default_rng(5)
    w = rng.standard_normal(n_pre * n_post) * 10.0 ** rng.uniform(-6, 6, n_pre * n_post)
    proj = sim.connect(pre, post, rule=all_to_all, weight=w * u.mV, delay=1.0 * u.ms)
    sim.connect(pre, sim.create(spike_recorder))
    desc = sim.build_desc()
    orc = oracle.OracleNet(desc)
    _, g, s = orc.run(14)
    assert len(g) == n_pre and np.all(s == 0)  # every pre neuron fires once, at step 0
    y = orc.state(1, 0, n_post)                # arrival at 0 + 1 + 10 = step 11
    F = desc.c.channels[desc.c.projs[desc.proj_index[id(proj)]].channel].fixed_shift
    pre_i, post_j = np.asarray(proj.pre_idx), np.asarray(proj.post_idx)
    wt = np.asarray(proj.weight, np.float64)
    for j in range(n_post):
        ws = wt[post_j == j]
        exact = math.fsum(ws.tolist())
        bound = len(ws) * 2.0 ** -(F + 1) + 0.5 * math.ulp(exact)
        assert abs(y[j] - exact) <= bound, (j, y[j], exact, bound)
        # and the quantisation bound is far below the magnitude of the sum
        assert bound <= 1e-12 * np.abs(ws).sum() + math.ulp(exact)


def _passive_post(sim, n):
    """iaf_psc_delta with decay exactly 1, E_L = 0 and no reachable
    threshold: V_m is the running sum of the delivered inputs."""
    return sim.create(iaf_psc_delta, n, E_L=0., V_th=1e300, V_reset=0., tau_m=1e30, t_ref=2.,
                      V_initializer=Constant(0.))


@pytest.mark.parametrize('comm', ['dense', 'sparse'])
def test_input_delay_fractional_frames_kat(comm):
    """delay_test.py:62-71: a delay of 1.5 steps reads 0.5 in[t-1] +
    0.5 in[t-2]. A neuron that fires once at step 0 reaches a passive
    target at steps 2 and 3 (one step of holder lag): with comm='dense'
    (x @ W, graded) as w/2 and w/2; with comm='sparse' (BinaryArray, any
    nonzero read is an event) as w and w. An axonal delay of 2.0 steps on the
    second source arrives whole at step 3."""
    from brainpy.state_amd import Simulator, spike_recorder, voltmeter
    from brainpy.state_amd import units as u
    sim = Simulator(dt=1.0 * u.ms, dtype='float64')
    pre = sim.create(iaf_psc_delta, 2, E_L=0., V_th=20., V_reset=0., tau_m=10., t_ref=1e6,
                     V_initializer=Constant(30.))
    post = _passive_post(sim, 2)
    sim.connect(pre[0], post[0], weight=3.0 * u.mV, delay=1.5 * u.ms, comm=comm)
    sim.connect(pre, post[1], weight=np.array([3.0, 5.0]) * u.mV, delay=np.array([1.5, 2.0]) * u.ms,
                comm=comm)
    sim.connect(pre, sim.create(spike_recorder))
    vm = sim.create(voltmeter)
    sim.connect(vm, post)
    desc = sim.build_desc()
    trace, g, s = oracle.OracleNet(desc).run(6)
    assert sorted(g.tolist()) == [0, 1] and np.all(s == 0)
    if comm == 'dense':
        np.testing.assert_array_equal(trace[:, 0], [0, 0, 1.5, 3.0, 3.0, 3.0])
        np.testing.assert_array_equal(trace[:, 1], [0, 0, 1.5, 3.0 + 5.0, 8.0, 8.0])
    else:
        np.testing.assert_array_equal(trace[:, 0], [0, 0, 3.0, 6.0, 6.0, 6.0])
        np.testing.assert_array_equal(trace[:, 1], [0, 0, 3.0, 6.0 + 5.0, 11.0, 11.0])


def test_input_delay_fractional_generator_kat():
    """A poisson_generator behind a 1.5-step delay (delay.py:114-128): the
    target receives (c[n-2] * 0.5 + c[n-3] * 0.5) * w, c[m] the generator's
    count of step m (zero outside its window), restated here from the
    KAT-pinned sampler; and at a whole-step delay, c[n-2] * w."""
    from brainpy.state_amd import Simulator, poisson_generator, voltmeter
    from brainpy.state_amd import units as u
    sim = Simulator(dt=1.0 * u.ms, dtype='float64')
    post = _passive_post(sim, 6)
    gen = sim.create(poisson_generator, rate=900. * u.Hz, start=3. * u.ms)
    p_frac = sim.connect(gen, post[:3], weight=2.0 * u.mV, delay=1.5 * u.ms)
    p_whole = sim.connect(gen, post[3:], weight=2.0 * u.mV, delay=1.0 * u.ms)
    vm = sim.create(voltmeter)
    sim.connect(vm, post)
    desc = sim.build_desc()
    T = 40
    trace, _, _ = oracle.OracleNet(desc).run(T)
    for k, (ch_i, frac) in enumerate([(0, 0.5), (1, 0.0)]):
        ch = desc.c.channels[ch_i]
        assert ch.delay_frac == frac and ch.delay_steps == 1

        def c(m, j):
            if m < 0 or not (ch.t_min_step < m <= ch.t_max_step):
                return 0.0
            return float(oracle.poisson(ch.seed, m, j, ch.lam))
        for j in range(3):
            v, ref = 0.0, []
            for n in range(T):
                y = (c(n - 2, j) * 0.5 + c(n - 3, j) * 0.5) * 2.0 if frac else c(n - 2, j) * 2.0
                v = v + y
                ref.append(v)
            np.testing.assert_array_equal(trace[:, 3 * k + j], ref)
        assert trace[-1, 3 * k:3 * k + 3].min() > 0
