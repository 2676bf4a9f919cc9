"""Pin the CPU BPTT reference (oracle/bptt_ref.py) before it checks the device.

* the forward reproduces the NEST iaf_psc_exp DC known answer (iaf_psc_exp_test.py:84-103)
  and holds V_m at V_reset for t_ref after the spike;
* a single input event charges V along the closed-form exp-PSC LIF response;
* torch-autograd gradients equal central finite differences of an
  independent float64 forward on subthreshold episodes (membrane_mse);
* the rate_mse surrogate gradient of one subthreshold neuron equals its
  closed form sum_n dL/ds * g'(m_n) / span * dV_n/dw.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip('torch')

from oracle import bptt_ref as R
from brainpy.state_amd.bptt import RecurrentIafExp, poisson_counts, random_sparse_weights


def _one(I_e=0.0, **kw):
    return RecurrentIafExp(1, 1, 0, dt=0.1, delay=1.0, delay_in=1.0, I_e=I_e, **kw)


def test_dc_first_spike_4_8_ms_and_refractory_hold():
    net = _one(I_e=1000.0)
    T = 80
    fwd = R.forward_f32(net, np.zeros((1, 1), np.float32), np.zeros((1, 1), np.float32),
                        np.zeros((T, 1, 1), np.uint8), np.zeros(1, np.float32))
    first = np.nonzero(fwd['spk'][:, 0, 0])[0][0]
    assert (first + 1) * 0.1 == pytest.approx(4.8, abs=1e-9)
    V = fwd['V'][:, 0, 0]
    # refractory for t_ref = 2 ms = 20 steps: V stays at V_reset, then charges again
    assert np.all(V[first:first + 21] == np.float32(-70.0))
    assert V[first + 21] > -70.0
    assert fwd['ref'][first + 1:first + 21, 0, 0].all() and not fwd['ref'][first + 21, 0, 0]


def _psc_response(w, n, d_in, p, h=0.1):
    """Closed form of V - E_L for an exp-PSC LIF after a jump w of I_ex at step d_in."""
    if n <= d_in:
        return 0.0
    t = (n - d_in) * h
    ts, tm, c = p['tau_syn_ex'], p['tau_m'], p['C_m']
    return w * ts * tm / (c * (tm - ts)) * (math.exp(-t / tm) - math.exp(-t / ts))


def test_single_event_follows_closed_form_psc_response():
    net = _one()
    T, w = 200, 500.0
    cnt = np.zeros((T, 1, 1), np.uint8)
    cnt[0, 0, 0] = 1
    fwd = R.forward_f32(net, np.full((1, 1), w, np.float32), np.zeros((1, 1), np.float32), cnt,
                        np.zeros(1, np.float32))
    V = fwd['V'][:, 0, 0].astype(np.float64) + 70.0
    ref = np.array([_psc_response(w, n, net.delay_in_steps, net.p) for n in range(T)])
    np.testing.assert_allclose(V, ref, atol=2e-5 * np.abs(ref).max())
    assert V.max() > 1.0


def _subthreshold_case(seed=0):
    net = RecurrentIafExp(6, 5, 3, dt=0.1, delay=1.0, delay_in=0.5)
    T, B = 120, 2
    cnt = poisson_counts(np.full(6, 300.0), T, B, 0.1, seed)
    w_in, _ = random_sparse_weights(6, 8, 0.6, 40.0, 10.0, seed)
    w_rec, _ = random_sparse_weights(8, 8, 0.5, 10.0, 3.0, seed + 1)
    target = np.full(8, -66.0, np.float32)
    return net, cnt, w_in, w_rec, target


def test_autograd_membrane_gradient_matches_finite_differences():
    # Recurrent weights zero: with count_ste the surrogate spikes of a silent
    # network still carry gradient through W_rec, a path the hard forward
    # (and so a finite difference) does not have.
    net, cnt, w_in, w_rec, target = _subthreshold_case()
    w_rec = np.zeros_like(w_rec)
    fwd = R.forward_f32(net, w_in, w_rec, cnt, target, loss='membrane_mse')
    assert not fwd['spk'].any()
    L, g_in, g_rec = R.grad_autograd(net, w_in, w_rec, cnt, target, fwd, loss='membrane_mse')
    assert L == pytest.approx(R.loss_f64(net, w_in, w_rec, cnt, target), rel=1e-12)
    assert not np.any(g_rec)                   # no spikes: recurrent weights do not act
    eps = 1e-3
    for (i, j) in [(0, 0), (1, 3), (2, 7), (5, 5), (3, 1)]:
        wp = w_in.astype(np.float64).copy()
        wm = wp.copy()
        wp[i, j] += eps
        wm[i, j] -= eps
        fd = (R.loss_f64(net, wp, w_rec, cnt, target) - R.loss_f64(net, wm, w_rec, cnt, target)) / (2 * eps)
        assert g_in[i, j] == pytest.approx(fd, rel=1e-6, abs=1e-12), (i, j)


def test_autograd_rate_surrogate_gradient_closed_form():
    """One neuron, one input event, no spike: dL/dw = sum_n c * g'(m_n) / span * K_n."""
    net = _one()
    T, w = 150, 900.0
    cnt = np.zeros((T, 1, 1), np.uint8)
    cnt[0, 0, 0] = 1
    target = np.array([7.0], np.float32)
    w_in = np.full((1, 1), w, np.float32)
    fwd = R.forward_f32(net, w_in, np.zeros((1, 1), np.float32), cnt, target)
    assert not fwd['spk'].any()
    L, g_in, _ = R.grad_autograd(net, w_in, np.zeros((1, 1), np.float32), cnt, target, fwd)
    T_s = T * 1e-4
    assert L == pytest.approx(49.0)
    dLds = 2.0 * (0.0 - 7.0) / T_s            # d mean((rate - 7)^2) / d spike at any step
    span, alpha = 15.0, 5.0 * 15.0
    tot = 0.0
    for n in range(T):
        K = _psc_response(1.0, n, net.delay_in_steps, net.p)
        m = (K * w - 15.0) / span               # V - V_th with E_L = -70, V_th = -55
        tot += dLds * (1.0 / (1.0 + alpha * abs(m)) ** 2) / span * K
    assert g_in[0, 0] == pytest.approx(tot, rel=2e-5)


def test_forward_loss_agrees_between_float32_and_float64_paths():
    net = RecurrentIafExp(10, 48, 16, dt=0.1, delay=1.0, delay_in=1.0)
    T, B = 300, 2
    cnt = poisson_counts(np.full(10, 400.0), T, B, 0.1, 3)
    w_in, _ = random_sparse_weights(10, 64, 0.5, 500.0, 100.0, 3)
    we, _ = random_sparse_weights(48, 64, 0.2, 40.0, 10.0, 4)
    wi, _ = random_sparse_weights(16, 64, 0.2, 120.0, 20.0, 5, sign=-1.0)
    w_rec = np.concatenate([we, wi])
    target = np.full(64, 20.0, np.float32)
    fwd = R.forward_f32(net, w_in, w_rec, cnt, target)
    assert fwd['spk'].sum() > 50
    L, g_in, g_rec = R.grad_autograd(net, w_in, w_rec, cnt, target, fwd)
    assert L == pytest.approx(fwd['loss'], rel=1e-12)   # rate loss depends on counts only
    assert np.abs(g_rec).sum() > 0 and np.abs(g_in).sum() > 0
