"""CPU reference of the surrogate-gradient BPTT -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` and ``bench.py``'s ``cpu_baseline`` leg may use this module.

Two independent restatements of the recurrent iaf_psc_exp episode that
``bs_bptt_*`` computes on the device:

``forward_f32``
    NumPy float32, one IEEE operation at a time in the device's order
    (inputs summed sources-ascending then presynaptic-ascending; the update
    of ``_kernel/models/iaf_psc_exp.py:136-151``; the spike rule of
    ``_kernel/spiking.py:364-488``), so spikes and states match the device
    bit for bit.

``grad_autograd``
    PyTorch float64 autograd of the same episode with the hard decisions
    (spikes, refractory flags) taken from ``forward_f32``: the spike is
    ``hard + g(m) - stop_gradient(g(m))`` with ``g(m) = m / (1 + alpha |m|)``
    (so ``dg/dm`` is the fast sigmoid ``1 / (1 + alpha |m|)^2``; ``count_ste``,
    spiking.py:65-79), the reset is detached (bptt.py:276), the loss is
    ``rate_mse`` or ``membrane_mse`` (bptt.py:170-215). Reverse mode comes
    from torch, not from the adjoint equations the device implements, so the
    two gradient derivations are independent.

``loss_f64``
    Plain float64 forward loss for central finite differences on
    subthreshold episodes (no spikes: the loss is smooth there).
"""
from __future__ import annotations

import numpy as np

F = np.float32


def _consts(net):
    k = net.constants()
    c = {key: F(k[key]) for key in ('E_L', 'V_th', 'V_reset', 'I_e', 'P22', 'P21ex', 'P21in',
                                      'P11ex', 'P11in', 'P20', 'surrogate_alpha')}
    c['reset_to'] = k['ref_steps'] + 1 if k['ref_steps'] > 0 else 0
    return c


def forward_f32(net, w_in, w_rec, counts, target, v0=None, loss='rate_mse'):
    """Returns dict(spk bool [T,B,N], ref bool [T,B,N], V float32 [T,B,N]
    (after the step), counts float32 [B,N], loss float)."""
    c = _consts(net)
    T, B, _ = counts.shape
    N, nE = net.n, net.n_exc
    d, d_in = net.delay_steps, net.delay_in_steps
    V = np.full((B, N), c['E_L'], F) if v0 is None else np.array(v0, F)
    ie = np.zeros((B, N), F)
    ii = np.zeros((B, N), F)
    r = np.zeros((B, N), np.int64)
    spk = np.zeros((T, B, N), bool)
    refs = np.zeros((T, B, N), bool)
    Vs = np.zeros((T, B, N), F)
    span = F(c['V_th'] - c['V_reset'])
    for n in range(T):
        xe = np.zeros((B, N), F)
        xi = np.zeros((B, N), F)
        if n - d_in >= 0:
            cn = counts[n - d_in]
            for i in range(net.n_in):
                col = cn[:, i]
                sel = col > 0
                if sel.any():
                    xe[sel] = xe[sel] + col[sel].astype(F)[:, None] * w_in[i][None, :]
        if n - d >= 0:
            s = spk[n - d]
            for p in np.nonzero(s.any(0))[0]:
                sel = s[:, p]
                if p < nE:
                    xe[sel] = xe[sel] + w_rec[p][None, :]
                else:
                    xi[sel] = xi[sel] + w_rec[p][None, :]
        ref = r > 0
        y = V - c['E_L']
        y_new = ((y * c['P22'] + ie * c['P21ex']) + ii * c['P21in']) + c['I_e'] * c['P20']
        y = np.where(ref, y, y_new).astype(F)
        ie = (ie * c['P11ex'] + xe).astype(F)
        ii = (ii * c['P11in'] + xi).astype(F)
        V = (y + c['E_L']).astype(F)
        m = ((V - c['V_th']) / span).astype(F)
        hard = (~ref) & (m >= 0)
        V = np.where(hard, c['V_reset'], V).astype(F)
        r = np.where(hard, c['reset_to'], r)
        r = np.maximum(r - 1, 0)
        spk[n], refs[n], Vs[n] = hard, ref, V
    cnt = spk.sum(0).astype(F)
    T_s = T * net.dt * 1e-3
    if loss == 'rate_mse':
        val = float(np.mean((cnt.astype(np.float64) / T_s - target.astype(np.float64)[None, :]) ** 2))
    else:
        val = float(np.mean((Vs.astype(np.float64) - target.astype(np.float64)[None, None, :]) ** 2))
    return dict(spk=spk, ref=refs, V=Vs, counts=cnt, loss=val)


def grad_autograd(net, w_in, w_rec, counts, target, fwd, v0=None, loss='rate_mse'):
    """(loss, g_in, g_rec) in float64 by torch autograd, hard decisions from ``fwd``."""
    import torch
    k = net.constants()
    T, B, _ = counts.shape
    N, nE = net.n, net.n_exc
    d, d_in = net.delay_steps, net.delay_in_steps
    dt64 = torch.float64
    Win = torch.tensor(np.asarray(w_in, np.float64), requires_grad=True)
    Wrec = torch.tensor(np.asarray(w_rec, np.float64), requires_grad=True)
    cnt = torch.tensor(counts.astype(np.float64))
    hard_all = torch.tensor(fwd['spk'])
    ref_all = torch.tensor(fwd['ref'])
    alpha = k['surrogate_alpha']
    span = k['V_th'] - k['V_reset']
    V = torch.full((B, N), k['E_L'], dtype=dt64) if v0 is None else torch.tensor(np.asarray(v0, np.float64))
    ie = torch.zeros((B, N), dtype=dt64)
    ii = torch.zeros((B, N), dtype=dt64)
    spikes = []
    total_v = torch.zeros((), dtype=dt64)
    tgt = torch.tensor(np.asarray(target, np.float64))
    for n in range(T):
        xe = torch.zeros((B, N), dtype=dt64)
        xi = torch.zeros((B, N), dtype=dt64)
        if n - d_in >= 0:
            xe = xe + cnt[n - d_in] @ Win
        if n - d >= 0:
            s = spikes[n - d]
            xe = xe + s[:, :nE] @ Wrec[:nE]
            xi = xi + s[:, nE:] @ Wrec[nE:]
        ref = ref_all[n]
        y = V - k['E_L']
        y_new = y * k['P22'] + ie * k['P21ex'] + ii * k['P21in'] + k['I_e'] * k['P20']
        y = torch.where(ref, y, y_new)
        ie = ie * k['P11ex'] + xe
        ii = ii * k['P11in'] + xi
        V = y + k['E_L']
        m = (V - k['V_th']) / span
        g = m / (1.0 + alpha * m.abs())
        hard = hard_all[n]
        s = hard.to(dt64) + (g - g.detach()) * (~ref).to(dt64)
        spikes.append(s)
        V = torch.where(hard, torch.full_like(V, k['V_reset']), V)
        if loss == 'membrane_mse':
            total_v = total_v + ((V - tgt[None, :]) ** 2).sum()
    T_s = T * net.dt * 1e-3
    if loss == 'rate_mse':
        rate = torch.stack(spikes).sum(0) / T_s
        L = ((rate - tgt[None, :]) ** 2).mean()
    else:
        L = total_v / (T * B * N)
    L.backward()
    return float(L.detach()), Win.grad.numpy(), Wrec.grad.numpy()


def loss_f64(net, w_in, w_rec, counts, target, v0=None):
    """float64 membrane_mse of a subthreshold episode (asserts no spike)."""
    k = net.constants()
    T, B, _ = counts.shape
    N = net.n
    d_in = net.delay_in_steps
    V = np.full((B, N), k['E_L']) if v0 is None else np.asarray(v0, np.float64).copy()
    ie = np.zeros((B, N))
    ii = np.zeros((B, N))
    acc = 0.0
    w_in = np.asarray(w_in, np.float64)
    for n in range(T):
        xe = counts[n - d_in].astype(np.float64) @ w_in if n - d_in >= 0 else 0.0
        y = V - k['E_L']
        y = y * k['P22'] + ie * k['P21ex'] + ii * k['P21in'] + k['I_e'] * k['P20']
        ie = ie * k['P11ex'] + xe
        ii = ii * k['P11in']
        V = y + k['E_L']
        assert np.all(V < k['V_th']), 'episode is not subthreshold'
        acc += float(((V - np.asarray(target, np.float64)[None, :]) ** 2).sum())
    return acc / (T * B * N)
