"""Surrogate-gradient BPTT of a recurrent ``iaf_psc_exp`` network on the GPU.

Restates, for one network family, the reference kernel learner
``BPTT._value_and_grad`` (``_kernel/learn/bptt.py:437-491``): the loss of an
episode rolled out with ``checkpointed_scan`` over ``Program.step``, and its
gradient with respect to the trainable projection weights. The model is the
kernel ``iaf_psc_exp`` (``_kernel/models/iaf_psc_exp.py:120-151``) with the
spike rule of ``_kernel/spiking.py:364-488`` (``count_ste`` forward/backward,
``detach_reset=True``) and a fast-sigmoid surrogate on the margin
``(V - V_th) / (V_th - V_reset)``; readouts ``rate_mse`` and ``membrane_mse``
(``bptt.py:170-215``).

The forward pass, the checkpoint recomputation, the reverse-time adjoint, the
transposed delayed delivery (one hand-written matrix-core GEMM per window,
``k_bptt_tdeliver``) and the weight
gradients all run on the device through ``bs_bptt_*`` of the C ABI. There is
no CPU path: without the engine or a device every call raises
:class:`~brainpy.state_amd._lib.EngineUnavailable`.

``torch`` only supplies device memory: weights, inputs and gradients are
``torch`` tensors on the HIP device; :class:`EpisodeLoss` wraps the pair of
calls as a ``torch.autograd.Function``.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from .models import propagator_exp

__all__ = ['RecurrentIafExp', 'BPTT', 'EpisodeLoss', 'poisson_counts', 'random_sparse_weights',
           'bptt_problem']

_NEST_IAF_PSC_EXP = dict(E_L=-70.0, C_m=250.0, tau_m=10.0, t_ref=2.0, V_th=-55.0, V_reset=-70.0,
                         tau_syn_ex=2.0, tau_syn_in=2.0, I_e=0.0)


class RecurrentIafExp:
    """A recurrent iaf_psc_exp network: ``n_in`` input sources, ``n_exc``
    excitatory then ``n_inh`` inhibitory neurons, one recurrent delay and one
    input delay (whole steps of ``dt``).

    Parameters take the kernel model's NEST defaults
    (``_kernel/models/iaf_psc_exp.py:106-111``); ``surrogate_slope`` is the
    fast sigmoid's slope in 1/mV, i.e. ``slope * (V_th - V_reset)`` on the
    dimensionless margin.
    """

    def __init__(self, n_in, n_exc, n_inh, *, dt=0.1, delay=1.0, delay_in=1.0,
                 surrogate_slope=5.0, **params):
        unknown = set(params) - set(_NEST_IAF_PSC_EXP)
        if unknown:
            raise KeyError(f'unknown iaf_psc_exp parameter(s) {sorted(unknown)}')
        self.p = dict(_NEST_IAF_PSC_EXP, **{k: float(v) for k, v in params.items()})
        self.n_in, self.n_exc, self.n_inh = int(n_in), int(n_exc), int(n_inh)
        self.n = self.n_exc + self.n_inh
        self.dt = float(dt)
        self.delay_steps = self._steps(delay, 'delay')
        self.delay_in_steps = self._steps(delay_in, 'delay_in')
        self.surrogate_slope = float(surrogate_slope)
        if not self.p['V_th'] > self.p['V_reset']:
            raise ValueError('V_th must exceed V_reset')

    def _steps(self, d, name):
        k = d / self.dt
        ki = int(round(k))
        if ki < 1 or abs(k - ki) > 1e-9 * max(1.0, k):
            raise ValueError(f'{name} must be a positive whole number of steps of dt')
        return ki

    def constants(self):
        """Propagators (float64, host) and thresholds of the update."""
        p, h = self.p, self.dt
        return dict(
            E_L=p['E_L'], V_th=p['V_th'], V_reset=p['V_reset'], I_e=p['I_e'],
            P22=math.exp(-h / p['tau_m']),
            P20=-p['tau_m'] / p['C_m'] * math.expm1(-h / p['tau_m']),
            P11ex=math.exp(-h / p['tau_syn_ex']), P11in=math.exp(-h / p['tau_syn_in']),
            P21ex=float(propagator_exp(p['tau_syn_ex'], p['tau_m'], p['C_m'], h)),
            P21in=float(propagator_exp(p['tau_syn_in'], p['tau_m'], p['C_m'], h)),
            # Refractory counter (spiking.py:124-146): ceil(t_ref/dt) grid steps
            ref_steps=int(round(p['t_ref'] / h)),
            surrogate_alpha=self.surrogate_slope * (p['V_th'] - p['V_reset']))


class BPTT:
    """Episode loss and weight gradients of a :class:`RecurrentIafExp`.

    Parameters
    ----------
    net : RecurrentIafExp
    batch, n_steps : int
        Trials per episode and steps per trial.
    checkpoint : int
        Steps between stored states; the backward pass recomputes each
        segment from its checkpoint (``checkpointed_scan``, bptt.py:455-458).
    loss : {'rate_mse', 'membrane_mse'}
    detach_reset : bool
        Only ``True`` (the learner's default, bptt.py:276) is implemented.
    gemm : {'bf16x3', 'fp32'}
        Arithmetic of the transposed delivery ``W_rec . ibar``: three bf16
        products of the hi/lo split of both operands accumulated in fp32
        (matrix-core rate, about 2^-16 relative error per product), or plain
        fp32.
    replay : {'auto', True, False}
        How the backward pass rebuilds a segment's surrogate factors from its
        checkpoint: True stores every step's summed inputs during the forward
        pass (8 bytes per neuron, trial and step) and replays the neuron steps
        from them; False recomputes the segment with its event gathers;
        'auto' stores when that takes at most 32 GiB. The gradients are the
        same bits either way.
    """

    def __init__(self, net, *, batch, n_steps, checkpoint=100, loss='rate_mse', detach_reset=True,
                 gemm='bf16x3', replay='auto'):
        if not detach_reset:
            raise NotImplementedError('only detach_reset=True is implemented')
        if loss not in ('rate_mse', 'membrane_mse'):
            raise ValueError(f"loss must be 'rate_mse' or 'membrane_mse', got {loss!r}")
        self.net, self.batch, self.n_steps = net, int(batch), int(n_steps)
        self.checkpoint, self.loss_name = int(checkpoint), loss
        k = net.constants()
        d = L.BpttDesc()
        d.abi_version = L.BS_ABI_VERSION
        d.n_in, d.n_exc, d.n_inh = net.n_in, net.n_exc, net.n_inh
        d.batch, d.n_steps = self.batch, self.n_steps
        d.delay_steps, d.delay_in_steps = net.delay_steps, net.delay_in_steps
        d.checkpoint = self.checkpoint
        d.loss = L.BS_LOSS_RATE_MSE if loss == 'rate_mse' else L.BS_LOSS_MEMBRANE_MSE
        d.ref_steps = k['ref_steps']
        if gemm not in ('bf16x3', 'fp32'):
            raise ValueError(f"gemm must be 'bf16x3' or 'fp32', got {gemm!r}")
        d.gemm = 0 if gemm == 'bf16x3' else 1
        if replay not in ('auto', True, False):
            raise ValueError(f"replay must be 'auto', True or False, got {replay!r}")
        d.replay = -1 if replay == 'auto' else int(bool(replay))
        d.dt_ms = net.dt
        for f in ('E_L', 'V_th', 'V_reset', 'I_e', 'P22', 'P21ex', 'P21in', 'P11ex', 'P11in', 'P20',
                  'surrogate_alpha'):
            setattr(d, f, float(k[f]))
        self.desc = d
        L.require_device()
        h = C.c_void_p()
        L.check(L.lib().bs_bptt_create(C.byref(d), C.byref(h)))
        self._h = h
        self._last = None

    def close(self):
        if getattr(self, '_h', None) and self._h.value:
            L.lib().bs_bptt_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ calls
    def _check(self, w_in, w_rec, in_counts, target, v0):
        import torch
        net = self.net
        want = {'w_in': (w_in, (net.n_in, net.n), torch.float32),
                'w_rec': (w_rec, (net.n, net.n), torch.float32),
                'in_counts': (in_counts, (self.n_steps, self.batch, net.n_in), torch.uint8),
                'target': (target, (net.n,), torch.float32)}
        if v0 is not None:
            want['v0'] = (v0, (self.batch, net.n), torch.float32)
        for k, (t, shape, dt) in want.items():
            if not isinstance(t, torch.Tensor) or t.device.type != 'cuda':
                raise TypeError(f'{k} must be a torch tensor on the HIP device')
            if tuple(t.shape) != shape or t.dtype != dt or not t.is_contiguous():
                raise ValueError(f'{k} must be a contiguous {dt} tensor of shape {shape}, '
                                 f'got {tuple(t.shape)} {t.dtype}')

    def forward(self, w_in, w_rec, in_counts, target, v0=None):
        """Loss of one episode (float). The tensors must stay alive and
        unchanged until :meth:`backward` has run."""
        import torch
        self._check(w_in, w_rec, in_counts, target, v0)
        torch.cuda.synchronize(w_rec.device)  # producers on torch's streams
        out = C.c_double()
        L.check(L.lib().bs_bptt_forward(self._h, w_in.data_ptr(), w_rec.data_ptr(),
                                        in_counts.data_ptr(), v0.data_ptr() if v0 is not None else None,
                                        target.data_ptr(), C.byref(out)))
        self._last = (w_in, w_rec, in_counts, target, v0)
        return out.value

    def backward(self):
        """Gradients (g_in, g_rec) of the last forward's loss, torch float32."""
        import torch
        if self._last is None:
            raise RuntimeError('backward() needs a forward() first')
        w_in, w_rec = self._last[0], self._last[1]
        g_in = torch.empty_like(w_in)
        g_rec = torch.empty_like(w_rec)
        L.check(L.lib().bs_bptt_backward(self._h, g_in.data_ptr(), g_rec.data_ptr()))
        return g_in, g_rec

    def loss_and_grad(self, w_in, w_rec, in_counts, target, v0=None):
        loss = self.forward(w_in, w_rec, in_counts, target, v0)
        return (loss,) + self.backward()

    def spikes(self):
        """(bitmaps uint64 [T, B, ceil(N/64)], counts float32 [B, N]) of the last forward."""
        nw = (self.net.n + 63) // 64
        bm = np.zeros((self.n_steps, self.batch, nw), np.uint64)
        cnt = np.zeros((self.batch, self.net.n), np.float32)
        L.check(L.lib().bs_bptt_spikes(self._h, bm.ctypes.data_as(C.POINTER(C.c_uint64)),
                                       cnt.ctypes.data_as(C.POINTER(C.c_float))))
        return bm, cnt


def _episode_fn():
    import torch

    class EpisodeLoss(torch.autograd.Function):
        """``loss = EpisodeLoss.apply(w_in, w_rec, bptt, in_counts, target)``;
        ``loss.backward()`` fills ``w_in.grad`` and ``w_rec.grad`` from the
        device adjoint."""

        @staticmethod
        def forward(ctx, w_in, w_rec, bptt, in_counts, target):
            ctx.bptt = bptt
            val = bptt.forward(w_in.detach(), w_rec.detach(), in_counts, target)
            return torch.tensor(val, dtype=torch.float64, device=w_rec.device)

        @staticmethod
        def backward(ctx, g):
            g_in, g_rec = ctx.bptt.backward()
            s = g.to(torch.float32)
            return g_in * s, g_rec * s, None, None, None

    return EpisodeLoss


class _Lazy:
    def __init__(self):
        self._cls = None

    def __getattr__(self, k):
        if self._cls is None:
            self._cls = _episode_fn()
        return getattr(self._cls, k)


EpisodeLoss = _Lazy()


# ------------------------------------------------------------------ problems
def poisson_counts(rates_hz, n_steps, batch, dt, seed):
    """Input events uint8 [n_steps, batch, n_in], Poisson(rate * dt) per step."""
    rng = np.random.default_rng(seed)
    lam = np.asarray(rates_hz, np.float64) * dt * 1e-3
    c = rng.poisson(lam[None, None, :], size=(n_steps, batch, len(lam)))
    return np.minimum(c, 255).astype(np.uint8)


def random_sparse_weights(n_pre, n_post, p, mean, std, seed, sign=1.0):
    """Dense (n_pre, n_post) float32 with Bernoulli(p) edges of |N(mean, std)| * sign."""
    rng = np.random.default_rng(seed)
    mask = rng.random((n_pre, n_post)) < p
    w = np.abs(rng.normal(mean, std, (n_pre, n_post))) * sign
    return (w * mask).astype(np.float32), mask


def bptt_problem(n_exc=3277, n_inh=819, n_in=200, batch=64, n_steps=1000, p=0.1, seed=0,
                 w_in=(260.0, 60.0), w_exc=(30.0, 8.0), w_inh=(120.0, 30.0)):
    """The differentiation workload: recurrent iaf_psc_exp, 10 % trainable
    recurrent and input weights, Poisson inputs at U[5, 50] Hz, rate targets
    U[2, 20] Hz, delay 1 ms, dt 0.1 ms. Returns numpy arrays."""
    net = RecurrentIafExp(n_in, n_exc, n_inh, dt=0.1, delay=1.0, delay_in=1.0)
    rng = np.random.default_rng(seed)
    rates = rng.uniform(5.0, 50.0, n_in)
    target = rng.uniform(2.0, 20.0, n_exc + n_inh).astype(np.float32)
    win, _ = random_sparse_weights(n_in, net.n, p, *w_in, seed=seed + 1)
    we, _ = random_sparse_weights(n_exc, net.n, p, *w_exc, seed=seed + 2)
    wi, _ = random_sparse_weights(n_inh, net.n, p, *w_inh, seed=seed + 3, sign=-1.0)
    wrec = np.concatenate([we, wi], 0)
    counts = poisson_counts(rates, n_steps, batch, net.dt, seed + 4)
    return net, win, wrec, counts, target
