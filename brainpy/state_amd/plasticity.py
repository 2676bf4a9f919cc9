"""``stdp_synapse``: pair-based STDP spec (``_nest_plasticity/stdp_synapse.py:34-179``).

Rule, per edge and step (reference ``update``): ``K+ = pre_trace - pre_spike``,
``K- = post_trace - post_spike``; facilitation where the post neuron fired,
``w/W <- w/W + lambda (1 - w/W)^mu_plus K+`` clipped at 1; then depression where
the (delayed) pre spike arrives, ``w/W <- w/W - alpha lambda (w/W)^mu_minus K-``
clipped at 0. Traces decay-then-add per neuron (``event_plastic.py:390-427``).
"""
from __future__ import annotations

from . import units as u

__all__ = ['stdp_synapse']


class stdp_synapse:
    """Pair-based STDP synapse spec (NEST ``stdp_synapse``)."""
    __module__ = 'brainpy.state_amd'
    is_homogeneous_weight = False
    stochastic = False

    def __init__(self, weight=1.0, delay=1.0, receptor_type=0, tau_plus=20.0, tau_minus=20.0,
                 lambda_=0.01, alpha=1.0, mu_plus=1.0, mu_minus=1.0, Wmax=100.0, Kplus=0.0):
        self.weight = weight
        d = u.scalar(delay, 'time')
        if not (d > 0.0) or d != d:
            raise ValueError('delay must be strictly positive.')
        self.delay = delay
        if int(receptor_type) < 0:
            raise ValueError('receptor_type must be >= 0.')
        self.receptor_type = int(receptor_type)
        self.tau_plus = u.scalar(tau_plus, 'time')
        self.tau_minus = u.scalar(tau_minus, 'time')
        self.lambda_ = float(lambda_)
        self.alpha = float(alpha)
        self.mu_plus = float(mu_plus)
        self.mu_minus = float(mu_minus)
        self.Wmax = float(Wmax)
        self.Kplus = float(Kplus)
        if self.tau_plus <= 0.0:
            raise ValueError("'tau_plus' must be > 0.")
        if self.tau_minus <= 0.0:
            raise ValueError("'tau_minus' must be > 0.")
        w0 = float(getattr(weight, 'value', weight)) if not hasattr(weight, 'sample') else 1.0
        # NEST's sign convention counts zero as positive: w and Wmax must lie
        # on the same side of the split [0, inf) | (-inf, 0)
        if (w0 < 0.0) != (self.Wmax < 0.0):
            raise ValueError(f'stdp_synapse: the weight ({w0}) and Wmax ({self.Wmax}) must have the same '
                             'sign (zero counts as positive)')
        if self.Kplus < 0.0:
            raise ValueError(f'stdp_synapse: Kplus must be >= 0, got {self.Kplus}')
        if self.Kplus != 0.0:
            raise NotImplementedError('Kplus != 0 (traces start at 0 in the engine)')

    # the rule itself, vectorised on NumPy arrays (host reference, used by tests)
    def facilitate(self, w, kplus):
        norm_w = w / self.Wmax + self.lambda_ * (1.0 - w / self.Wmax) ** self.mu_plus * kplus
        return norm_w * self.Wmax if norm_w < 1.0 else self.Wmax

    def depress(self, w, kminus):
        norm_w = w / self.Wmax - self.alpha * self.lambda_ * (w / self.Wmax) ** self.mu_minus * kminus
        return norm_w * self.Wmax if norm_w > 0.0 else 0.0
