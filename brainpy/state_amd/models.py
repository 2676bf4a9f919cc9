"""NEST neuron model classes of the network step, as parameter specs.

Each class keeps the reference's constructor signature (NEST parameter names,
defaults and validation) and turns itself into the engine's parameter block:
a ``double[40]`` whose layout is fixed by ``csrc/bs_models.h`` and whose
propagators are computed here on the host in FP64 with the reference's own
expressions. The step itself runs on the device (``k_update``).

References (paths relative to the brainspike tree):
  iaf_psc_delta   _nest_neuron/iaf_psc_delta.py:357-521
  iaf_psc_exp     _nest_neuron/iaf_psc_exp.py:385-682
  iaf_psc_alpha   _nest_neuron/iaf_psc_alpha.py:349-684
  aeif_cond_alpha _nest_neuron/aeif_cond_alpha.py:371-747
  propagators     _nest_base/utils.py:126-235
"""
from __future__ import annotations

import math

import numpy as np

from . import units as u
from ._lib import (BS_NPARAM, BS_NSTATE, BS_IAF_PSC_DELTA, BS_IAF_PSC_EXP, BS_IAF_PSC_ALPHA,
                   BS_AEIF_COND_ALPHA)
from .init import Constant, resolve_init

__all__ = ['iaf_psc_delta', 'iaf_psc_exp', 'iaf_psc_alpha', 'aeif_cond_alpha',
           'propagator_exp', 'alpha_propagator_p31_p32', 'Neuron']


# --------------------------------------------------------------- propagators
def propagator_exp(tau_syn, tau_m, c_m, h_ms):
    """P21 of the exp-PSC membrane system with the singular fallback.

    Restates ``_nest_base/utils.py:126-174``: the regular form
    ``gamma * exp(-h/tau_syn) * expm1(h * inv_beta)`` unless it is not a
    positive normal number, then ``h / C_m * exp(-h / tau_m)``.
    """
    tau_syn = np.asarray(tau_syn, dtype=np.float64)
    tau_m = np.asarray(tau_m, dtype=np.float64)
    c_m = np.asarray(c_m, dtype=np.float64)
    h = np.asarray(h_ms, dtype=np.float64)
    with np.errstate(divide='ignore', invalid='ignore', over='ignore', under='ignore'):
        beta = tau_syn * tau_m / (tau_m - tau_syn)
        gamma = beta / c_m
        inv_beta = (tau_m - tau_syn) / (tau_syn * tau_m)
        p32_raw = gamma * np.exp(-h / tau_syn) * np.expm1(h * inv_beta)
        regular = np.isfinite(p32_raw) & (np.abs(p32_raw) >= np.finfo(np.float64).tiny) & (p32_raw > 0.0)
        p32_singular = h / c_m * np.exp(-h / tau_m)
        return np.where(regular, p32_raw, p32_singular)


def alpha_propagator_p31_p32(tau_syn, tau_m, c_m, h_ms):
    """P31, P32 of the alpha-PSC system (NEST IAFPropagatorAlpha).

    Restates ``_nest_base/utils.py:177-235`` including both singular limits.
    """
    tau_syn = np.asarray(tau_syn, dtype=np.float64)
    tau_m = np.asarray(tau_m, dtype=np.float64)
    c_m = np.asarray(c_m, dtype=np.float64)
    h = np.asarray(h_ms, dtype=np.float64)
    with np.errstate(divide='ignore', invalid='ignore', over='ignore', under='ignore'):
        beta = tau_syn * tau_m / (tau_m - tau_syn)
        gamma = beta / c_m
        inv_beta = (tau_m - tau_syn) / (tau_syn * tau_m)
        exp_h_tau_syn = np.exp(-h / tau_syn)
        expm1_h_tau = np.expm1(h * inv_beta)
        p32_raw = gamma * exp_h_tau_syn * expm1_h_tau
        exp_h_tau_m = np.exp(-h / tau_m)
        p32_singular = h / c_m * exp_h_tau_m
        tiny = np.finfo(np.float64).tiny
        p32 = np.where(np.isfinite(p32_raw) & (np.abs(p32_raw) >= tiny) & (p32_raw > 0.0),
                       p32_raw, p32_singular)
        h_min_regular = 1e-7 * tau_m * tau_m / np.abs(tau_m - tau_syn)
        p31_regular = gamma * exp_h_tau_syn * (beta * expm1_h_tau - h)
        p31_singular = 0.5 * h * h / c_m * exp_h_tau_m
        p31 = np.where(np.isfinite(h_min_regular) & (h > h_min_regular), p31_regular, p31_singular)
    return p31, p32


def _scalar(x, dim, name):
    m = np.asarray(u.magnitude(x, dim), dtype=np.float64)
    if m.size != 1:
        raise NotImplementedError(
            f'{name}: per-neuron parameter arrays are not supported by the engine; '
            'create one population per parameter value')
    return float(m.reshape(()))


class Neuron:
    """Common population spec: size, parameters, initial state, recordables."""
    __module__ = 'brainpy.state_amd'
    model_id = 0
    #: NEST recordable -> engine state variable (simulator.py:151-216 aliases)
    recordables = {'V_m': 0}

    def __init__(self, in_size, name=None):
        if isinstance(in_size, (tuple, list)):
            n = 1
            for s in in_size:
                n *= int(s)
        else:
            n = int(in_size)
        if n < 0:
            raise ValueError('population size must be >= 0')
        self.in_size = n
        self.varshape = (n,)
        self.name = name

    @property
    def n(self):
        return self.in_size

    def ref_count(self, dt):
        return int(math.ceil(self.t_ref / dt))

    def param_block(self, dt: float) -> np.ndarray:  # pragma: no cover - abstract
        raise NotImplementedError

    def state0(self, dt: float) -> np.ndarray:  # pragma: no cover - abstract
        raise NotImplementedError

    def _v0(self):
        return resolve_init(self.V_initializer, self.n, 'voltage')


class iaf_psc_delta(Neuron):
    """Leaky integrate-and-fire with delta (voltage-jump) synapses (NEST ``iaf_psc_delta``).

    Step (``_nest_neuron/iaf_psc_delta.py:478-521``):
    ``v = E_L + (V - E_L) e^{-dt/tau_m} + (I/C_m) tau_m (1 - e^{-dt/tau_m}) + sum(delta)``,
    held while refractory, spike iff ``v >= V_th``, reset to ``V_reset`` for
    ``ceil(t_ref/dt)`` steps.
    """
    __module__ = 'brainpy.state_amd'
    model_id = BS_IAF_PSC_DELTA

    def __init__(self, in_size, E_L=-70., C_m=250., tau_m=10., t_ref=2., V_th=-55.,
                 V_reset=-70., I_e=0., V_min=None, V_initializer=Constant(-70.),
                 spk_fun=None, spk_reset='hard', refractory_input=False, ref_var=False,
                 name=None):
        super().__init__(in_size, name)
        self.E_L = _scalar(E_L, 'voltage', 'E_L')
        self.C_m = _scalar(C_m, 'capacitance', 'C_m')
        self.tau_m = _scalar(tau_m, 'time', 'tau_m')
        self.t_ref = _scalar(t_ref, 'time', 't_ref')
        self.V_th = _scalar(V_th, 'voltage', 'V_th')
        self.V_reset = _scalar(V_reset, 'voltage', 'V_reset')
        self.I_e = _scalar(I_e, 'current', 'I_e')
        self.V_min = None if V_min is None else _scalar(V_min, 'voltage', 'V_min')
        self.V_initializer = V_initializer
        self.refractory_input = bool(refractory_input)
        self.spk_reset = spk_reset
        if spk_reset != 'hard':
            raise NotImplementedError("only spk_reset='hard' is supported")

    def param_block(self, dt):
        P = np.zeros(BS_NPARAM)
        decay = float(np.exp(-dt / self.tau_m))
        i_total = self.I_e + 0.0
        P[0] = self.E_L
        P[1] = decay
        P[2] = (i_total / self.C_m) * self.tau_m * (1. - decay)
        P[3] = self.V_th
        P[4] = self.V_reset
        P[5] = -np.inf if self.V_min is None else self.V_min
        P[6] = self.ref_count(dt)
        P[7] = 1.0 if self.refractory_input else 0.0
        if self.refractory_input:
            if self.ref_count(dt) > 23:
                raise NotImplementedError('refractory_input with more than 23 refractory steps')
            for k in range(24):     # exp(-ref_steps * dt / tau_m), iaf_psc_delta.py:501
                P[8 + k] = float(np.exp(-k * dt / self.tau_m))
        P[32], P[33], P[34], P[35] = self.C_m, self.tau_m, self.I_e, self.t_ref
        return P

    def state0(self, dt):
        S = np.zeros((BS_NSTATE, self.n))
        S[0] = self._v0()
        return S


class iaf_psc_exp(Neuron):
    """LIF with exponential PSCs (NEST ``iaf_psc_exp``), deterministic threshold.

    Propagators as in ``iaf_psc_exp._precompute_propagators``
    (``_nest_neuron/iaf_psc_exp.py:492-515``); step ``:604-682``.
    """
    __module__ = 'brainpy.state_amd'
    model_id = BS_IAF_PSC_EXP
    recordables = {'V_m': 0, 'I_syn_ex': 1, 'I_syn_in': 2}

    def __init__(self, in_size, E_L=-70., C_m=250., tau_m=10., t_ref=2., V_th=-55.,
                 V_reset=-70., tau_syn_ex=2., tau_syn_in=2., I_e=0., rho=0.01 / u.second,
                 delta=0., V_initializer=Constant(-70.), spk_fun=None, spk_reset='hard',
                 ref_var=False, name=None):
        super().__init__(in_size, name)
        self.E_L = _scalar(E_L, 'voltage', 'E_L')
        self.C_m = _scalar(C_m, 'capacitance', 'C_m')
        self.tau_m = _scalar(tau_m, 'time', 'tau_m')
        self.t_ref = _scalar(t_ref, 'time', 't_ref')
        self.V_th = _scalar(V_th, 'voltage', 'V_th')
        self.V_reset = _scalar(V_reset, 'voltage', 'V_reset')
        self.tau_syn_ex = _scalar(tau_syn_ex, 'time', 'tau_syn_ex')
        self.tau_syn_in = _scalar(tau_syn_in, 'time', 'tau_syn_in')
        self.I_e = _scalar(I_e, 'current', 'I_e')
        self.rho = _scalar(rho, 'frequency', 'rho')
        self.delta = _scalar(delta, 'voltage', 'delta')
        self.V_initializer = V_initializer
        # iaf_psc_exp._validate_parameters (iaf_psc_exp.py:427-451)
        if self.V_reset >= self.V_th:
            raise ValueError('Reset potential must be smaller than threshold.')
        if self.C_m <= 0.0:
            raise ValueError('Capacitance must be strictly positive.')
        if self.tau_m <= 0.0:
            raise ValueError('Membrane time constant must be strictly positive.')
        if self.tau_syn_ex <= 0.0 or self.tau_syn_in <= 0.0:
            raise ValueError('Synaptic time constants must be strictly positive.')
        if self.t_ref < 0.0:
            raise ValueError('Refractory time must not be negative.')
        if self.rho < 0.0:
            raise ValueError('Stochastic firing intensity rho must not be negative.')
        if self.delta < 0.0:
            raise ValueError('Threshold width delta must not be negative.')
        if self.delta >= 1e-10:
            raise NotImplementedError('escape-noise threshold (delta > 0) is not supported')

    def propagators(self, dt):
        h = dt
        P11_ex = float(np.exp(-h / self.tau_syn_ex))
        P11_in = float(np.exp(-h / self.tau_syn_in))
        P22 = float(np.exp(-h / self.tau_m))
        P21_ex = float(propagator_exp(self.tau_syn_ex, self.tau_m, self.C_m, h))
        P21_in = float(propagator_exp(self.tau_syn_in, self.tau_m, self.C_m, h))
        P20 = float(self.tau_m / self.C_m * (1.0 - np.exp(-h / self.tau_m)))
        return dict(P11_ex=P11_ex, P11_in=P11_in, P22=P22, P21_ex=P21_ex, P21_in=P21_in, P20=P20)

    def param_block(self, dt):
        p = self.propagators(dt)
        P = np.zeros(BS_NPARAM)
        P[0] = self.E_L
        P[1] = self.V_th - self.E_L
        P[2] = self.V_reset - self.E_L
        P[3], P[4], P[5], P[6] = p['P22'], p['P21_ex'], p['P21_in'], p['P20']
        P[7], P[8] = p['P11_ex'], p['P11_in']
        P[9] = self.I_e
        P[10] = self.ref_count(dt)
        P[11] = 1.0 - p['P11_ex']
        return P

    def state0(self, dt):
        S = np.zeros((BS_NSTATE, self.n))
        S[0] = self._v0()
        return S


class iaf_psc_alpha(Neuron):
    """LIF with alpha-shaped PSCs (NEST ``iaf_psc_alpha``).

    Propagators as in ``iaf_psc_alpha._precompute_propagators``
    (``_nest_neuron/iaf_psc_alpha.py:451-479``); step ``:604-684``.
    """
    __module__ = 'brainpy.state_amd'
    model_id = BS_IAF_PSC_ALPHA
    recordables = {'V_m': 0, 'I_syn_ex': 1, 'I_syn_in': 2}

    def __init__(self, in_size, E_L=-70., C_m=250., tau_m=10., t_ref=2., V_th=-55.,
                 V_reset=-70., tau_syn_ex=2., tau_syn_in=2., I_e=0., V_min=None,
                 V_initializer=Constant(-70.), spk_fun=None, spk_reset='hard', ref_var=False,
                 name=None):
        super().__init__(in_size, name)
        self.E_L = _scalar(E_L, 'voltage', 'E_L')
        self.C_m = _scalar(C_m, 'capacitance', 'C_m')
        self.tau_m = _scalar(tau_m, 'time', 'tau_m')
        self.t_ref = _scalar(t_ref, 'time', 't_ref')
        self.V_th = _scalar(V_th, 'voltage', 'V_th')
        self.V_reset = _scalar(V_reset, 'voltage', 'V_reset')
        self.tau_syn_ex = _scalar(tau_syn_ex, 'time', 'tau_syn_ex')
        self.tau_syn_in = _scalar(tau_syn_in, 'time', 'tau_syn_in')
        self.I_e = _scalar(I_e, 'current', 'I_e')
        self.V_min = None if V_min is None else _scalar(V_min, 'voltage', 'V_min')
        self.V_initializer = V_initializer
        # iaf_psc_alpha._validate_parameters (iaf_psc_alpha.py:390-410)
        if self.C_m <= 0.0:
            raise ValueError('Capacitance must be > 0.')
        if self.tau_m <= 0.0:
            raise ValueError('Membrane time constant must be > 0.')
        if self.tau_syn_ex <= 0.0 or self.tau_syn_in <= 0.0:
            raise ValueError('All synaptic time constants must be > 0.')
        if self.t_ref < 0.0:
            raise ValueError("The refractory time t_ref can't be negative.")
        if self.V_reset >= self.V_th:
            raise ValueError('Reset potential must be smaller than threshold.')

    def propagators(self, dt):
        h = dt
        P11_ex = float(np.exp(-h / self.tau_syn_ex))
        P11_in = float(np.exp(-h / self.tau_syn_in))
        expm1_tau_m = float(np.expm1(-h / self.tau_m))
        P31_ex, P32_ex = alpha_propagator_p31_p32(self.tau_syn_ex, self.tau_m, self.C_m, h)
        P31_in, P32_in = alpha_propagator_p31_p32(self.tau_syn_in, self.tau_m, self.C_m, h)
        return dict(P11_ex=P11_ex, P22_ex=P11_ex, P21_ex=h * P11_ex,
                    P11_in=P11_in, P22_in=P11_in, P21_in=h * P11_in,
                    expm1_tau_m=expm1_tau_m, P30=-self.tau_m / self.C_m * expm1_tau_m,
                    P31_ex=float(P31_ex), P32_ex=float(P32_ex),
                    P31_in=float(P31_in), P32_in=float(P32_in),
                    epsc_init=np.e / self.tau_syn_ex, ipsc_init=np.e / self.tau_syn_in)

    def param_block(self, dt):
        p = self.propagators(dt)
        P = np.zeros(BS_NPARAM)
        P[0] = self.E_L
        P[1] = self.V_th - self.E_L
        P[2] = self.V_reset - self.E_L
        P[3], P[4], P[5] = p['P11_ex'], p['P21_ex'], p['P22_ex']
        P[6], P[7], P[8] = p['P11_in'], p['P21_in'], p['P22_in']
        P[9], P[10] = p['expm1_tau_m'], p['P30']
        P[11], P[12], P[13], P[14] = p['P31_ex'], p['P32_ex'], p['P31_in'], p['P32_in']
        P[15], P[16] = p['epsc_init'], p['ipsc_init']
        P[17] = self.I_e
        P[18] = self.ref_count(dt)
        P[19] = -np.inf if self.V_min is None else self.V_min - self.E_L
        return P

    def state0(self, dt):
        S = np.zeros((BS_NSTATE, self.n))
        S[0] = self._v0()
        return S


class aeif_cond_alpha(Neuron):
    """Adaptive exponential IF with alpha conductances (NEST ``aeif_cond_alpha``).

    Integrated per neuron with adaptive RKF45 (``_nest_base/utils.py:583-723``,
    atol = ``gsl_error_tol``, rtol = 0, min step 1e-8 ms); spike/reset/
    adaptation inside the substep loop (``aeif_cond_alpha.py:594-625``).
    Receptor 1 feeds ``g_ex``, receptor 2 ``g_in`` (nS weights).
    """
    __module__ = 'brainpy.state_amd'
    model_id = BS_AEIF_COND_ALPHA
    recordables = {'V_m': 0, 'g_ex': 2, 'g_in': 4, 'w': 5}
    n_receptors = 2
    _MIN_H = 1e-8
    _MAX_ITERS = 100000

    def __init__(self, in_size, V_peak=0.0, V_reset=-60.0, t_ref=0.0, g_L=30.0, C_m=281.0,
                 E_ex=0.0, E_in=-85.0, E_L=-70.6, Delta_T=2.0, tau_w=144.0, a=4.0, b=80.5,
                 V_th=-50.4, tau_syn_ex=0.2, tau_syn_in=2.0, I_e=0.0, gsl_error_tol=1e-6,
                 V_initializer=Constant(-70.6), g_ex_initializer=Constant(0.0),
                 g_in_initializer=Constant(0.0), w_initializer=Constant(0.0), spk_fun=None,
                 spk_reset='hard', ref_var=False, name=None):
        super().__init__(in_size, name)
        s = _scalar
        self.V_peak = s(V_peak, 'voltage', 'V_peak')
        self.V_reset = s(V_reset, 'voltage', 'V_reset')
        self.t_ref = s(t_ref, 'time', 't_ref')
        self.g_L = s(g_L, 'conductance', 'g_L')
        self.C_m = s(C_m, 'capacitance', 'C_m')
        self.E_ex = s(E_ex, 'voltage', 'E_ex')
        self.E_in = s(E_in, 'voltage', 'E_in')
        self.E_L = s(E_L, 'voltage', 'E_L')
        self.Delta_T = s(Delta_T, 'voltage', 'Delta_T')
        self.tau_w = s(tau_w, 'time', 'tau_w')
        self.a = s(a, 'conductance', 'a')
        self.b = s(b, 'current', 'b')
        self.V_th = s(V_th, 'voltage', 'V_th')
        self.tau_syn_ex = s(tau_syn_ex, 'time', 'tau_syn_ex')
        self.tau_syn_in = s(tau_syn_in, 'time', 'tau_syn_in')
        self.I_e = s(I_e, 'current', 'I_e')
        self.gsl_error_tol = float(gsl_error_tol)
        self.V_initializer = V_initializer
        self.g_ex_initializer = g_ex_initializer
        self.g_in_initializer = g_in_initializer
        self.w_initializer = w_initializer
        # aeif_cond_alpha._validate_parameters (aeif_cond_alpha.py:449-492)
        if self.V_reset >= self.V_peak:
            raise ValueError('Ensure that: V_reset < V_peak .')
        if self.Delta_T < 0.0:
            raise ValueError('Delta_T must be positive.')
        if self.V_peak < self.V_th:
            raise ValueError('V_peak >= V_th required.')
        if self.C_m <= 0.0:
            raise ValueError('Capacitance must be strictly positive.')
        if self.t_ref < 0.0:
            raise ValueError('Refractory time cannot be negative.')
        if self.tau_syn_ex <= 0.0 or self.tau_syn_in <= 0.0 or self.tau_w <= 0.0:
            raise ValueError('All time constants must be strictly positive.')
        if self.gsl_error_tol <= 0.0:
            raise ValueError('The gsl_error_tol must be strictly positive.')
        if self.Delta_T > 0.0 and (self.V_peak - self.V_th) / self.Delta_T >= math.log(
                np.finfo(np.float64).max / 1e20):
            raise ValueError('The current combination of V_peak, V_th and Delta_T '
                             'will lead to numerical overflow at spike time.')

    def param_block(self, dt):
        P = np.zeros(BS_NPARAM)
        vals = [self.V_peak, self.V_reset, self.g_L, self.C_m, self.E_ex, self.E_in, self.E_L,
                self.Delta_T, self.tau_w, self.a, self.b, self.V_th, self.tau_syn_ex,
                self.tau_syn_in, self.I_e, self.gsl_error_tol, float(self.ref_count(dt)),
                np.e / self.tau_syn_ex, np.e / self.tau_syn_in, dt, self._MIN_H,
                self.V_peak if self.Delta_T > 0.0 else self.V_th,
                1.0 if self.Delta_T == 0.0 else self.Delta_T, float(self._MAX_ITERS),
                self.g_L * self.Delta_T, -self.g_L]
        P[:len(vals)] = vals
        return P

    def state0(self, dt):
        S = np.zeros((BS_NSTATE, self.n))
        S[0] = self._v0()
        S[2] = resolve_init(self.g_ex_initializer, self.n, 'conductance')
        S[4] = resolve_init(self.g_in_initializer, self.n, 'conductance')
        S[5] = resolve_init(self.w_initializer, self.n, 'current')
        S[6] = dt   # integration_step initialised to dt (aeif_cond_alpha.py:524)
        return S
