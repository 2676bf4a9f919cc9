"""ctypes binding of ``libbrainspike_hip.so`` (C ABI in ``include/brainspike_hip.h``).

The library is built in-tree by ``__graft_entry__.build()`` (``hipcc
--offload-arch=gfx950``). There is no CPU fallback: every call that needs the
engine raises :class:`EngineUnavailable` when the library or a GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import os

__all__ = ['lib', 'EngineUnavailable', 'EngineError', 'check', 'BS_NPARAM', 'BS_NSTATE',
           'Population', 'Channel', 'Projection', 'Stdp', 'NetDesc', 'BpttDesc', 'LIB_PATH']

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libbrainspike_hip.so')

BS_ABI_VERSION = 9
BS_NPARAM = 40
BS_NSTATE = 9
BS_F64, BS_F32 = 0, 1
BS_IAF_PSC_DELTA, BS_IAF_PSC_EXP, BS_IAF_PSC_ALPHA, BS_AEIF_COND_ALPHA = 1, 2, 3, 4
BS_CH_COUNT, BS_CH_FIXED, BS_CH_POISSON, BS_CH_CURRENT, BS_CH_AC, BS_CH_SCHEDULE = 1, 2, 3, 4, 5, 6

_ERRORS = {-1: 'invalid argument', -2: 'out of memory', -3: 'HIP error',
           -4: 'capacity exceeded', -5: 'numerical instability', -6: 'no device'}


class EngineUnavailable(RuntimeError):
    """The HIP engine library is not built or no gfx950 device is visible."""


class EngineError(RuntimeError):
    """A C-ABI call returned an error code."""

    def __init__(self, code, msg):
        super().__init__(f'{_ERRORS.get(code, "error")} ({code}): {msg}')
        self.code = code


class Population(C.Structure):
    _fields_ = [('model', C.c_int32), ('n', C.c_int32), ('params', C.c_double * BS_NPARAM),
                ('state0', C.POINTER(C.c_double)), ('ch_begin', C.c_int32),
                ('ch_count', C.c_int32), ('n_vm', C.c_int32),
                ('vm_idx', C.POINTER(C.c_int32)), ('record_spikes', C.c_int32),
                ('remote', C.c_int32)]


class Channel(C.Structure):
    _fields_ = [('kind', C.c_int32), ('receptor', C.c_int32), ('weight', C.c_double),
                ('max_count', C.c_int32), ('fixed_shift', C.c_int32), ('lam', C.c_double),
                ('t_min_step', C.c_int64), ('t_max_step', C.c_int64), ('seed', C.c_uint32),
                ('delay_steps', C.c_int32), ('gen_index', C.POINTER(C.c_int32)),
                ('ac', C.c_double * 4), ('sched', C.POINTER(C.c_int64)), ('sched_val', C.POINTER(C.c_double)),
                ('sched_len', C.c_int32), ('_pad_sched', C.c_int32), ('delay_frac', C.c_double)]


class Stdp(C.Structure):
    _fields_ = [(k, C.c_double) for k in ('tau_plus', 'tau_minus', 'lambda_', 'alpha',
                                          'mu_plus', 'mu_minus', 'Wmax', 'decay_plus',
                                          'decay_minus')]


class Projection(C.Structure):
    _fields_ = [('pre_pop', C.c_int32), ('post_pop', C.c_int32), ('channel', C.c_int32),
                ('plastic', C.c_int32), ('n_edges', C.c_int64),
                ('pre', C.POINTER(C.c_int32)), ('post', C.POINTER(C.c_int32)),
                ('delay', C.POINTER(C.c_int32)), ('delay_steps', C.c_int32),
                ('on_device', C.c_int32), ('weight', C.POINTER(C.c_double)), ('stdp', Stdp)]


class NetDesc(C.Structure):
    _fields_ = [('abi_version', C.c_int32), ('dtype', C.c_int32), ('dt_ms', C.c_double),
                ('n_pops', C.c_int32), ('n_channels', C.c_int32), ('n_projs', C.c_int32),
                ('max_window', C.c_int32), ('pops', C.POINTER(Population)),
                ('channels', C.POINTER(Channel)), ('projs', C.POINTER(Projection)),
                ('tables', C.POINTER(C.c_double)), ('table_off', C.POINTER(C.c_int64)),
                ('n_tables', C.c_int64), ('tile_size', C.c_int32), ('plastic_records', C.c_int32)]


class BpttDesc(C.Structure):
    _fields_ = ([(k, C.c_int32) for k in ('abi_version', 'n_in', 'n_exc', 'n_inh', 'batch', 'n_steps',
                                          'delay_steps', 'delay_in_steps', 'checkpoint', 'loss',
                                          'ref_steps', 'gemm')] +
                [(k, C.c_double) for k in ('dt_ms', 'E_L', 'V_th', 'V_reset', 'I_e', 'P22', 'P21ex',
                                           'P21in', 'P11ex', 'P11in', 'P20', 'surrogate_alpha')] +
                [('replay', C.c_int32), ('_pad', C.c_int32)])


BS_LOSS_RATE_MSE, BS_LOSS_MEMBRANE_MSE = 0, 1

_P = C.c_void_p
_SIGS = {
    'bs_abi_version': (C.c_int, []),
    'bs_source_hash': (C.c_char_p, []),
    'bs_last_error': (C.c_char_p, []),
    'bs_device_count': (C.c_int, []),
    'bs_net_create': (C.c_int, [C.POINTER(NetDesc), C.POINTER(_P)]),
    'bs_net_destroy': (C.c_int, [_P]),
    'bs_net_reset': (C.c_int, [_P]),
    'bs_net_run': (C.c_int, [_P, C.c_int64, C.POINTER(C.c_double)]),
    'bs_net_step': (C.c_int, [_P, C.POINTER(C.c_int64)]),
    'bs_net_spike_count': (C.c_int, [_P, C.POINTER(C.c_int64)]),
    'bs_net_copy_spikes': (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int64]),
    'bs_net_get_state': (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    'bs_net_get_weights': (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double)]),
    'bs_net_set_state': (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    'bs_net_state_bytes': (C.c_int, [_P, C.POINTER(C.c_int64)]),
    'bs_net_save_state': (C.c_int, [_P, _P, C.c_int64]),
    'bs_net_load_state': (C.c_int, [_P, _P, C.c_int64]),
    'bs_net_run_device': (C.c_int, [_P, C.c_int64, _P, _P, _P, _P]),
    'bs_net_set_weight_taps': (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32)]),
    'bs_net_set_drive': (C.c_int, [_P, C.c_int32, C.c_int64, C.c_int64, C.c_int32, C.POINTER(C.c_int32)]),
    'bs_net_copy_weight_trace': (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double)]),
    'bs_net_set_timing': (C.c_int, [_P, C.c_int]),
    'bs_net_last_timing': (C.c_int, [_P] + [C.POINTER(C.c_double)] * 3 + [C.POINTER(C.c_int64)] * 2),
    'bs_net_last_bytes': (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_int64)]),
    'bs_bptt_create': (C.c_int, [C.POINTER(BpttDesc), C.POINTER(_P)]),
    'bs_bptt_destroy': (C.c_int, [_P]),
    'bs_bptt_forward': (C.c_int, [_P, _P, _P, _P, _P, _P, C.POINTER(C.c_double)]),
    'bs_bptt_backward': (C.c_int, [_P, _P, _P]),
    'bs_bptt_spikes': (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_float)]),
    'bs_net_window': (C.c_int, [_P, C.POINTER(C.c_int32)]),
    'bs_net_inject_spikes': (C.c_int, [_P, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    'bs_conn_fixed_indegree': (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                         _P, _P, _P, _P]),
    'bs_conn_bernoulli_count': (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_int32, _P,
                                          C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    'bs_conn_bernoulli_count_noalloc': (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_int32, _P,
                                                  C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    'bs_conn_bernoulli_fill': (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_int32, _P, _P, _P,
                                         _P, _P]),
    'bs_debug_math': (C.c_int, [C.c_int32, C.c_int64, C.POINTER(C.c_double),
                                C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    'bs_debug_threefry': (C.c_int, [C.c_int64] + [C.POINTER(C.c_uint32)] * 6),
    'bs_debug_poisson': (C.c_int, [C.c_int64] + [C.POINTER(C.c_uint32)] * 3 +
                         [C.c_double, C.POINTER(C.c_uint32)]),
}

_lib = None


def lib():
    """Load the engine (once). Raises EngineUnavailable if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EngineUnavailable(
            f'{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950)')
    # torch ships its own HIP runtime. When the engine's (system ROCm) runtime is
    # loaded and initialised first, torch's later device initialisation in the
    # same process finds no GPU; loading torch's libraries first avoids that.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    handle = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    if handle.bs_abi_version() != BS_ABI_VERSION:
        raise EngineUnavailable('libbrainspike_hip.so ABI version mismatch; rebuild it')
    from .build import source_hash
    built, here = handle.bs_source_hash().decode(), source_hash()
    if built != here:
        raise EngineUnavailable(f'libbrainspike_hip.so was built from other sources (hash {built[:12]}, '
                                f'sources here {here[:12]}); rebuild it with __graft_entry__.build()')
    _lib = handle
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().bs_last_error()
        raise EngineError(rc, msg.decode() if msg else '')
    return rc


def require_device():
    """Raise EngineUnavailable unless a HIP device is visible to the engine."""
    if lib().bs_device_count() < 1:
        raise EngineUnavailable('no HIP (gfx950) device is visible; the engine has no CPU path')
