"""CPU oracle of the network step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package; the product (``brainpy.state_amd``) never does.
It loads ``oracle/liboracle.so`` (``oracle.cpp``, built by ``oracle/build.py``
with g++ -ffp-contract=off) and runs the reference's algorithm on one CPU
thread from the same ``bs_net_desc`` the engine consumes.

Pinning of each restated function (tests/test_oracle_kat.py,
tests/test_stdp_oracle.py, tests/test_bptt_oracle.py, tests/test_dc_generator.py):

* threefry2x32 -- Random123 published known-answer vectors;
* exp / log / pow -- agreement with libm to about one ulp;
* Poisson sampler -- moments, Fano factor, pmf against scipy, window rule;
* iaf_psc_delta -- closed-form recurrence and refractory cases
  (iaf_psc_delta_test.py:88-224);
* iaf_psc_exp -- DC first spike at 4.8 ms (iaf_psc_exp_test.py:84-103) and the
  closed-form PSC kernel;
* iaf_psc_alpha -- the NEST V_m table (iaf_psc_alpha_test.py:200-245);
* aeif_cond_alpha -- an independent RKF45 restated from
  aeif_cond_alpha_test.py:40-200, and the one-step current known answer;
* delivery -- latency s + 1 + d, event conservation, count tables equal to
  sequential sums;
* current inputs (dc_generator) -- the reference's dc_generator tests: the
  same trace as I_e, the closed-form iaf_psc_delta recurrence inside
  start/stop/origin windows, steady state, additivity, and the one-step latch
  of iaf_psc_exp / iaf_psc_alpha (the I_e trace one step later);
* stdp_synapse -- stdp_synapse_rule_test.py:65-131 known answers and a
  closed-form two-neuron trajectory;
* bptt_ref -- DC / refractory / PSC closed forms for the forward, central
  finite differences and a closed-form surrogate gradient for autograd;
* connectivity samplers (tests/test_conn_oracle.py) -- Python big-integer
  restatements of the index map on tiny grids (on the KAT-pinned Threefry),
  exact in-degree, self-exclusion, row-major order, p = 0 / p = 1 grids, and
  chi-square / binomial bounds on the index and edge-count distributions.

Parity unpinned: the reference's JAX random streams (Poisson counts,
connectivity), XLA's exp, brainevent's float accumulation order for per-edge
weights, jax.grad of the kernel learner and braintools' surrogate definition;
none of them can be run or imported here.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, 'liboracle.so')
_lib = None


DEPS = [os.path.join(HERE, 'oracle.cpp'), os.path.join(HERE, 'oracle.h'),
        os.path.join(HERE, '..', 'include', 'brainspike_hip.h')]


def source_hash():
    """SHA-256 over the oracle's source files; compiled into liboracle.so so
    that a library built from older sources is rebuilt before use."""
    import hashlib
    h = hashlib.sha256()
    for f in DEPS:
        h.update(os.path.basename(f).encode() + b'\0')
        with open(f, 'rb') as fh:
            h.update(fh.read())
    return h.hexdigest()


def build():
    tmp = LIB + f'.tmp{os.getpid()}'
    cmd = ['g++', '-O2', '-std=c++17', '-fPIC', '-shared', '-ffp-contract=off',
           f'-DORC_SOURCE_HASH="{source_hash()}"', '-o', tmp, os.path.join(HERE, 'oracle.cpp')]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('oracle build failed:\n' + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        want = source_hash()
        if not os.path.exists(LIB) or _built_hash(LIB) != want:
            build()
        h = C.CDLL(LIB)
        P = C.c_void_p
        sig = {
            'orc_last_error': (C.c_char_p, []),
            'orc_create': (C.c_int, [C.c_void_p, C.POINTER(P)]),
            'orc_destroy': (C.c_int, [P]),
            'orc_run': (C.c_int, [P, C.c_int64, C.POINTER(C.c_double)]),
            'orc_spike_count': (C.c_int, [P, C.POINTER(C.c_int64)]),
            'orc_copy_spikes': (C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int64]),
            'orc_get_state': (C.c_int, [P, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
            'orc_get_weights': (C.c_int, [P, C.c_int32, C.POINTER(C.c_double)]),
            'orc_set_drive': (C.c_int, [P, C.c_int32, C.c_int64, C.c_int64, C.c_int32, C.POINTER(C.c_int32)]),
            'orc_threefry2x32': (None, [C.c_uint32] * 4 + [C.POINTER(C.c_uint32)] * 2),
            'orc_neuron_step': (C.c_int, [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_int), C.c_double, C.c_double, C.c_int]),
            'orc_poisson': (C.c_uint32, [C.c_uint32] * 3 + [C.c_double]),
            'orc_exp': (C.c_double, [C.c_double]),
            'orc_log': (C.c_double, [C.c_double]),
            'orc_sin': (C.c_double, [C.c_double]),
            'orc_pow': (C.c_double, [C.c_double, C.c_double]),
            'orc_stdp_edge': (C.c_double, [C.c_double] * 5 + [C.c_void_p]),
            'orc_conn_fixed_indegree': (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                                  C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
            'orc_conn_bernoulli': (C.c_int64, [C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_int32,
                                               C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int64]),
        }
        for k, (res, args) in sig.items():
            f = getattr(h, k)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def _built_hash(path):
    """The source hash compiled into a built liboracle.so, read in a child
    process so that a stale library is never loaded into this one."""
    import sys
    code = ('import ctypes, sys; f = ctypes.CDLL(sys.argv[1]).orc_source_hash; '
            'f.restype = ctypes.c_char_p; print(f().decode())')
    res = subprocess.run([sys.executable, '-c', code, path], capture_output=True, text=True)
    return res.stdout.strip() if res.returncode == 0 else None


def _check(rc):
    if rc != 0:
        raise RuntimeError(f'oracle error {rc}: {lib().orc_last_error().decode()}')


def threefry2x32(k0, k1, c0, c1):
    a, b = C.c_uint32(), C.c_uint32()
    lib().orc_threefry2x32(k0, k1, c0, c1, C.byref(a), C.byref(b))
    return a.value, b.value


def poisson(seed, step, elem, lam):
    return int(lib().orc_poisson(seed, step, elem, lam))


def exp(x):
    return lib().orc_exp(float(x))


def log(x):
    return lib().orc_log(float(x))


def sin(x):
    return lib().orc_sin(float(x))


def pow(x, y):  # noqa: A001
    return lib().orc_pow(float(x), float(y))


def conn_fixed_indegree(seed, n_pre, n_post, K, exclude_self=False):
    """The device fixed in-degree sampler's edges ``(pre, post)`` (int32),
    post-major."""
    E = int(n_post) * int(K)
    pre = np.zeros(max(1, E), np.int32)
    post = np.zeros(max(1, E), np.int32)
    P = C.POINTER(C.c_int32)
    _check(lib().orc_conn_fixed_indegree(int(seed), int(n_pre), int(n_post), int(K), int(bool(exclude_self)),
                                         pre.ctypes.data_as(P), post.ctypes.data_as(P)))
    return pre[:E], post[:E]


def conn_bernoulli(seed, n_pre, n_post, p, exclude_self=False):
    """The device pairwise-Bernoulli sampler's edges ``(pre, post)`` (int32),
    row-major."""
    P = C.POINTER(C.c_int32)
    dummy = np.zeros(1, np.int32)
    n = lib().orc_conn_bernoulli(int(seed), int(n_pre), int(n_post), float(p), int(bool(exclude_self)),
                                 dummy.ctypes.data_as(P), dummy.ctypes.data_as(P), 0)
    if n < 0:
        _check(n)
    pre = np.zeros(max(1, n), np.int32)
    post = np.zeros(max(1, n), np.int32)
    lib().orc_conn_bernoulli(int(seed), int(n_pre), int(n_post), float(p), int(bool(exclude_self)),
                             pre.ctypes.data_as(P), post.ctypes.data_as(P), n)
    return pre[:n], post[:n]


def stdp_edge(w, pre_spike, post_spike, pre_trace, post_trace, Wmax=100., lambda_=0.01,
              alpha=1., mu_plus=1., mu_minus=1.):
    """One edge's stdp_synapse update with the given spikes and traces."""
    from brainpy.state_amd import _lib as L
    S = L.Stdp()
    S.Wmax, S.lambda_, S.alpha, S.mu_plus, S.mu_minus = Wmax, lambda_, alpha, mu_plus, mu_minus
    return lib().orc_stdp_edge(float(w), float(pre_spike), float(post_spike), float(pre_trace),
                               float(post_trace), C.addressof(S))


class Neuron:
    """One neuron of a model spec driven step by step through the oracle's update."""

    def __init__(self, model, dt, dtype='float64', index=0):
        self.model = model
        self.P = np.ascontiguousarray(model.param_block(dt), dtype=np.float64)
        self.s = np.ascontiguousarray(model.state0(dt)[:, index], dtype=np.float64)
        self.r = C.c_int(0)
        self.dtype = 0 if dtype == 'float64' else 1

    def step(self, in0=0.0, in1=0.0):
        rc = lib().orc_neuron_step(self.model.model_id, self.P.ctypes.data_as(C.POINTER(C.c_double)),
                                   self.s.ctypes.data_as(C.POINTER(C.c_double)), C.byref(self.r),
                                   float(in0), float(in1), self.dtype)
        if rc < 0:
            raise RuntimeError(f'oracle neuron step failed ({rc})')
        return bool(rc)

    @property
    def V(self):
        return float(self.s[0])


class OracleNet:
    """The oracle instance of a network descriptor (``Simulator.build_desc()``)."""

    def __init__(self, desc):
        self.desc = desc
        self.h = C.c_void_p()
        _check(lib().orc_create(C.cast(C.pointer(desc.c), C.c_void_p), C.byref(self.h)))

    def run(self, n_steps):
        nt = self.desc.n_taps
        vm = np.zeros((n_steps, nt))
        ptr = vm.ctypes.data_as(C.POINTER(C.c_double)) if nt else None
        _check(lib().orc_run(self.h, int(n_steps), ptr))
        n = C.c_int64()
        _check(lib().orc_spike_count(self.h, C.byref(n)))
        g = np.zeros(max(1, n.value), np.int32)
        s = np.zeros(max(1, n.value), np.int32)
        _check(lib().orc_copy_spikes(self.h, g.ctypes.data_as(C.POINTER(C.c_int32)),
                                     s.ctypes.data_as(C.POINTER(C.c_int32)), len(g)))
        return vm, g[:n.value], s[:n.value]

    def state(self, pop, var, n):
        out = np.zeros(max(1, n))
        _check(lib().orc_get_state(self.h, pop, var, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out[:n]

    def set_drive(self, channel, step0, counts):
        """Inject generator counts [T][n_gen] for channel ``channel`` from
        generator step ``step0`` on (see orc_set_drive)."""
        c = np.ascontiguousarray(counts, dtype=np.int32)
        if c.ndim != 2:
            raise ValueError('counts must be [n_steps][n_gen]')
        _check(lib().orc_set_drive(self.h, int(channel), int(step0), c.shape[0], c.shape[1],
                                   c.ctypes.data_as(C.POINTER(C.c_int32))))

    def weights(self, proj, n):
        out = np.zeros(max(1, n))
        _check(lib().orc_get_weights(self.h, proj, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out[:n]

    def close(self):
        if self.h.value:
            lib().orc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
