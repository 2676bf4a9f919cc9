"""CPU tests: the C-ABI library loads and exports its header, and the host-side
Simulator logic (rules, seeds, descriptor layout, validation) is right."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from brainpy.state_amd import (Simulator, all_to_all, explicit_edges, fixed_indegree,
                               fixed_total_number, iaf_psc_delta, iaf_psc_exp, one_to_one,
                               pairwise_bernoulli, poisson_generator, spike_recorder, units as u,
                               voltmeter, aeif_cond_alpha)
from brainpy.state_amd import _lib as L
from brainpy.state_amd.simulator import input_delay_steps, nest_delay_steps, seq_sum_table

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, 'include', 'brainspike_hip.h')).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char\*)\s+(bs_\w+)\(', text, re.M)))


def test_library_exports_every_header_function():
    if not os.path.exists(L.LIB_PATH):
        pytest.skip('engine not built')
    lib = C.CDLL(L.LIB_PATH)
    names = _header_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n


def test_abi_version_and_error_path_without_device():
    lib = L.lib()
    assert lib.bs_abi_version() == L.BS_ABI_VERSION
    assert lib.bs_device_count() >= 0
    h = C.c_void_p()
    rc = lib.bs_net_create(None, C.byref(h))
    assert rc == -1
    assert b'null' in lib.bs_last_error()


def test_descriptor_abi_rejected_on_version_mismatch():
    lib = L.lib()
    d = L.NetDesc()
    d.abi_version = 999
    h = C.c_void_p()
    assert lib.bs_net_create(C.byref(d), C.byref(h)) == -1
    assert b'ABI' in lib.bs_last_error()


def test_derive_seed_matches_reference_formula():
    # simulator.py:926-935
    assert Simulator._derive_seed(None, 0) == 1
    assert Simulator._derive_seed(1, 2) == (1_000_003 + 3) & 0x7FFFFFFF
    assert Simulator._derive_seed(2**40, 5) == (2**40 * 1_000_003 + 6) & 0x7FFFFFFF


def test_fixed_indegree_properties():
    spec = fixed_indegree(7).sample(50, 40, seed=3, pre_is_post=True, allow_autapses=False,
                                    allow_multapses=True)
    assert spec.n_edges == 280
    assert np.all(np.bincount(spec.post_idx, minlength=40) == 7)
    assert not np.any(spec.pre_idx == spec.post_idx)
    assert spec.pre_idx.min() >= 0 and spec.pre_idx.max() < 50
    spec2 = fixed_indegree(7).sample(50, 40, seed=3, pre_is_post=False, allow_autapses=True,
                                     allow_multapses=False)
    for j in range(40):
        pre = spec2.pre_idx[spec2.post_idx == j]
        assert len(set(pre.tolist())) == 7


def test_pairwise_bernoulli_density_and_order():
    spec = pairwise_bernoulli(0.1).sample(300, 200, seed=1, pre_is_post=False,
                                          allow_autapses=True, allow_multapses=True)
    assert abs(spec.n_edges / 60000 - 0.1) < 0.01
    flat = spec.pre_idx.astype(np.int64) * 200 + spec.post_idx
    assert np.all(np.diff(flat) > 0)      # row-major order, no duplicates


def test_rules_misc():
    a = all_to_all.sample(3, 3, seed=0, pre_is_post=True, allow_autapses=False, allow_multapses=True)
    assert a.n_edges == 6
    o = one_to_one.sample(4, 4, seed=0, pre_is_post=False, allow_autapses=True, allow_multapses=True)
    assert np.array_equal(o.pre_idx, o.post_idx)
    f = fixed_total_number(25).sample(10, 10, seed=0, pre_is_post=False, allow_autapses=True,
                                      allow_multapses=True)
    assert f.n_edges == 25
    e = explicit_edges(np.array([0, 0, 2]), np.array([1, 4, 3]))
    assert e.sample(5, 5, seed=0, pre_is_post=True, allow_autapses=True, allow_multapses=True).n_edges == 3
    with pytest.raises(ValueError):
        explicit_edges(np.array([0, 1]), np.array([1]))
    with pytest.raises(ValueError):
        pairwise_bernoulli(1.5)
    with pytest.raises(ValueError):
        fixed_indegree(-1)


def test_delay_conversion():
    """Homogeneous / axonal delays are read as InputDelay reads them
    (delay.py:114-128): frames floor(d/dt) and the next one, with the fraction
    computed in the network dtype; per-edge delays take NEST's rounding
    (_kernel/delays.py:50-108, whose docstring example is restated here)."""
    assert input_delay_steps(1.5 * u.ms, 0.1) == (15, 0.0)
    assert input_delay_steps(None, 0.1) == (0, 0.0)
    lo, fr = input_delay_steps(np.array([1.0, 2.0]) * u.ms, 0.1)
    assert np.array_equal(lo, [10, 20]) and np.all(fr == 0)
    assert input_delay_steps(1.5 * u.ms, 1.0) == (1, 0.5)          # the reference's 1.5-step KAT
    lo, fr = input_delay_steps(1.2 * u.ms, 0.1)                      # 11.999999999999998 in float64
    assert lo == 11 and fr == 1.2 / 0.1 - 11.0 and 0.99 < fr < 1.0
    lo32, fr32 = input_delay_steps(1.2 * u.ms, 0.1, np.float32)      # whole in float32
    assert (lo32, fr32) == (12, 0.0) and isinstance(fr32, np.float32)
    lo32, fr32 = input_delay_steps(1.3 * u.ms, 0.1, np.float32)      # 12.999999 in float32
    assert lo32 == 12 and fr32 == np.float32(1.3) / np.float32(0.1) - np.float32(12) and fr32 > 0.99
    with pytest.raises(ValueError):
        input_delay_steps(-0.1 * u.ms, 0.1)
    assert nest_delay_steps(np.array([0.0, 0.04, 0.05, 0.15, 1.5]) * u.ms, 0.1).tolist() == [0, 0, 1, 2, 15]
    assert nest_delay_steps(np.array([1.2, 0.7, 2.9]), 0.1).tolist() == [12, 7, 29]


def test_seq_sum_table_is_sequential_accumulation():
    t = seq_sum_table(0.1, 30, np.float64)
    acc = 0.0
    for k in range(31):
        assert t[k] == acc
        acc = acc + 0.1
    t32 = seq_sum_table(0.1, 30, np.float32)
    acc = np.float32(0.0)
    for k in range(31):
        assert t32[k] == float(acc)
        acc = np.float32(acc + np.float32(0.1))


def test_descriptor_channel_order_and_generator_fanout():
    sim = Simulator(dt=0.1 * u.ms)
    ne = sim.create(iaf_psc_delta, 20)
    ni = sim.create(iaf_psc_delta, 5)
    noise = sim.create(poisson_generator, rate=8000. * u.Hz)
    sim.connect(noise, ne + ni, weight=0.1 * u.mV, delay=1.5 * u.ms)     # ordinals 0, 1
    sim.connect(ne, ne + ni, weight=0.1, delay=1.5, rule=fixed_indegree(4), seed=1)  # 2, 3
    sim.connect(ni, ne + ni, weight=-0.5, delay=1.5, rule=fixed_indegree(1), seed=2)  # 4, 5
    d = sim.build_desc()
    c = d.c
    assert c.n_pops == 2 and c.n_channels == 6 and c.n_projs == 4
    kinds = [c.channels[k].kind for k in range(6)]
    assert kinds == [L.BS_CH_POISSON, L.BS_CH_COUNT, L.BS_CH_COUNT] * 2
    # generators realised per target segment with derived seeds (simulator.py:1437-1463)
    assert c.channels[0].seed == Simulator._derive_seed(0, 0)
    assert c.channels[3].seed == Simulator._derive_seed(0, 1)
    assert c.channels[0].lam == pytest.approx(0.8)
    assert c.channels[0].t_min_step == 0
    assert c.channels[1].weight == 0.1 and c.channels[5].weight == -0.5
    assert c.projs[0].delay_steps == 15


def test_validation_errors():
    sim = Simulator(dt=0.1)
    a = sim.create(aeif_cond_alpha, 3)
    with pytest.raises(ValueError):
        sim.connect(a, a, weight=1.0, delay=1.0)        # aeif needs a receptor port
    with pytest.raises(ValueError):
        iaf_psc_exp(1, V_reset=-55., V_th=-55.)
    with pytest.raises(ValueError):
        Simulator(dt=0.1, dtype='float16')
    with pytest.raises(NotImplementedError):
        iaf_psc_exp(1, delta=1.0)


def test_simulate_without_device_fails_loudly():
    if L.lib().bs_device_count() > 0:
        pytest.skip('a device is present')
    sim = Simulator(dt=0.1)
    p = sim.create(iaf_psc_delta, 3)
    with pytest.raises(L.EngineUnavailable):
        sim.simulate(1.0)


def test_recorder_taps_registered():
    sim = Simulator(dt=0.1)
    p = sim.create(iaf_psc_delta, 10)
    sr = sim.create(spike_recorder)
    vm = sim.create(voltmeter)
    sim.connect(p[:4], sr)
    sim.connect(vm, p[2:5])
    d = sim.build_desc()
    assert d.c.pops[0].record_spikes == 1
    assert d.c.pops[0].n_vm == 3 and d.n_taps == 3


def test_stats_rate_cv_and_divergence():
    from brainpy.state_amd import stats
    # two neurons: regular (CV 0) and alternating ISIs 2, 4 (CV 1/3)
    g = np.array([0, 0, 0, 0, 1, 1, 1, 1])
    s = np.array([0, 5, 10, 15, 0, 2, 6, 8])
    assert stats.rate(s, 2, 100, 0.1) == pytest.approx(8 / 2 / 0.01)
    assert stats.cv_isi(g, s, 2) == pytest.approx((0.0 + np.std([2, 4, 2]) / np.mean([2, 4, 2])) / 2)
    o = np.lexsort((g, s))
    a_g, a_s = g[o], s[o]
    assert stats.first_divergence(a_g, a_s, a_g, a_s) is None
    b_s = a_s.copy()
    b_s[-1] += 1
    assert stats.first_divergence(a_g, a_s, a_g, b_s) == 15


def test_nodeview_algebra_and_indexing():
    """NodeView: concatenation, sizes, and indexing over the concatenated
    positions (single and multi-population views, slices, masks, arrays)."""
    from brainpy.state_amd import NodeView
    from brainpy.state_amd.models import iaf_psc_delta
    a, b = iaf_psc_delta(5), iaf_psc_delta(3)
    va, vb = NodeView.of(a), NodeView.of(b)
    v = va + vb
    assert len(v) == v.size == 8 and len(v.segments) == 2
    assert [s.population for s in v.segments] == [a, b]
    # single population: the reference's slicing
    s = va[1:4]
    assert len(s.segments) == 1 and s.segments[0].population is a
    np.testing.assert_array_equal(s.segments[0].indices, [1, 2, 3])
    np.testing.assert_array_equal(va[-1].segments[0].indices, [4])
    # across populations: positions 3..6 = a[3], a[4], b[0], b[1]
    x = v[3:7]
    assert [sg.population for sg in x.segments] == [a, b]
    np.testing.assert_array_equal(x.segments[0].indices, [3, 4])
    np.testing.assert_array_equal(x.segments[1].indices, [0, 1])
    # a mask and an index array with an order change split into runs
    m = np.zeros(8, bool)
    m[[0, 6, 7]] = True
    y = v[m]
    np.testing.assert_array_equal(y.segments[0].indices, [0])
    np.testing.assert_array_equal(y.segments[1].indices, [1, 2])
    z = v[np.array([6, 0, 1])]
    assert [sg.population for sg in z.segments] == [b, a]
    np.testing.assert_array_equal(z.segments[1].indices, [0, 1])
    assert len(v[8:]) == 0
    with pytest.raises(IndexError):
        v[8]
    with pytest.raises(TypeError):
        va + 1


def test_recorders_created_inline_keep_distinct_taps():
    """Recorders created and dropped inside a loop (the usual
    ``sim.connect(pop, sim.create(spike_recorder))``) each keep their tap:
    the Simulator holds the devices, so a later device never reuses an
    earlier one's id (which keys the tap)."""
    import gc
    sim = Simulator(dt=0.1 * u.ms)
    pops = [sim.create(iaf_psc_delta, 3 + k) for k in range(6)]
    for p in pops:
        sim.connect(p, sim.create(spike_recorder))
        gc.collect()
    assert len(sim._taps) == 6
    vm = [sim.create(voltmeter) for _ in range(2)]
    sim.connect(vm[0], pops[0])
    sim.connect(sim.create(voltmeter), pops[1])
    gc.collect()
    sim.connect(sim.create(voltmeter), pops[2])
    assert len(sim._analog_taps) == 3
    assert sim.build_desc().n_taps == 3 + 4 + 5
