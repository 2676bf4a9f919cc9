"""GPU parity of the engine against the CPU oracle (bit-exact raster, V_m 1e-12).

Every case builds one network with the Simulator, runs it through the engine
(``bs_net_run``) and through the oracle from the same descriptor, and compares
the spike lists ``(gid, step)`` exactly and the voltmeter traces to 1e-12
relative in FP64 (bit-exact in practice).
"""
import numpy as np
import pytest

import oracle
from brainpy.state_amd import units as u
from brainpy.state_amd.networks import brunel_delta

pytestmark = pytest.mark.gpu


def _both(sim, n_steps):
    desc = sim.build_desc()
    eng = sim._ensure_engine()
    eng.reset()
    vm_g, (g_g, s_g) = eng.run(n_steps)
    orc = oracle.OracleNet(desc)
    vm_o, g_o, s_o = orc.run(n_steps)
    return (vm_g, g_g, s_g), (vm_o, g_o, s_o)


def _assert_same(gpu, orc, rtol=1e-12):
    vm_g, g_g, s_g = gpu
    vm_o, g_o, s_o = orc
    assert len(g_g) == len(g_o), f'spike count differs: gpu {len(g_g)} oracle {len(g_o)}'
    assert np.array_equal(s_g, s_o) and np.array_equal(g_g, g_o), 'raster differs'
    np.testing.assert_allclose(vm_g, vm_o, rtol=rtol, atol=0)


def test_device_math_matches_oracle_bitwise():
    from brainpy.state_amd import _lib
    import ctypes as C
    lib = _lib.lib()
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-50, 50, 4000), rng.uniform(-745, 709, 2000),
                        [0.0, -0.0, 1e-300, -1e-300, 709.0, -744.0]])
    out = np.zeros_like(x)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    _lib.check(lib.bs_debug_math(0, len(x), P(x), None, P(out)))
    ref = np.array([oracle.exp(v) for v in x])
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    y = np.abs(x) + 1e-3
    _lib.check(lib.bs_debug_math(1, len(y), P(y), None, P(out)))
    assert np.array_equal(out.view(np.uint64), np.array([oracle.log(v) for v in y]).view(np.uint64))
    e = np.full_like(y, 0.2)
    _lib.check(lib.bs_debug_math(2, len(y), P(y), P(e), P(out)))
    assert np.array_equal(out.view(np.uint64), np.array([oracle.pow(v, 0.2) for v in y]).view(np.uint64))


def test_reciprocal_division_is_the_ieee_quotient():
    """The aeif update divides through a precomputed reciprocal plus one FMA
    correction; it must give the IEEE quotient's bits for every input,
    including zeros, infinities, NaN and subnormal quotients."""
    from brainpy.state_amd import _lib
    import ctypes as C
    lib = _lib.lib()
    rng = np.random.default_rng(7)
    n = 1 << 20
    a = rng.standard_normal(n) * np.exp2(rng.integers(-1000, 1000, n).astype(np.float64))
    b = np.abs(rng.standard_normal(n)) * np.exp2(rng.integers(-40, 40, n).astype(np.float64)) + 1e-300
    b[: n // 4] = rng.choice([281.0, 0.2, 2.0, 144.0, 0.5, 1e-3], n // 4)  # the model's divisors
    a[:8] = [0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, -3e-310, 1.7e308]
    a[8:16] = [1e-307, 2.2e-308, -1e-306, 3e-300, 1e300, -4e299, 1e-320, 7e-309]
    b[8:16] = [0.2, 281.0, 144.0, 1e7, 1e-8, 1e-9, 0.5, 3.0]
    out = np.zeros_like(a)
    P = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
    _lib.check(lib.bs_debug_math(3, n, P(a), P(b), P(out)))
    with np.errstate(all='ignore'):
        ref = a / b
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(out), nan)
    assert np.array_equal(out[~nan].view(np.uint64), ref[~nan].view(np.uint64))


def test_device_threefry_and_poisson_match_oracle():
    from brainpy.state_amd import _lib
    import ctypes as C
    lib = _lib.lib()
    rng = np.random.default_rng(1)
    n = 5000
    k0, k1, c0, c1 = (rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32) for _ in range(4))
    o0 = np.zeros(n, np.uint32)
    o1 = np.zeros(n, np.uint32)
    U = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
    _lib.check(lib.bs_debug_threefry(n, U(k0), U(k1), U(c0), U(c1), U(o0), U(o1)))
    for i in range(0, n, 97):
        assert (o0[i], o1[i]) == oracle.threefry2x32(int(k0[i]), int(k1[i]), int(c0[i]), int(c1[i]))
    for lam in (0.8, 2.0, 17.5, 40.0):
        out = np.zeros(n, np.uint32)
        _lib.check(lib.bs_debug_poisson(n, U(k0), U(c0), U(c1), lam, U(out)))
        ref = np.array([oracle.poisson(int(k0[i]), int(c0[i]), int(c1[i]), lam) for i in range(n)])
        assert np.array_equal(out, ref)


def test_brunel_1000_delta_fp64_10k_steps():
    """Brunel-1000 iaf_psc_delta, float64, 1000 ms: bit-exact raster."""
    sim, h = brunel_delta(order=200, n_vm=10, dtype='float64')
    gpu, orc = _both(sim, 10000)
    assert len(orc[1]) > 10000
    _assert_same(gpu, orc)


def test_brunel_1000_delta_fp32_bitexact_vs_fp32_oracle():
    sim, h = brunel_delta(order=200, n_vm=10, dtype='float32')
    gpu, orc = _both(sim, 3000)
    _assert_same(gpu, orc, rtol=0)


def test_pipelined_chunks_and_split_runs():
    """Runs of several chunks (both record buffers and both captured graphs in
    use, then a partial chunk), split across run calls at points inside a
    chunk: the raster and V_m rows equal the oracle's single run."""
    sim, h = brunel_delta(order=50, n_vm=6, dtype='float64')
    desc = sim.build_desc()
    eng = sim._ensure_engine()
    eng.reset()
    parts = [eng.run(n) for n in (30011, 9000, 1)]
    vm_g = np.concatenate([p[0] for p in parts])
    g_g = np.concatenate([p[1][0] for p in parts])
    s_g = np.concatenate([p[1][1] for p in parts])
    vm_o, g_o, s_o = oracle.OracleNet(desc).run(39012)
    assert len(g_o) > 5000
    _assert_same((vm_g, g_g, s_g), (vm_o, g_o, s_o))


def test_brunel_per_edge_delays_small():
    sim, h = brunel_delta(order=100, delay=('uniform', 1.0, 3.0), n_vm=5)
    gpu, orc = _both(sim, 3000)
    assert len(orc[1]) > 1000
    _assert_same(gpu, orc)


def test_brunel_11250_per_edge_delays_short():
    """Brunel-11250 with per-edge delays 1-3 ms (12.7M synapses), 300 ms."""
    sim, h = brunel_delta(order=2250, delay=('uniform', 1.0, 3.0), n_vm=10)
    gpu, orc = _both(sim, 3000)
    assert len(orc[1]) > 10000
    _assert_same(gpu, orc)


def test_simulate_api_raster_and_trace():
    sim, h = brunel_delta(order=50, n_vm=3)
    res = sim.simulate(200. * u.ms)
    spk = res.spikes(h['esr'])
    assert spk.shape == (2000, 200)
    tr = res.trace(h['vm'])
    assert tr.shape == (2000, 3)
    assert res.rate(h['esr']) > 0
    res2 = sim.cont(100. * u.ms)
    assert res2.spikes(h['esr']).shape == (1000, 200)
    assert float(res2.times.to_decimal(u.ms)[0]) == pytest.approx(200.0)


# ---------------------------------------------------------------- other models
def _mixed_net(model, dtype='float64', hetero=False, n=300, delay=1.0, gen_subset=False):
    from brainpy.state_amd import (Simulator, aeif_cond_alpha, dist, fixed_indegree,
                                   poisson_generator, spike_recorder, voltmeter)
    from brainpy.state_amd.init import Uniform
    sim = Simulator(dt=0.1 * u.ms, dtype=dtype)
    cond = model is aeif_cond_alpha
    kw = {} if cond else dict(V_initializer=Uniform(-70., -56., seed=5))
    a = sim.create(model, n, **kw)
    b = sim.create(model, n // 4, **kw)
    noise = sim.create(poisson_generator, rate=12000. * u.Hz)
    rt = dict(receptor_type=1) if cond else {}
    rti = dict(receptor_type=2) if cond else {}
    wE = 3.0 if cond else 60.0
    wI = 4.0 if cond else -240.0
    if model.__name__ == 'iaf_psc_delta':
        wE, wI = 0.5, -2.0
    if gen_subset:
        sim.connect(noise, a[:n // 2], weight=wE, delay=delay, **rt)
        sim.connect(noise, a[n // 2:], weight=wE, delay=2 * delay, **rt)
    else:
        sim.connect(noise, a, weight=wE, delay=delay, **rt)
    sim.connect(noise, b, weight=wE, delay=delay, **rt)
    w = dist.Normal(wE, 0.2 * wE) if hetero else wE
    sim.connect(a, a + b, weight=w, delay=delay, rule=fixed_indegree(30), seed=4, **rt)
    wi = dist.Uniform(0.5 * wI, 1.5 * wI) if hetero else wI
    sim.connect(b, a + b, weight=wi, delay=2 * delay, rule=fixed_indegree(10), seed=5, **rti)
    ra = sim.create(spike_recorder)
    sim.connect(a, ra)
    rb = sim.create(spike_recorder)
    sim.connect(b, rb)
    vm = sim.create(voltmeter)
    sim.connect(vm, a[:5])
    return sim


@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_iaf_psc_exp_network(dtype):
    from brainpy.state_amd import iaf_psc_exp
    sim = _mixed_net(iaf_psc_exp, dtype=dtype)
    gpu, orc = _both(sim, 3000)
    assert len(orc[1]) > 100
    _assert_same(gpu, orc, rtol=1e-12 if dtype == 'float64' else 0)


def test_iaf_psc_alpha_network_fp64():
    from brainpy.state_amd import iaf_psc_alpha
    sim = _mixed_net(iaf_psc_alpha)
    gpu, orc = _both(sim, 3000)
    assert len(orc[1]) > 100
    _assert_same(gpu, orc)


def test_aeif_cond_alpha_network_fp64():
    from brainpy.state_amd import aeif_cond_alpha
    sim = _mixed_net(aeif_cond_alpha, n=200)
    gpu, orc = _both(sim, 1500)
    assert len(orc[1]) > 50
    _assert_same(gpu, orc)


def test_heterogeneous_weights_fixed_point_sums():
    from brainpy.state_amd import iaf_psc_exp
    sim = _mixed_net(iaf_psc_exp, hetero=True)
    gpu, orc = _both(sim, 3000)
    assert len(orc[1]) > 100
    _assert_same(gpu, orc)


def test_generator_subsets_and_zero_delay_window():
    """Generators on sub-segments (gen_index) and delay 0 (window of one step)."""
    from brainpy.state_amd import iaf_psc_delta
    sim = _mixed_net(iaf_psc_delta, gen_subset=True, delay=0.0)
    gpu, orc = _both(sim, 1000)
    assert len(orc[1]) > 100
    _assert_same(gpu, orc)


def test_brunel_11250_fp32_statistics_within_distributional_band():
    """float32 against float64 over 4 s: mean rate and ISI CV within 5 %
    (the reference's distributional category); V_m agrees to float32 precision
    until the rasters first diverge."""
    from brainpy.state_amd import stats
    runs = {}
    for dt in ('float64', 'float32'):
        sim, h = brunel_delta(order=2250, delay=('uniform', 1.0, 3.0), n_vm=10, dtype=dt)
        eng = sim._ensure_engine()
        eng.reset()
        vm, (g, s) = eng.run(40000)
        runs[dt] = (g, s, vm)
    rep = stats.compare_runs(runs['float64'], runs['float32'], 11250, 40000, 0.1)
    assert rep['rate_rel_dev'] < 0.05 and rep['cv_isi_rel_dev'] < 0.05, rep
    assert rep['first_divergence_step'] is not None and rep['first_divergence_step'] > 10
    assert rep['vm_max_abs_err_mV_before_divergence'] < 1e-3, rep


def test_hub_population_with_five_outgoing_projections():
    """One population feeds five others (five gather sources reading the same
    spike buckets, different delays) and receives from all of them."""
    from brainpy.state_amd import (Simulator, fixed_indegree, iaf_psc_delta, poisson_generator,
                                   spike_recorder)
    from brainpy.state_amd.init import Uniform
    sim = Simulator(dt=0.1 * u.ms)
    kw = dict(E_L=0., V_th=20., V_reset=10., tau_m=20., t_ref=2., V_initializer=Uniform(0., 19., seed=3))
    hub = sim.create(iaf_psc_delta, 400, **kw)
    posts = [sim.create(iaf_psc_delta, 90 + 30 * k, **kw) for k in range(5)]
    noise = sim.create(poisson_generator, rate=20000. * u.Hz)
    sim.connect(noise, hub, weight=0.1, delay=1.5)
    for k, p in enumerate(posts):
        sim.connect(noise, p, weight=0.1, delay=1.5)
        sim.connect(hub, p, weight=0.1, delay=1.0 + 0.5 * k, rule=fixed_indegree(40), seed=10 + k)
        sim.connect(p, hub, weight=-0.4, delay=1.2, rule=fixed_indegree(5), seed=20 + k)
    for pop in [hub] + posts:
        sim.connect(pop, sim.create(spike_recorder))
    gpu, orc = _both(sim, 2000)
    assert len(orc[1]) > 500
    _assert_same(gpu, orc)
