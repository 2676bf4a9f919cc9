"""GPU parity at the benchmark configurations' full sizes.

* Brunel-11250 iaf_psc_delta with per-edge delays 1-3 ms, float64, the full
  10 s (100,000 steps): raster bit-exact and V_m within 1e-12 of the oracle.
* aeif_cond_alpha 40,000 E + 10,000 I, p = 0.01 (~25M synapses), stdp_synapse
  on E->E (~16M plastic edges), delays 1-5 ms, Poisson 8 kHz: at the specified
  1 nS drive (subthreshold from rest: V_m of every tap) and at 3 nS (the first
  synchronous wave of spikes, then plasticity): raster, V_m and the final
  E->E weights against the oracle's literal per-edge restatement.
* The 256-network iaf_psc_exp ensemble (N log-uniform 500-20,000, float32):
  every member's raster inside the merged engine equals its own plain run over
  5000 steps; six members (smallest, median, largest and three more) equal
  the oracle bit for bit over their first 1000 steps.
* BPTT 4096 neurons x batch 64 x 1000 steps: forward spikes of all 64 trials
  bit-exact against the float32 restatement; the gradient of a two-trial
  episode within 1e-4 (relative, L2) of torch float64 autograd, and the
  gradient of all 64 trials within 1e-4 of torch float32 autograd of the same
  episode on the GPU.
"""
import numpy as np
import pytest

import oracle
from brainpy.state_amd.networks import aeif_stdp, brunel_delta, brunel_exp_ensemble

pytestmark = pytest.mark.gpu


def _run_both(sim, n_steps):
    desc = sim.build_desc()
    eng = sim._ensure_engine()
    eng.reset()
    vm_g, (g_g, s_g) = eng.run(n_steps)
    orc = oracle.OracleNet(desc)
    vm_o, g_o, s_o = orc.run(n_steps)
    return desc, orc, (vm_g, g_g, s_g), (vm_o, g_o, s_o)


def _same_raster(gpu, orc):
    vm_g, g_g, s_g = gpu
    vm_o, g_o, s_o = orc
    assert len(g_g) == len(g_o), f'spike count differs: gpu {len(g_g)} oracle {len(g_o)}'
    assert np.array_equal(s_g, s_o) and np.array_equal(g_g, g_o), 'raster differs'


def test_brunel_11250_per_edge_delays_full_10s_fp64():
    sim, h = brunel_delta(order=2250, delay=('uniform', 1.0, 3.0), n_vm=10, dtype='float64')
    print('brunel-11250 f64, 100000 steps', flush=True)
    _, _, gpu, orc = _run_both(sim, 100000)
    assert len(orc[1]) > 3_000_000  # ~40 Hz over 10 s
    _same_raster(gpu, orc)
    np.testing.assert_allclose(gpu[0], orc[0], rtol=1e-12, atol=0)
    sim._engine.close()
    sim._engine = None


@pytest.mark.parametrize('w_noise,n_steps', [(1.0, 200), (3.0, 300)])
def test_aeif_50k_stdp_fp64(w_noise, n_steps):
    sim, h = aeif_stdp(n=50000, w_noise=w_noise, n_vm=10)
    print(f'aeif-50k + stdp, drive {w_noise} nS, {n_steps} steps', flush=True)
    desc, orc, gpu, ref = _run_both(sim, n_steps)
    _same_raster(gpu, ref)
    np.testing.assert_allclose(gpu[0], ref[0], rtol=1e-12, atol=0)
    pee = h['pee']
    w_g = sim.get_weights(pee)
    w_o = orc.weights(desc.proj_index[id(pee)], pee.n_edges)
    np.testing.assert_allclose(w_g, w_o, rtol=1e-12, atol=0)
    if w_noise == 1.0:
        assert len(ref[1]) == 0  # the specified drive stays subthreshold from rest
        assert np.ptp(ref[0]) > 1.0  # but V_m moves
    else:
        assert len(ref[1]) > 10000, 'the 3 nS drive must make the first wave of spikes'
        assert np.any(w_o != np.asarray(pee.weight, np.float64)), 'no plasticity happened'
    sim._engine.close()
    sim._engine = None


def test_ensemble_256_networks_equal_plain_runs_and_oracle():
    from brainpy.state_amd import Ensemble
    sims, prm = brunel_exp_ensemble(256, seed=0)
    print(f'ensemble: 256 networks, {int(prm["N"].sum())} neurons', flush=True)
    ens = Ensemble(sims)
    parts = ens.run(5000)
    ens.close()
    n_spk = 0
    for k, sim in enumerate(sims):
        eng = sim._ensure_engine()
        eng.reset()
        _, (g, s) = eng.run(5000)
        eng.close()
        sim._engine = None
        assert np.array_equal(parts[k][1], g) and np.array_equal(parts[k][2], s), f'network {k}'
        n_spk += len(g)
    assert n_spk > 1_000_000
    order = np.argsort(prm['N'])
    rng = np.random.default_rng(3)
    pick = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    pick += [int(k) for k in rng.choice(np.setdiff1d(np.arange(256), pick), 3, replace=False)]
    print(f'oracle members {pick}', flush=True)
    for k in pick:
        _, g_o, s_o = oracle.OracleNet(sims[k].build_desc()).run(1000)
        first = parts[k][2] < 1000
        assert len(g_o) > 0, f'network {k} silent in the first 100 ms'
        assert np.array_equal(parts[k][1][first], g_o) and np.array_equal(parts[k][2][first], s_o), k


def test_bptt_4096_batch64_t1000_forward_and_gradient():
    torch = pytest.importorskip('torch')
    from oracle import bptt_ref as R
    from brainpy.state_amd.bptt import BPTT, bptt_problem
    net, w_in, w_rec, cnt, target = bptt_problem()
    T, B = cnt.shape[:2]
    print(f'bptt {net.n} neurons, batch {B}, {T} steps', flush=True)
    dev = [torch.tensor(a).cuda() for a in (w_in, w_rec, cnt, target)]
    bp = BPTT(net, batch=B, n_steps=T, checkpoint=100)
    loss, g_in, g_rec = bp.loss_and_grad(*dev)
    bm, counts = bp.spikes()
    bp.close()
    fwd = R.forward_f32(net, w_in, w_rec, cnt, target)
    bits = np.unpackbits(bm.view(np.uint8), bitorder='little').reshape(T, B, -1)[:, :, :net.n]
    assert fwd['spk'].sum() > 100000
    assert np.array_equal(bits.astype(bool), fwd['spk'])
    assert np.array_equal(counts, fwd['counts'])
    assert loss == pytest.approx(fwd['loss'], rel=1e-12)
    # the gradient of the first two trials as their own episode
    c2 = np.ascontiguousarray(cnt[:, :2])
    bp2 = BPTT(net, batch=2, n_steps=T, checkpoint=100)
    l2, gi, gr = bp2.loss_and_grad(dev[0], dev[1], torch.tensor(c2).cuda(), dev[3])
    bp2.close()
    f2 = R.forward_f32(net, w_in, w_rec, c2, target)
    L, gi_ref, gr_ref = R.grad_autograd(net, w_in, w_rec, c2, target, f2)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    assert l2 == pytest.approx(L, rel=1e-6)
    e_in, e_rec = rel(gi.cpu().numpy(), gi_ref), rel(gr.cpu().numpy(), gr_ref)
    print(f'gradient relative error: input {e_in:.2e}, recurrent {e_rec:.2e}', flush=True)
    assert e_in < 1e-4 and e_rec < 1e-4


def test_bptt_4096_batch64_gradient_vs_autograd_on_the_gpu():
    """All 64 trials: the fused adjoint's gradients against torch autograd of
    the same episode in float32 on the same GPU (the benchmark's framework
    reverse-mode leg, which computes its own forward spikes)."""
    torch = pytest.importorskip('torch')
    import bench
    from brainpy.state_amd.bptt import BPTT, bptt_problem
    net, w_in, w_rec, cnt, target = bptt_problem()
    T, B = cnt.shape[:2]
    print(f'bptt {net.n} neurons, batch {B}, {T} steps: fused adjoint vs torch autograd', flush=True)
    dev = [torch.tensor(a).cuda() for a in (w_in, w_rec, cnt, target)]
    bp = BPTT(net, batch=B, n_steps=T, checkpoint=100)
    loss, g_in, g_rec = bp.loss_and_grad(*dev)
    bp.close()
    ad = bench.run_autodiff(net, dev, 1, loss, g_in, g_rec)
    print(f"gradient relative difference: input {ad['grad_rel_diff_in']:.2e}, "
          f"recurrent {ad['grad_rel_diff_rec']:.2e}", flush=True)
    assert ad['loss'] == pytest.approx(loss, rel=1e-5)
    assert ad['grad_rel_diff_in'] < 1e-4 and ad['grad_rel_diff_rec'] < 1e-4
