"""GPU parity of the BPTT adjoint (bs_bptt_*) against the CPU reference.

Forward: spike bitmaps and counts bit-exact vs the float32 restatement.
Backward: weight gradients vs torch float64 autograd of the same episode
(relative error of the gradient <= 1e-4), vs central finite differences on
subthreshold episodes, checkpoint-interval invariance and run-to-run
determinism.
"""
import numpy as np
import pytest

torch = pytest.importorskip('torch')

from oracle import bptt_ref as R
from brainpy.state_amd.bptt import (BPTT, EpisodeLoss, RecurrentIafExp, poisson_counts,
                                    random_sparse_weights)

pytestmark = pytest.mark.gpu


def _problem(n_exc=256, n_inh=64, n_in=20, B=3, T=400, seed=0, rate_in=60.0, loss='rate_mse',
             w_scale=1.0):
    net = RecurrentIafExp(n_in, n_exc, n_inh, dt=0.1, delay=1.0, delay_in=1.0)
    N = n_exc + n_inh
    cnt = poisson_counts(np.full(n_in, rate_in), T, B, 0.1, seed)
    w_in, _ = random_sparse_weights(n_in, N, 0.3, 900.0 * w_scale, 200.0 * w_scale, seed + 1)
    we, _ = random_sparse_weights(n_exc, N, 0.1, 60.0 * w_scale, 15.0, seed + 2)
    wi, _ = random_sparse_weights(n_inh, N, 0.1, 200.0 * w_scale, 40.0, seed + 3, sign=-1.0)
    w_rec = np.concatenate([we, wi])
    rng = np.random.default_rng(seed + 4)
    if loss == 'rate_mse':
        target = rng.uniform(2.0, 20.0, N).astype(np.float32)
    else:
        target = np.full(N, -66.0, np.float32)
    return net, cnt, w_in, w_rec, target


def _dev(*arrs):
    return [torch.tensor(a).cuda() for a in arrs]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize('checkpoint,gemm,replay', [(100, 'bf16x3', True), (37, 'bf16x3', True),
                                                    (37, 'bf16x3', False), (100, 'fp32', 'auto')])
def test_bptt_forward_bitexact_and_gradient_vs_autograd(checkpoint, gemm, replay):
    net, cnt, w_in, w_rec, target = _problem()
    T, B = cnt.shape[:2]
    bp = BPTT(net, batch=B, n_steps=T, checkpoint=checkpoint, gemm=gemm, replay=replay)
    win_d, wrec_d, cnt_d, tgt_d = _dev(w_in, w_rec, cnt, target)
    loss, g_in, g_rec = bp.loss_and_grad(win_d, wrec_d, cnt_d, tgt_d)
    fwd = R.forward_f32(net, w_in, w_rec, cnt, target)
    bm, counts = bp.spikes()
    bits = np.unpackbits(bm.view(np.uint8), bitorder='little').reshape(T, B, -1)[:, :, :net.n]
    assert fwd['spk'].sum() > 200, 'network too quiet for a useful check'
    assert np.array_equal(bits.astype(bool), fwd['spk'])
    assert np.array_equal(counts, fwd['counts'])
    assert loss == pytest.approx(fwd['loss'], rel=1e-12)
    L, gi_ref, gr_ref = R.grad_autograd(net, w_in, w_rec, cnt, target, fwd)
    gi, gr = g_in.cpu().numpy(), g_rec.cpu().numpy()
    assert _rel(gi, gi_ref) < 1e-4 and _rel(gr, gr_ref) < 1e-4, (_rel(gi, gi_ref), _rel(gr, gr_ref))
    assert np.abs(gr_ref).max() > 0


def test_bptt_membrane_loss_matches_finite_differences():
    net, cnt, w_in, w_rec, target = _problem(n_exc=128, n_inh=64, B=2, T=200, loss='membrane_mse',
                                             w_scale=0.05)
    w_rec = np.zeros_like(w_rec)
    T, B = cnt.shape[:2]
    bp = BPTT(net, batch=B, n_steps=T, checkpoint=50, loss='membrane_mse')
    loss, g_in, g_rec = bp.loss_and_grad(*_dev(w_in, w_rec, cnt, target))
    assert loss == pytest.approx(R.loss_f64(net, w_in, w_rec, cnt, target), rel=1e-5)
    gi = g_in.cpu().numpy()
    eps = 0.5
    rng = np.random.default_rng(0)
    for _ in range(6):
        i, j = int(rng.integers(0, net.n_in)), int(rng.integers(0, net.n))
        wp = w_in.astype(np.float64).copy()
        wm = wp.copy()
        wp[i, j] += eps
        wm[i, j] -= eps
        fd = (R.loss_f64(net, wp, w_rec, cnt, target) - R.loss_f64(net, wm, w_rec, cnt, target)) / (2 * eps)
        assert gi[i, j] == pytest.approx(fd, rel=1e-3, abs=1e-7), (i, j)


def test_bptt_deterministic_and_torch_function():
    net, cnt, w_in, w_rec, target = _problem(T=300)
    T, B = cnt.shape[:2]
    bp = BPTT(net, batch=B, n_steps=T, checkpoint=100)
    win_d, wrec_d, cnt_d, tgt_d = _dev(w_in, w_rec, cnt, target)
    l1, a_in, a_rec = bp.loss_and_grad(win_d, wrec_d, cnt_d, tgt_d)
    l2, b_in, b_rec = bp.loss_and_grad(win_d, wrec_d, cnt_d, tgt_d)
    assert l1 == l2
    assert torch.equal(a_in, b_in) and torch.equal(a_rec, b_rec)
    wi = win_d.clone().requires_grad_(True)
    wr = wrec_d.clone().requires_grad_(True)
    loss = EpisodeLoss.apply(wi, wr, bp, cnt_d, tgt_d)
    (3.0 * loss).backward()
    assert float(loss.detach()) == l1
    assert torch.allclose(wi.grad, 3.0 * a_in, rtol=1e-6, atol=0)
    assert torch.allclose(wr.grad, 3.0 * a_rec, rtol=1e-6, atol=0)


@pytest.mark.parametrize('loss', ['rate_mse', 'membrane_mse'])
def test_bptt_segment_replay_gives_the_recompute_bits(loss):
    """Replaying checkpoint segments from the stored summed inputs and
    recomputing them with their event gathers give identical gradients."""
    net, cnt, w_in, w_rec, target = _problem(T=300)
    T, B = cnt.shape[:2]
    if loss == 'membrane_mse':
        target = np.full_like(target, -65.0)
    win_d, wrec_d, cnt_d, tgt_d = _dev(w_in, w_rec, cnt, target)
    out = {}
    for replay in (True, False):
        bp = BPTT(net, batch=B, n_steps=T, checkpoint=64, loss=loss, replay=replay)
        out[replay] = bp.loss_and_grad(win_d, wrec_d, cnt_d, tgt_d)
        bp.close()
    assert out[True][0] == out[False][0]
    assert torch.equal(out[True][1], out[False][1]) and torch.equal(out[True][2], out[False][2])
    assert float(out[True][2].abs().max()) > 0
