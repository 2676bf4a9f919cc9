"""Pins of the oracle's connectivity samplers (the device samplers' definition).

The reference samples with jax.random, which is absent here, so its exact edge
lists are parity-unpinned; what its tests fix are properties
(connectivity_test.py:48-64: Bernoulli density 0.1 +- 0.02 on 100 x 100,
exactly K edges per post for fixed in-degree) and the rule semantics of
connectivity.py:81-136 (uniform choice with replacement over the candidates,
the diagonal excluded without autapses, row-major Bernoulli edges). Those are
checked below, together with a Python big-integer evaluation of the index map
on tiny grids over the KAT-pinned Threefry, which catches a wrong key, counter
order, edge order or rounding in the C restatement.
"""
import numpy as np
import pytest
from scipy import stats

import oracle

TAG = 0xE36F6E6E


def _key(seed):
    return seed & 0xFFFFFFFF, TAG ^ (seed >> 32)


def _indegree_py(seed, n_pre, n_post, K, excl):
    k0, k1 = _key(seed)
    n_cand = n_pre - 1 if excl else n_pre
    pre, post = [], []
    for j in range(n_post):
        for k in range(K):
            x0, x1 = oracle.threefry2x32(k0, k1, j, k)
            c = ((x1 << 32 | x0) * n_cand) >> 64
            if excl and c >= j:
                c += 1
            pre.append(c)
            post.append(j)
    return np.array(pre), np.array(post)


def _bernoulli_py(seed, n_pre, n_post, p, excl):
    k0, k1 = _key(seed)
    pre, post = [], []
    for i in range(n_pre):
        for j in range(n_post):
            if excl and i == j:
                continue
            x0, _ = oracle.threefry2x32(k0, k1, i, j)
            if x0 < p * 2 ** 32:           # the uniform x0 * 2^-32 < p, exactly
                pre.append(i)
                post.append(j)
    return np.array(pre, dtype=np.int64), np.array(post, dtype=np.int64)


@pytest.mark.parametrize('seed,n_pre,n_post,K,excl', [
    (1, 7, 5, 3, False), (2, 7, 7, 4, True), (2 ** 40 + 17, 1000, 3, 5, True), (9, 2, 6, 2, True),
    (5, 1, 4, 3, False)])
def test_fixed_indegree_matches_big_integer_map(seed, n_pre, n_post, K, excl):
    pre, post = oracle.conn_fixed_indegree(seed, n_pre, n_post, K, excl)
    rp, rq = _indegree_py(seed, n_pre, n_post, K, excl)
    assert np.array_equal(pre, rp) and np.array_equal(post, rq)


@pytest.mark.parametrize('seed,n_pre,n_post,p,excl', [
    (3, 9, 11, 0.3, False), (4, 8, 8, 0.5, True), (2 ** 33 + 5, 5, 40, 0.05, False),
    (6, 6, 6, 0.999, True)])
def test_bernoulli_matches_big_integer_map(seed, n_pre, n_post, p, excl):
    pre, post = oracle.conn_bernoulli(seed, n_pre, n_post, p, excl)
    rp, rq = _bernoulli_py(seed, n_pre, n_post, p, excl)
    assert np.array_equal(pre, rp) and np.array_equal(post, rq)


def test_fixed_indegree_each_post_has_K():
    """connectivity_test.py:57-64, with multapses allowed."""
    pre, post = oracle.conn_fixed_indegree(1, 50, 20, 10, False)
    assert np.array_equal(np.bincount(post, minlength=20), np.full(20, 10))
    assert pre.min() >= 0 and pre.max() < 50
    assert np.array_equal(post, np.repeat(np.arange(20), 10))     # post-major order


def test_fixed_indegree_self_exclusion_and_uniform_candidates():
    """No autapse, and each post's candidates (all pre but itself) drawn
    uniformly: chi-square over the candidate rank."""
    n, K = 40, 2000
    pre, post = oracle.conn_fixed_indegree(11, n, n, K, True)
    assert not np.any(pre == post)
    rank = pre - (pre > post)                                        # index among the n - 1 candidates
    counts = np.bincount(rank, minlength=n - 1)
    assert stats.chisquare(counts).pvalue > 1e-4
    # both ends of the candidate range are reachable
    assert pre[post == 0].min() == 1 and pre[post == n - 1].max() == n - 2


def test_fixed_indegree_draws_are_uncorrelated():
    pre, post = oracle.conn_fixed_indegree(12, 1000, 200, 100, False)
    a = pre.reshape(200, 100).astype(float)
    r = np.corrcoef(a[:, :-1].ravel(), a[:, 1:].ravel())[0, 1]
    assert abs(r) < 0.02
    assert stats.chisquare(np.bincount(pre, minlength=1000)).pvalue > 1e-4


def test_bernoulli_density_and_order():
    """connectivity_test.py:48-55 (density 0.1 +- 0.02 on 100 x 100), plus
    row-major order and a binomial bound on the count."""
    pre, post = oracle.conn_bernoulli(0, 100, 100, 0.1, False)
    assert abs(len(pre) / 1e4 - 0.1) < 0.02
    flat = pre.astype(np.int64) * 100 + post
    assert np.all(np.diff(flat) > 0)
    n_big = len(oracle.conn_bernoulli(1, 400, 500, 0.01, False)[0])
    assert stats.binomtest(n_big, 200000, 0.01).pvalue > 1e-4


def test_bernoulli_grids_and_diagonal():
    pre, post = oracle.conn_bernoulli(5, 7, 9, 1.0, False)
    assert len(pre) == 63 and np.array_equal(pre * 9 + post, np.arange(63))
    pre, post = oracle.conn_bernoulli(5, 7, 7, 1.0, True)
    assert len(pre) == 42 and not np.any(pre == post)
    assert len(oracle.conn_bernoulli(5, 7, 9, 0.0, False)[0]) == 0
    pre, post = oracle.conn_bernoulli(8, 300, 300, 0.5, True)
    assert not np.any(pre == post)
    # the diagonal is skipped, not resampled: off-diagonal pairs keep their draws
    p2, q2 = oracle.conn_bernoulli(8, 300, 300, 0.5, False)
    keep = p2 != q2
    assert np.array_equal(p2[keep], pre) and np.array_equal(q2[keep], post)


def test_seed_changes_edges_and_is_reproducible():
    a = oracle.conn_fixed_indegree(21, 100, 10, 5)
    b = oracle.conn_fixed_indegree(21, 100, 10, 5)
    c = oracle.conn_fixed_indegree(22, 100, 10, 5)
    assert np.array_equal(a[0], b[0]) and not np.array_equal(a[0], c[0])
    # the high seed word enters the key
    d = oracle.conn_fixed_indegree(21 + 2 ** 32, 100, 10, 5)
    assert not np.array_equal(a[0], d[0])
