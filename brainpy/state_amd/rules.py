"""Connection rules and their host samplers.

Rule values and constructors mirror ``_nest_network/rules.py:24-204``. The
reference samples with ``jax.random`` (threefry, ``choice`` / ``uniform``);
here every rule samples with a NumPy PCG64 stream seeded by the projection's
derived seed, vectorised so that ~10^7-edge networks build in well under a
second. Edge order follows the reference's (post-major for fixed in-degree,
pre-major row order for pairwise Bernoulli and all-to-all).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ['ConnRule', 'all_to_all', 'one_to_one', 'fixed_indegree', 'pairwise_bernoulli',
           'fixed_total_number', 'explicit_edges', 'ConnSpec']


@dataclass
class ConnSpec:
    pre_idx: np.ndarray
    post_idx: np.ndarray
    n_edges: int


def _rng(seed):
    return np.random.Generator(np.random.PCG64(0 if seed is None else int(seed)))


class ConnRule:
    """Base class: ``sample(n_pre, n_post, *, seed, pre_is_post, allow_autapses,
    allow_multapses) -> ConnSpec`` of segment-local indices. Random rules with
    a device sampler also implement ``sample_device(...)`` returning
    :class:`~brainpy.state_amd.device_conn.DeviceEdges` of population-local
    indices (``pre_map`` / ``post_map`` translate the segment-local ones)."""
    __module__ = 'brainpy.state_amd'
    random = False           # draws random numbers (deterministic rules build on the host)

    def sample(self, n_pre, n_post, *, seed, pre_is_post, allow_autapses, allow_multapses):
        raise NotImplementedError

    def sample_device(self, n_pre, n_post, *, seed, pre_is_post, allow_autapses, allow_multapses,
                      pre_map=None, post_map=None):
        raise NotImplementedError(f'{type(self).__name__.lstrip("_")} has no device sampler; '
                                  "use Simulator(connectivity='host')")


class _AllToAll(ConnRule):
    def sample(self, n_pre, n_post, *, seed, pre_is_post, allow_autapses, allow_multapses):
        pre = np.repeat(np.arange(n_pre, dtype=np.int32), n_post)
        post = np.tile(np.arange(n_post, dtype=np.int32), n_pre)
        if pre_is_post and not allow_autapses:
            keep = pre != post
            pre, post = pre[keep], post[keep]
        return ConnSpec(pre, post, int(pre.shape[0]))


class _OneToOne(ConnRule):
    def sample(self, n_pre, n_post, *, seed, pre_is_post, allow_autapses, allow_multapses):
        if n_pre != n_post:
            raise ValueError(f'one_to_one requires equal sizes, got n_pre={n_pre}, n_post={n_post}')
        idx = np.arange(n_pre, dtype=np.int32)
        return ConnSpec(idx, idx.copy(), int(n_pre))


class _FixedIndegree(ConnRule):
    """Each post neuron receives exactly ``K`` edges (connectivity.py:101-140)."""
    random = True

    def __init__(self, K: int):
        self.K = int(K)

    def sample(self, n_pre, n_post, *, seed, pre_is_post, allow_autapses, allow_multapses):
        K = self.K
        if K < 0:
            raise ValueError(f'K must be >= 0, got {K}')
        rng = _rng(seed)
        excl = pre_is_post and not allow_autapses
        n_cand = n_pre - 1 if excl else n_pre
        post = np.repeat(np.arange(n_post, dtype=np.int32), K)
        if K == 0 or n_post == 0:
            return ConnSpec(np.zeros(0, np.int32), np.zeros(0, np.int32), 0)
        if allow_multapses:
            if n_cand <= 0:
                raise ValueError('no candidate presynaptic neurons')
            pre = rng.integers(0, n_cand, size=(n_post, K), dtype=np.int64)
        else:
            if K > n_cand:
                raise ValueError(f'cannot pick {K} unique pre for a post from {n_cand} '
                                 'candidates with allow_multapses=False')
            pre = np.empty((n_post, K), dtype=np.int64)
            for j in range(n_post):
                pre[j] = rng.choice(n_cand, K, replace=False)
        if excl:   # candidates are arange(0, j) ++ arange(j+1, n_pre)
            pre = pre + (pre >= np.arange(n_post)[:, None])
        return ConnSpec(pre.reshape(-1).astype(np.int32), post, int(K * n_post))

    def sample_device(self, n_pre, n_post, *, seed, pre_is_post, allow_autapses, allow_multapses,
                      pre_map=None, post_map=None):
        from .device_conn import sample_fixed_indegree
        if self.K < 0:
            raise ValueError(f'K must be >= 0, got {self.K}')
        if not allow_multapses:
            raise NotImplementedError('the device fixed_indegree sampler draws with replacement; '
                                      "use allow_multapses=True or Simulator(connectivity='host')")
        excl = pre_is_post and not allow_autapses
        if self.K > 0 and n_post > 0 and (n_pre - 1 if excl else n_pre) <= 0:
            raise ValueError('no candidate presynaptic neurons')
        return sample_fixed_indegree(seed, n_pre, n_post, self.K, exclude_self=excl,
                                     pre_map=pre_map, post_map=post_map)


class _PairwiseBernoulli(ConnRule):
    """Each ordered pair connected with probability ``p`` (connectivity.py:81-99)."""
    random = True

    def __init__(self, p: float):
        self.p = float(p)

    def sample(self, n_pre, n_post, *, seed, pre_is_post, allow_autapses, allow_multapses):
        p = self.p
        total = int(n_pre) * int(n_post)
        if p <= 0.0 or total == 0:
            return ConnSpec(np.zeros(0, np.int32), np.zeros(0, np.int32), 0)
        rng = _rng(seed)
        if p >= 1.0:
            flat = np.arange(total, dtype=np.int64)
        else:
            # geometric gaps between successes over the row-major pair index
            parts = []
            pos = -1
            est = int(total * p + 10 * np.sqrt(total * p) + 16)
            while pos < total:
                gaps = rng.geometric(p, size=est).astype(np.int64)
                idx = pos + np.cumsum(gaps)
                parts.append(idx[idx < total])
                pos = int(idx[-1])
            flat = np.concatenate(parts) if parts else np.zeros(0, np.int64)
        pre = (flat // n_post).astype(np.int32)
        post = (flat % n_post).astype(np.int32)
        if pre_is_post and not allow_autapses:
            keep = pre != post
            pre, post = pre[keep], post[keep]
        return ConnSpec(pre, post, int(pre.shape[0]))

    def sample_device(self, n_pre, n_post, *, seed, pre_is_post, allow_autapses, allow_multapses,
                      pre_map=None, post_map=None):
        from .device_conn import sample_pairwise_bernoulli
        return sample_pairwise_bernoulli(seed, n_pre, n_post, self.p,
                                         exclude_self=pre_is_post and not allow_autapses,
                                         pre_map=pre_map, post_map=post_map)


class _FixedTotalNumber(ConnRule):
    """Exactly ``N`` edges drawn uniformly over the (pre, post) grid."""
    random = True

    def __init__(self, N: int):
        self.N = int(N)

    def sample(self, n_pre, n_post, *, seed, pre_is_post, allow_autapses, allow_multapses):
        rng = _rng(seed)
        N = self.N
        excl = pre_is_post and not allow_autapses
        total = n_pre * n_post - (min(n_pre, n_post) if excl else 0)
        if not allow_multapses and N > total:
            raise ValueError(f'cannot draw {N} unique edges from {total} pairs')
        if allow_multapses:
            pre = np.empty(0, np.int64)
            post = np.empty(0, np.int64)
            while pre.shape[0] < N:
                a = rng.integers(0, n_pre, size=N)
                b = rng.integers(0, n_post, size=N)
                if excl:
                    keep = a != b
                    a, b = a[keep], b[keep]
                pre = np.concatenate([pre, a])
                post = np.concatenate([post, b])
            pre, post = pre[:N], post[:N]
        else:
            flat = rng.choice(n_pre * n_post, size=min(n_pre * n_post, N + (n_pre if excl else 0)),
                              replace=False)
            a, b = flat // n_post, flat % n_post
            if excl:
                keep = a != b
                a, b = a[keep], b[keep]
            pre, post = a[:N], b[:N]
        return ConnSpec(pre.astype(np.int32), post.astype(np.int32), int(N))


class _ExplicitEdges(ConnRule):
    """A precomputed edge list returned verbatim (rules.py:139-204)."""

    def __init__(self, spec: ConnSpec):
        self._spec = spec

    def sample(self, n_pre, n_post, *, seed, pre_is_post, allow_autapses, allow_multapses):
        s = self._spec
        if s.n_edges and (s.pre_idx.max() >= n_pre or s.post_idx.max() >= n_post or
                          s.pre_idx.min() < 0 or s.post_idx.min() < 0):
            raise ValueError('explicit edge index out of range for the connected segments')
        return s


all_to_all = _AllToAll()
one_to_one = _OneToOne()


def fixed_indegree(K: int) -> _FixedIndegree:
    if int(K) < 0:
        raise ValueError(f'K must be >= 0, got {K}')
    return _FixedIndegree(int(K))


def pairwise_bernoulli(p: float) -> _PairwiseBernoulli:
    p = float(p)
    if not (0.0 <= p <= 1.0):
        raise ValueError(f'p must be in [0, 1], got {p}')
    return _PairwiseBernoulli(p)


def fixed_total_number(N: int) -> _FixedTotalNumber:
    if int(N) < 0:
        raise ValueError(f'N must be >= 0, got {N}')
    return _FixedTotalNumber(int(N))


def explicit_edges(pre_idx, post_idx) -> _ExplicitEdges:
    pre = np.asarray(pre_idx)
    post = np.asarray(post_idx)
    if pre.ndim != 1 or post.ndim != 1:
        raise ValueError(f'pre_idx/post_idx must be 1-D, got {pre.ndim}-D and {post.ndim}-D')
    if pre.shape[0] != post.shape[0]:
        raise ValueError(f'pre_idx/post_idx must have equal length, got {pre.shape[0]} and '
                         f'{post.shape[0]}')
    if not (np.issubdtype(pre.dtype, np.integer) and np.issubdtype(post.dtype, np.integer)):
        raise ValueError(f'pre_idx/post_idx must be integer arrays, got dtypes {pre.dtype} '
                         f'and {post.dtype}')
    return _ExplicitEdges(ConnSpec(pre.astype(np.int32), post.astype(np.int32), int(pre.shape[0])))
