"""Connectivity sampled on the device: edges that stay in device memory.

``Simulator(connectivity='device')`` realises the random rules with the
engine's samplers (``bs_conn_fixed_indegree``, ``bs_conn_bernoulli_count`` /
``_fill``; the definitions are in ``include/brainspike_hip.h``) instead of the
host's NumPy streams, and ``bs_net_create`` builds its gather segments from
the device arrays in place (``bs_projection.on_device``), so neither the
sampling nor the per-(pre, tile) segment sort touches the host. The edge
lists are a different random draw from the host sampler's (both differ from
the reference's jax.random draw, which cannot be reproduced here); rule
semantics and edge order are the reference's (connectivity.py:81-136).

torch provides the device memory; the engine's calls synchronise before they
return.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L

__all__ = ['DeviceEdges', 'sample_fixed_indegree', 'sample_pairwise_bernoulli']


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise L.EngineUnavailable('device connectivity needs a visible HIP device')
    return torch


class DeviceEdges:
    """Population-local ``pre`` / ``post`` int32 tensors in device memory, in
    the rule's edge order, with the largest in-degree of any post."""

    def __init__(self, pre, post, max_indegree):
        self.pre, self.post = pre, post
        self.max_indegree = int(max_indegree)
        self._host = None

    @property
    def n_edges(self) -> int:
        return int(self.pre.shape[0])

    def host(self):
        """``(pre, post)`` as NumPy int32 arrays (downloaded once)."""
        if self._host is None:
            self._host = (self.pre.cpu().numpy(), self.post.cpu().numpy())
        return self._host


def _map_arg(torch, sel, n):
    """A device copy of a segment -> population index map, or None for the
    identity."""
    if sel is None:
        return None, None
    sel = np.asarray(sel, dtype=np.int32)
    if sel.shape[0] == n and np.array_equal(sel, np.arange(n, dtype=np.int32)):
        return None, None
    t = torch.from_numpy(np.ascontiguousarray(sel)).cuda()
    return t, t.data_ptr()


def sample_fixed_indegree(seed, n_pre, n_post, K, *, exclude_self, pre_map=None, post_map=None) -> DeviceEdges:
    """``K`` edges into every post, pre drawn uniformly with replacement from
    the candidates (connectivity.py:101-136 with multapses)."""
    torch = _torch()
    lib = L.lib()
    E = int(n_post) * int(K)
    pre = torch.empty(max(1, E), dtype=torch.int32, device='cuda')
    post = torch.empty(max(1, E), dtype=torch.int32, device='cuda')
    pm, pm_ptr = _map_arg(torch, pre_map, n_pre)
    qm, qm_ptr = _map_arg(torch, post_map, n_post)
    torch.cuda.synchronize()        # torch's copies and allocations are complete
    L.check(lib.bs_conn_fixed_indegree(int(seed), int(n_pre), int(n_post), int(K), int(bool(exclude_self)),
                                       pm_ptr, qm_ptr, pre.data_ptr(), post.data_ptr()))
    del pm, qm
    return DeviceEdges(pre[:E], post[:E], K if E else 0)


def sample_pairwise_bernoulli(seed, n_pre, n_post, p, *, exclude_self, pre_map=None,
                              post_map=None) -> DeviceEdges:
    """Each ordered pair independently with probability ``p``, row-major
    (connectivity.py:81-99)."""
    import ctypes as C
    torch = _torch()
    lib = L.lib()
    row_ptr = torch.empty(int(n_pre) + 1, dtype=torch.int64, device='cuda')
    pm, pm_ptr = _map_arg(torch, pre_map, n_pre)
    qm, qm_ptr = _map_arg(torch, post_map, n_post)
    torch.cuda.synchronize()
    n = C.c_int64()
    kmax = C.c_int32()
    L.check(lib.bs_conn_bernoulli_count(int(seed), int(n_pre), int(n_post), float(p), int(bool(exclude_self)),
                                        row_ptr.data_ptr(), C.byref(n), C.byref(kmax)))
    E = int(n.value)
    pre = torch.empty(max(1, E), dtype=torch.int32, device='cuda')
    post = torch.empty(max(1, E), dtype=torch.int32, device='cuda')
    torch.cuda.synchronize()
    L.check(lib.bs_conn_bernoulli_fill(int(seed), int(n_pre), int(n_post), float(p), int(bool(exclude_self)),
                                       row_ptr.data_ptr(), pm_ptr, qm_ptr, pre.data_ptr(), post.data_ptr()))
    del pm, qm, row_ptr
    return DeviceEdges(pre[:E], post[:E], kmax.value)
