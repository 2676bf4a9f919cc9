"""One network advanced as several shards that exchange each window's spikes.

This is the single-process form of the distributed step: every shard is an
engine that owns some populations (their state, inputs and updates) and holds
the others as remote populations (``bs_population.remote``). Windows are the
min-delay slices of the step (W <= d_min + 1 steps, so nothing emitted in a
window lands in it); after each window every shard passes its spike list --
``(gid, step)`` pairs sorted by step then gid, the exchange format -- to the
others, which inject the spikes of their remote populations
(``bs_net_inject_spikes``) before their next window. Spikes then reach their
targets through the same buckets and edge segments as local ones, so the
union of the shards' rasters equals the unsharded engine's bit for bit.

Across GPUs the exchange is one all-gather of the window's lists over the
shards (RCCL over xGMI); here the shards share one device and the lists pass
through host memory.
"""
from __future__ import annotations

import numpy as np

from .nodeview import NodeView
from .simulator import Simulator, _Engine, _NetBuilder

__all__ = ['Shards']


def _population(x):
    if isinstance(x, NodeView):
        segs = x.segments
        if len(segs) != 1 or len(np.asarray(segs[0].indices)) != segs[0].population.n:
            raise ValueError('a shard owns whole populations')
        return segs[0].population
    return x


class Shards:
    """``groups``: one list of populations (handles from ``Simulator.create``)
    per shard; every neuron population of ``sim`` in exactly one group."""
    __module__ = 'brainpy.state_amd'

    def __init__(self, sim: Simulator, groups):
        pops = sim._pops
        owner = {}
        for s, grp in enumerate(groups):
            for x in grp:
                pop = _population(x)
                if id(pop) in owner:
                    raise ValueError('a population is assigned to two shards')
                owner[id(pop)] = s
        if len(owner) != len(pops) or any(id(p) not in owner for p in pops):
            raise ValueError('every neuron population must be assigned to one shard')
        for r in sim._records:                 # plastic projections need both ends on one shard
            if getattr(r, 'plastic', False) and owner[id(r.pre_pop)] != owner[id(r.post_pop)]:
                raise NotImplementedError('a plastic projection across shards is not supported')
        self.sim = sim
        self.n_shards = len(groups)
        self.descs = [_NetBuilder(sim, remote=[i for i, p in enumerate(pops) if owner[id(p)] != s]).build()
                      for s in range(self.n_shards)]
        self.engines = [_Engine(d) for d in self.descs]
        ws = {e.window() for e in self.engines}
        if len(ws) != 1:
            self.close()
            raise RuntimeError(f'the shards were given different window lengths {sorted(ws)}; '
                               'set Simulator(max_window=...) to a common bound')
        self.W = ws.pop()

    def reset(self):
        for e in self.engines:
            e.reset()

    def run(self, n_steps):
        """Advance every shard ``n_steps``; returns ``(gid, step, vms)``: the
        union of the shards' spikes sorted by (step, gid), and each shard's
        voltmeter rows (its own taps)."""
        gl, sl = [], []
        vms = [[] for _ in self.engines]
        done = 0
        while done < n_steps:
            w = min(self.W, n_steps - done)
            res = [e.run(w) for e in self.engines]
            spikes = []
            for k, (vm, (g, st)) in enumerate(res):
                vms[k].append(vm)
                spikes.append((g, st))
            for k, e in enumerate(self.engines):
                parts = [spikes[j] for j in range(self.n_shards) if j != k]
                g = np.concatenate([p[0] for p in parts]) if parts else np.zeros(0, np.int32)
                st = np.concatenate([p[1] for p in parts]) if parts else np.zeros(0, np.int32)
                o = np.lexsort((g, st))
                e.inject(g[o], st[o])
            for g, st in spikes:
                gl.append(g)
                sl.append(st)
            done += w
        g = np.concatenate(gl) if gl else np.zeros(0, np.int32)
        st = np.concatenate(sl) if sl else np.zeros(0, np.int32)
        o = np.lexsort((g, st))
        return g[o], st[o], [np.concatenate(v) if v else np.zeros((0, 0)) for v in vms]

    def close(self):
        for e in self.engines:
            e.close()
        self.engines = []
