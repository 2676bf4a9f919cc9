"""Batched ensembles: many independent networks advanced together on one GPU.

The reference batches trials of one program with ``jax.vmap`` (the kernel
runner's ``trials`` axis, ``_kernel/runner.py:571-700``); here the networks
may differ in size and structure. Each network is built by its own
:class:`~brainpy.state_amd.Simulator` (so its connectivity, generator seeds
and recorders are exactly those of a plain run), the descriptors are
concatenated into one ``bs_net_desc``, and one engine advances all of them:
every window is a single ``k_window`` launch over the tiles of all networks,
and chunks of windows replay as one HIP graph. Results are split back per
network and are identical to the plain runs of the same simulators.

:meth:`Ensemble.autotune` times candidate tile sizes and window bounds on the
networks (or a sample of them) and returns the fastest; the choice does not
change any result.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib as L
from . import units as u
from .simulator import Simulator, _Desc, _Engine

__all__ = ['Ensemble', 'merge_descs']


def _copy(dst, src, ctype):
    C.memmove(C.addressof(dst), C.addressof(src), C.sizeof(ctype))


def merge_descs(descs, *, tile_size=0, max_window=0) -> _Desc:
    """One descriptor holding every network of ``descs`` (population, channel
    and projection indices offset; tables concatenated)."""
    if not descs:
        raise ValueError('an ensemble needs at least one network')
    if len({d.c.dtype for d in descs}) != 1 or len({d.c.dt_ms for d in descs}) != 1:
        raise ValueError('all networks of an ensemble need the same dt and dtype')
    npop = sum(d.c.n_pops for d in descs)
    nch = sum(d.c.n_channels for d in descs)
    nproj = sum(d.c.n_projs for d in descs)
    pops = (L.Population * npop)()
    chans = (L.Channel * max(1, nch))()
    projs = (L.Projection * max(1, nproj))()
    tables, toff = [], []
    M = _Desc()
    M.instances = []
    po = co = ro = to = go = tap = 0
    for D in descs:
        c = D.c
        for p in range(c.n_pops):
            _copy(pops[po + p], c.pops[p], L.Population)
            pops[po + p].ch_begin += co
        for k in range(c.n_channels):
            _copy(chans[co + k], c.channels[k], L.Channel)
            o = int(D.table_off_np[k])
            toff.append(o + to if o >= 0 else -1)
        for r in range(c.n_projs):
            _copy(projs[ro + r], c.projs[r], L.Projection)
            projs[ro + r].pre_pop += po
            projs[ro + r].post_pop += po
            projs[ro + r].channel += co
        tables.append(np.asarray(D.tables_np, dtype=np.float64))
        M.instances.append(dict(gid0=go, ntot=D.ntot, tap0=tap, n_taps=D.n_taps, pop0=po, proj0=ro))
        po += c.n_pops
        co += c.n_channels
        ro += c.n_projs
        to += len(D.tables_np)
        go += D.ntot
        tap += D.n_taps
        M.keep.append(D)  # the per-network arrays the copied structs point into
    tab = np.concatenate(tables)
    d = L.NetDesc()
    d.abi_version = L.BS_ABI_VERSION
    d.dtype = descs[0].c.dtype
    d.dt_ms = descs[0].c.dt_ms
    d.n_pops, d.n_channels, d.n_projs = npop, nch, nproj
    d.max_window = int(max_window)
    d.tile_size = int(tile_size)
    d.pops, d.channels, d.projs = pops, chans, projs
    d.tables = M.arr(tab, C.c_double)
    d.table_off = M.arr(np.asarray(toff or [-1], dtype=np.int64), C.c_int64)
    d.n_tables = len(tab)
    M.keep += [pops, chans, projs]
    M.tables_np = tab
    M.table_off_np = np.asarray(toff or [-1], dtype=np.int64)
    M.c = d
    M.ntot = go
    M.n_taps = tap
    return M


class Ensemble:
    """Independent :class:`Simulator` networks advanced by one engine.

    Parameters
    ----------
    sims : sequence of Simulator
        Fully connected networks with the same ``dt`` and ``dtype``.
    tile_size, max_window : int
        Launch shape of the window kernel (0 = automatic); see :meth:`autotune`.
    """

    def __init__(self, sims, *, tile_size=0, max_window=0):
        self.sims = list(sims)
        if not all(isinstance(s, Simulator) for s in self.sims):
            raise TypeError('an Ensemble is built from Simulator instances')
        self.descs = [s.build_desc() for s in self.sims]
        self.desc = merge_descs(self.descs, tile_size=tile_size, max_window=max_window)
        self.tile_size, self.max_window = int(tile_size), int(max_window)
        L.require_device()
        self._engine = _Engine(self.desc)
        self._ready = False

    @property
    def dt(self):
        return self.sims[0].dt

    def close(self):
        if self._engine is not None:
            self._engine.close()
            self._engine = None

    def reset(self):
        self._engine.reset()
        for s in self.sims:
            s._base_step = 0
        self._ready = True

    def run_raw(self, n_steps):
        """Advance every network ``n_steps``; returns ``(vm, (gids, steps))``
        with ensemble-global gids (network k owns ``desc.instances[k]``'s
        range), spikes sorted by (step, gid)."""
        if not self._ready:
            self.reset()
        return self._engine.run(int(n_steps))

    def split(self, vm, g, st):
        """Per-network ``(vm, gids, steps)`` of a :meth:`run_raw` result."""
        inst = self.desc.instances
        bounds = np.array([i['gid0'] for i in inst] + [self.desc.ntot])
        which = np.searchsorted(bounds, g, side='right') - 1
        order = np.argsort(which, kind='stable')
        g, st, which = g[order], st[order], which[order]
        cuts = np.searchsorted(which, np.arange(len(inst) + 1))
        out = []
        for k, i in enumerate(inst):
            a, b = cuts[k], cuts[k + 1]
            out.append((vm[:, i['tap0']:i['tap0'] + i['n_taps']], g[a:b] - i['gid0'], st[a:b]))
        return out

    def run(self, n_steps):
        """Advance every network ``n_steps``; returns per network
        ``(vm [n_steps, n_taps], gids, steps)`` with network-local gids,
        spikes sorted by (step, gid)."""
        vm, (g, st) = self.run_raw(n_steps)
        return self.split(vm, g, st)

    def simulate(self, duration):
        """Reset every network and run ``duration``; one SimulationResult each."""
        self.reset()
        return self.cont(duration)

    def cont(self, duration):
        T = u.scalar(duration, 'time')
        n = int(np.ceil(T / self.dt))
        parts = self.run(n)
        return [s._result(D, n, T, vm, g, st) for s, D, (vm, g, st) in zip(self.sims, self.descs, parts)]

    def engine_bytes(self):
        return self._engine.bytes()

    # -------------------------------------------------------------- autotune
    @staticmethod
    def autotune(sims, *, candidates=None, steps=1000, warmup=50, repeats=2):
        """Time each (tile_size, max_window) candidate on ``sims`` and return
        ``(best, table)``; ``table`` maps each candidate to seconds per step,
        the best of ``repeats`` timed runs of ``steps`` steps after one reset.
        Candidates whose spike counts differ from the first one's raise (the
        launch shape must not change results)."""
        if candidates is None:
            candidates = [(ts, mw) for ts in (64, 128, 256) for mw in (0, 8)]
        table, ref = {}, None
        for ts, mw in candidates:
            ens = Ensemble(sims, tile_size=ts, max_window=mw)
            try:
                best_t, g = None, None
                for _ in range(max(1, repeats)):
                    ens.reset()
                    ens._engine.run(warmup)
                    t0 = time.perf_counter()
                    _, (g, _) = ens._engine.run(steps)
                    dt_ = (time.perf_counter() - t0) / steps
                    best_t = dt_ if best_t is None else min(best_t, dt_)
                table[(ts, mw)] = best_t
            finally:
                ens.close()
            if ref is None:
                ref = len(g)
            elif len(g) != ref:
                raise RuntimeError(f'launch shape {(ts, mw)} changed the result ({len(g)} vs {ref} spikes)')
        best = min(table, key=table.get)
        return best, table
