"""Connection introspection: ``Simulator.get_connections`` and ``SynapseCollection``.

Mirrors the reference's NEST-style view (``connection_introspection.py:362-571``,
``simulator.py:832-866``): a lazy, filtered view over realised neuron-to-neuron
projections that reads ``source`` / ``target`` / ``weight`` / ``delay`` live
and writes ``weight`` / ``delay`` back with the same guards.

Edge order is the canonical one (``canonical_order``,
``connection_introspection.py:104-112``): a stable sort by population-local
source, which keeps each source's edges in their sampled order. Weights are
floats in the post population's NEST unit (mV for iaf_psc_delta, pA for the
current-based models, nS for aeif_cond_alpha); delays are in ms.

The engine builds its delivery tables (per-(pre, tile) segments, count tables,
fixed-point weights) from the projections when it is created, so a write marks
the network changed: the next ``simulate`` builds a new engine from the edited
projections (``simulate`` starts a new rollout, as in the reference), and
``cont`` refuses to continue a rollout whose connections changed under it.
"""
from __future__ import annotations

import numpy as np

from . import units as u

__all__ = ['SynapseCollection', 'collect_connections']

# weight-evolving plastic rules refuse weight writes (connection_introspection.py:297-326)
_STATIC_PLASTIC_MODELS = ('static_synapse', 'static_synapse_hom_w')


def _membership(view):
    """``{id(population): set(local indices)}`` of a NodeView, ``None`` for all."""
    if view is None:
        return None
    out = {}
    for seg in view.segments:
        out.setdefault(id(seg.population), set()).update(int(i) for i in np.asarray(seg.indices))
    return out


def _model_name(proj):
    return type(proj.rule).__name__ if proj.plastic else 'static_synapse'


class _Spanned:
    """One projection's share of a collection: canonical positions kept, and
    their population-local source / target."""

    def __init__(self, proj, order, kept):
        self.proj = proj
        self.order = order                      # canonical position -> sampled position
        self.kept = kept                        # canonical positions in the view
        self.source = np.asarray(proj.pre_idx)[order][kept]
        self.target = np.asarray(proj.post_idx)[order][kept]


def collect_connections(sim, source=None, target=None, synapse=None):
    """The ``SynapseCollection`` of ``sim``'s projections matching the filters
    (``collect_connections``, connection_introspection.py:362-406)."""
    from .simulator import StaticProjection
    src_m, tgt_m = _membership(source), _membership(target)
    spanned = []
    for proj in sim._records:
        if not isinstance(proj, StaticProjection):
            continue                            # generator channels are not synapses
        if synapse is not None and _model_name(proj) != synapse:
            continue
        if src_m is not None and id(proj.pre_pop) not in src_m:
            continue
        if tgt_m is not None and id(proj.post_pop) not in tgt_m:
            continue
        pre = np.asarray(proj.pre_idx)
        order = np.argsort(pre, kind='stable')
        keep = np.ones(pre.shape[0], dtype=bool)
        if src_m is not None:
            keep &= np.isin(pre[order], np.fromiter(src_m[id(proj.pre_pop)], dtype=np.int64))
        if tgt_m is not None:
            post = np.asarray(proj.post_idx)[order]
            keep &= np.isin(post, np.fromiter(tgt_m[id(proj.post_pop)], dtype=np.int64))
        kept = np.nonzero(keep)[0]
        if kept.shape[0]:
            spanned.append(_Spanned(proj, order, kept))
    return SynapseCollection(sim, spanned)


class SynapseCollection:
    """A filtered, lazy view over realised synapses (NEST ``SynapseCollection``).

    ``get`` reads weights and delays live: a plastic projection's weights come
    from the engine once one exists (post-simulation evolved values), else from
    their initial values. ``set('weight', v)`` takes a scalar (broadcast) or one
    value per edge of the view; it is refused on a weight-evolving plastic
    projection, and a per-edge value is refused on a homogeneous-weight
    projection. ``set('delay', v)`` takes a scalar, rounded to the ``dt`` grid,
    and applies to static projections only.
    """
    __module__ = 'brainpy.state_amd'

    def __init__(self, sim, spanned):
        self._sim = sim
        self._spanned = list(spanned)

    def __len__(self):
        return int(sum(sp.kept.shape[0] for sp in self._spanned))

    @property
    def source(self):
        """Population-local source index per edge."""
        if not self._spanned:
            return np.zeros(0, dtype=np.int64)
        return np.concatenate([sp.source for sp in self._spanned]).astype(np.int64)

    @property
    def target(self):
        """Population-local target index per edge."""
        if not self._spanned:
            return np.zeros(0, dtype=np.int64)
        return np.concatenate([sp.target for sp in self._spanned]).astype(np.int64)

    def __iter__(self):
        for s, t in zip(self.source.tolist(), self.target.tolist()):
            yield (s, t)

    def __repr__(self):
        return f'SynapseCollection({len(self)} edges over {len(self._spanned)} projection(s))'

    # -- read ------------------------------------------------------------
    def get(self, key):
        """``'source'``, ``'target'``, ``'weight'`` or ``'delay'`` (or a list of
        them, giving a dict)."""
        if isinstance(key, (list, tuple)):
            return {k: self.get(k) for k in key}
        if key == 'source':
            return self.source
        if key == 'target':
            return self.target
        if key == 'weight':
            parts = [self._weights(sp) for sp in self._spanned]
        elif key == 'delay':
            parts = [self._delays(sp) for sp in self._spanned]
        else:
            raise KeyError(f"unknown connection key {key!r}; expected one of "
                           "'source', 'target', 'weight', 'delay'")
        return np.concatenate(parts) if parts else np.zeros(0)

    def _weights(self, sp):
        p = sp.proj
        if p.plastic and self._sim._engine is not None:
            w = self._sim.get_weights(p)        # canonical (stable sorted-by-pre) order
            return np.asarray(w, dtype=np.float64)[sp.kept]
        if p.homogeneous:
            return np.full(sp.kept.shape[0], float(p.weight))
        return np.asarray(p.weight, dtype=np.float64)[sp.order][sp.kept]

    def _delays(self, sp):
        d = np.asarray(sp.proj.delay, dtype=np.float64)
        fr = getattr(sp.proj, 'delay_frac', None)
        if fr is not None:                      # an InputDelay read between two frames
            d = d + np.asarray(fr, dtype=np.float64)
        steps = np.full(sp.kept.shape[0], float(d)) if np.ndim(d) == 0 else d[sp.order][sp.kept]
        return steps * self._sim.dt

    # -- write -----------------------------------------------------------
    def set(self, key, value):
        """Write ``'weight'`` or ``'delay'`` back into the projections."""
        if key == 'weight':
            self._set_weight(value)
        elif key == 'delay':
            self._set_delay(value)
        else:
            raise KeyError(f"cannot set {key!r}; only 'weight' and 'delay' are settable")

    def _set_weight(self, value):
        is_scalar = np.ndim(value) == 0
        for sp in self._spanned:               # validate all before writing any
            p = sp.proj
            if p.plastic and _model_name(p) not in _STATIC_PLASTIC_MODELS:
                raise ValueError(
                    f"weight of the '{_model_name(p)}' projection is managed by its plastic rule "
                    "and is overwritten during simulate(); set it through the synapse spec at "
                    "connect() time, not SynapseCollection.set().")
            if not is_scalar and p.homogeneous:
                raise ValueError(
                    "cannot set a per-edge weight on a homogeneous-weight projection (it shares a "
                    "single scalar weight); pass a scalar to set them all, or rebuild it with a "
                    "per-edge rule.")
        if not is_scalar and np.shape(value) != (len(self),):
            raise ValueError(f'expected a scalar or {len(self)} per-edge weights, got shape '
                             f'{np.shape(value)}')
        off = 0
        for sp in self._spanned:
            p = sp.proj
            m = int(sp.kept.shape[0])
            dim = _weight_unit(p.post_pop)
            if is_scalar:
                w = float(u.magnitude(value, dim))
                if p.homogeneous:
                    p.weight = w
                else:
                    arr = np.array(p.weight, dtype=np.float64)
                    arr[sp.order[sp.kept]] = w
                    p.weight = arr
            else:
                chunk = np.asarray(u.magnitude(value, dim), dtype=np.float64).reshape(-1)[off:off + m]
                arr = np.array(p.weight, dtype=np.float64)
                arr[sp.order[sp.kept]] = chunk
                p.weight = arr
            off += m
        self._sim._connections_changed()

    def _set_delay(self, value):
        if np.ndim(value) != 0:
            raise ValueError('set("delay", ...) takes a scalar (a homogeneous delay per projection)')
        for sp in self._spanned:
            if sp.proj.plastic:
                raise ValueError(f"delay of the '{_model_name(sp.proj)}' projection is not settable "
                                 "via SynapseCollection (a plastic delay is fixed at connect() time).")
        from .simulator import input_delay_steps
        q_ms = float(u.magnitude(value, 'time'))
        # rounded to the dt grid, then the projection's InputDelay is rebuilt
        # with that delay (connection_introspection.py:132-151), which reads
        # it as delay.py:114-128 does: steps * dt / dt can carry a fraction in
        # the network dtype (7 * 0.1 / 0.1 = 7.000000000000001)
        steps = max(int(round(q_ms / self._sim.dt)), 0)
        rounded = steps * self._sim.dt
        lo, frac = input_delay_steps(rounded, self._sim.dt, self._sim.dtype)
        for sp in self._spanned:
            if frac != 0 and not sp.proj.graded and sp.proj.pre_pop.ref_count(self._sim.dt) < 1:
                raise NotImplementedError(
                    "comm='sparse' with a fractional delay needs a presynaptic refractory period of "
                    'at least one step')
        for sp in self._spanned:
            sp.proj.delay = lo
            sp.proj.delay_frac = frac if frac != 0 else None
            if frac != 0 and sp.proj.device_edges is not None:
                sp.proj.pre_idx                 # the two-frame edges are built from host arrays
                sp.proj.device_edges = None
        self._sim._connections_changed()
        return rounded


def _weight_unit(pop):
    from .simulator import _weight_dim
    return _weight_dim(pop)
