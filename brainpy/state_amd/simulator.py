"""Simulator -- the NEST-style create / connect / simulate front end.

Drop-in for the reference's ``Simulator`` (``_nest_network/simulator.py``):
same method names, argument meanings, one-step holder lag, projection-order
input summation and recorder taps. Building is host-side (connectivity is
sampled here, once); ``simulate`` / ``cont`` hand a flat network descriptor to
``libbrainspike_hip.so`` (``bs_net_create``) and advance it on the GPU
(``bs_net_run``). There is no CPU fallback: without the HIP library or a GPU,
running raises :class:`~brainpy.state_amd._lib.EngineUnavailable`.
"""
from __future__ import annotations

import ctypes as C
import itertools
import math

import numpy as np

from . import units as u
from . import _lib as L
from .devices import dc_generator, multimeter, poisson_generator, spike_recorder, step_current_generator
from .dist import Distribution
from .models import Neuron, aeif_cond_alpha, iaf_psc_delta
from .nodeview import NodeView, Selection, population_size
from .plasticity import stdp_synapse
from .rules import ConnRule, _OneToOne, all_to_all

__all__ = ['Simulator', 'SimulationResult', 'StaticProjection', 'PlasticProjection']


# --------------------------------------------------------------------- helpers
class _GeneratorSpec:
    def __init__(self, model_cls, params):
        self.model_cls = model_cls
        self.params = params


class _GenSegment(Selection):
    """A generator not yet realised: its spec, no population or indices."""

    def __init__(self, spec):
        super().__init__(None, ())
        self.spec = spec


def _is_generator(model_cls) -> bool:
    name = getattr(model_cls, '__name__', '')
    return getattr(model_cls, '_is_generator', False) or 'generator' in name


def _weight_dim(pop):
    if isinstance(pop, iaf_psc_delta):
        return 'voltage'
    if isinstance(pop, aeif_cond_alpha):
        return 'conductance'
    return 'current'


def _map_local(sel, seg_idx):
    """Population-local indices of segment-local ones (``sel[seg_idx]``, int32);
    a segment that is the whole population in order maps to itself."""
    seg_idx = np.asarray(seg_idx)
    n = sel.shape[0]
    if n and sel[0] == 0 and sel[-1] == n - 1 and (n < 3 or np.array_equal(sel, np.arange(n))):
        return np.asarray(seg_idx, dtype=np.int32)
    return sel[np.asarray(seg_idx, dtype=np.int64)].astype(np.int32, copy=False)


def nest_delay_steps(delay, dt):
    """Per-edge delays (ms) -> whole steps with NEST's rounding, as the
    reference's per-edge delay path converts them (_kernel/delays.py:50-108):
    ``floor(d * (1/dt) + 0.5)`` (round half up) in float64, with 1/dt snapped
    to the nearest integer when it is within 1e-6 of one."""
    d = np.asarray(u.magnitude(delay, 'time'), dtype=np.float64)
    if d.size and (not np.all(np.isfinite(d)) or np.min(d) < 0):
        raise ValueError(f'delays must be finite and non-negative, got min {np.min(d)} ms')
    spm = 1.0 / float(dt)
    k = np.rint(spm)
    if k >= 1 and abs(spm - k) <= 1e-6 * k:
        spm = k
    out = np.floor(d * spm + 0.5)
    if out.size and out.max() >= (1 << 24):
        raise ValueError('delay must be below 2^24 steps')
    return int(out) if out.ndim == 0 else out.astype(np.int32)


def input_delay_steps(delay, dt, dtype=np.float64):
    """A homogeneous or per-pre-neuron (axonal) delay as the reference's
    InputDelay reads it (delay.py:114-128): ``steps = delay / dt`` in the
    network dtype, frames ``lo = floor(steps)`` and ``lo + 1``, output
    ``x[lo] * (1 - frac) + x[lo + 1] * frac`` with ``frac = steps - lo``.

    Returns ``(lo, frac)``: ``lo`` int (scalar) or int32 array, ``frac`` in
    ``dtype`` (scalar or array), 0 where ``delay / dt`` is integral. Ratios
    that are integral in exact arithmetic need not be in floating point (in
    float64, 1.2 / 0.1 = 11.999999999999998): those lerp too, as the
    reference does."""
    if delay is None:
        return 0, dtype(0.0)
    d = np.asarray(u.magnitude(delay, 'time'), dtype=np.float64)
    if d.size and (not np.all(np.isfinite(d)) or np.min(d) < 0):
        raise ValueError('delays must be >= 0')
    steps = d.astype(dtype) / dtype(dt)
    lo = np.floor(steps)
    frac = (steps - lo).astype(dtype)
    if lo.size and lo.max() >= (1 << 24) - 1:
        raise ValueError('delay must be below 2^24 steps')
    if lo.ndim == 0:
        return int(lo), dtype(frac)
    return lo.astype(np.int32), frac


def seq_sum_table(w, kmax, dtype):
    """S[k] = ((0 + w) + w) + ... (k terms), accumulated in ``dtype``.

    The CSR event matmul over equal per-edge weights produces exactly this
    value for a post neuron with k arriving events, whatever their order.
    """
    wt = dtype(w)
    out = np.zeros(kmax + 1, dtype=dtype)
    acc = dtype(0.0)
    for k in range(1, kmax + 1):
        acc = dtype(acc + wt)
        out[k] = acc
    return out.astype(np.float64)


# ---------------------------------------------------------------- projections
class StaticProjection:
    """Handle of a realised static projection (reference: EventProjection).

    ``pre_idx`` / ``post_idx`` are population-local indices in the sampled
    order. Edges sampled on the device (``device_edges``) stay there; the
    host arrays are downloaded the first time they are read."""
    __module__ = 'brainpy.state_amd'
    plastic = False

    def __init__(self, ordinal, pre_pop, post_pop, pre, post, weight, delay, receptor, device_edges=None):
        self.ordinal = ordinal
        self.pre_pop, self.post_pop = pre_pop, post_pop
        self._pre, self._post = pre, post
        self.device_edges = device_edges
        self.weight = weight                         # float or (E,) array
        self.delay = delay                           # int steps or (E,) int array
        self.receptor = receptor
        # fractional part of an InputDelay read (scalar or (E,) array in the
        # network dtype; None when every delay is whole steps): the edge then
        # delivers from frames delay and delay + 1 (delay.py:114-128), as two
        # events of its weight (comm='sparse', binary) or as the weighted
        # parts w (1 - frac) and w frac (comm='dense' and receptor scatter)
        self.delay_frac = None
        self.graded = False

    @property
    def pre_idx(self):
        if self._pre is None:
            self._pre, self._post = self.device_edges.host()
        return self._pre

    @property
    def post_idx(self):
        if self._post is None:
            self._pre, self._post = self.device_edges.host()
        return self._post

    @property
    def n_edges(self):
        if self._pre is None:
            return self.device_edges.n_edges
        return int(self._pre.shape[0])

    @property
    def homogeneous(self):
        return np.ndim(self.weight) == 0


class PlasticProjection(StaticProjection):
    """Handle of a realised plastic projection (reference: EventPlasticProj)."""
    __module__ = 'brainpy.state_amd'
    plastic = True

    def __init__(self, *args, rule=None, **kw):
        super().__init__(*args, **kw)
        self.rule = rule


class _CurChannel:
    """A current generator wired into a population's current-input seam
    (simulator.py:1612-1642): no projection, no delay."""

    def __init__(self, ordinal, gen, post_pop, post_idx, weight):
        self.ordinal = ordinal
        self.gen = gen
        self.post_pop = post_pop
        self.post_idx = post_idx
        self.weight = weight            # None: the generator's current as is
        self.receptor = 0


class _GenChannel:
    def __init__(self, ordinal, gen, post_pop, post_idx, weight, delay, receptor, seed):
        self.ordinal = ordinal
        self.gen = gen
        self.post_pop = post_pop
        self.post_idx = post_idx
        self.weight = weight
        self.delay = delay
        self.receptor = receptor
        self.seed = seed
        self.delay_frac = 0.0   # InputDelay lerp weight of frame delay + 1


# --------------------------------------------------------------------- result
class SimulationResult:
    """Recorded spikes and analog traces of one ``simulate`` / ``cont`` window."""
    __module__ = 'brainpy.state_amd'

    def __init__(self, recordings, duration_ms, dt, traces=None, times=None, weights=None):
        self._rec = recordings
        self._T = duration_ms
        self._dt = dt
        self._traces = dict(traces or {})
        self._times = times
        self._weights = dict(weights or {})

    @staticmethod
    def _key(node):
        if isinstance(node, NodeView):
            return id(node.segments[0].population)
        return id(node)

    def spikes(self, node):
        """Per-step spike matrix ``(n_steps, n_recorded)`` of a recorder."""
        return self._rec[self._key(node)]

    def n_events(self, node) -> int:
        return int(np.sum(self._rec[self._key(node)] > 0))

    def rate(self, node) -> float:
        """Mean rate in spikes/s over the recorded neurons (simulator.py:343-348)."""
        spk = self._rec[self._key(node)]
        return float(np.sum(spk > 0)) / spk.shape[1] / (self._T / 1000.0)

    def trace(self, recorder, recordable='V_m'):
        rid = self._key(recorder)
        key = f'{rid}|{recordable}'
        if key not in self._traces:
            available = sorted(k.split('|', 1)[1] for k in self._traces if k.startswith(f'{rid}|'))
            raise KeyError(f'recordable {recordable!r} was not recorded by this recorder; '
                           f'recorded: {available}')
        return self._traces[key]

    def weight_trace(self, proj):
        rid = id(proj)
        if rid not in self._weights:
            raise KeyError("this projection's weight was not recorded; call "
                           'sim.record_weight(proj) before simulate()')
        return self._weights[rid]

    @property
    def times(self):
        return self._times


# ------------------------------------------------------------------ simulator
class Simulator:
    """Explicit NEST-flavoured network builder and runner on the MI355X engine.

    Parameters
    ----------
    dt : Quantity or float
        Time step (ms).
    dtype : {'float64', 'float32'}
        Arithmetic type of the neuron state and synaptic sums (the reference's
        ``brainstate.environ`` precision).
    max_window : int
        Upper bound on the steps advanced per update launch (0 = automatic).
    tile_size : int
        Post neurons per block of the window kernel: 32, 64, 128 or 256 (0 =
        automatic). Results do not depend on it; ``Ensemble.autotune`` picks it.
    connectivity : {'host', 'device'}
        Where the random rules (fixed_indegree, pairwise_bernoulli) draw their
        edges: the host's NumPy streams, or the engine's counter-based device
        samplers, whose edges stay in device memory through engine creation.
        The two give different (equally distributed) networks for one seed.
    plastic_records : int
        Plastic arrivals one block of the window kernel holds at once (0 =
        what its LDS leaves). A window with more arrivals runs in
        sub-windows; results do not depend on it.
    """
    __module__ = 'brainpy.state_amd'

    def __init__(self, *, dt, dtype='float64', max_window=0, tile_size=0, connectivity='host',
                 plastic_records=0):
        self._dt_q = dt
        self.dt = u.scalar(dt, 'time')
        if self.dt <= 0:
            raise ValueError('dt must be > 0')
        if dtype in ('float64', 64, np.float64):
            self.dtype = np.float64
        elif dtype in ('float32', 32, np.float32):
            self.dtype = np.float32
        else:
            raise ValueError(f"dtype must be 'float64' or 'float32', got {dtype!r}")
        self.max_window = int(max_window)
        if int(tile_size) not in (0, 32, 64, 128, 256):
            raise ValueError('tile_size must be 0, 32, 64, 128 or 256')
        self.tile_size = int(tile_size)
        if int(plastic_records) < 0:
            raise ValueError('plastic_records must be >= 0')
        self.plastic_records = int(plastic_records)
        if connectivity not in ('host', 'device'):
            raise ValueError(f"connectivity must be 'host' or 'device', got {connectivity!r}")
        self.connectivity = connectivity
        self._pops = []              # neuron populations, creation order
        self._devices = []           # recorders and meters, kept alive: their ids key the taps
        self._taps = {}              # id(spike_recorder) -> (pop, idx)
        self._analog_taps = {}       # id(multimeter) -> (pop, idx, names)
        self._weight_taps = {}
        self._records = []           # projections and generator channels, ordinal order
        self._drives = {}            # id(generator connection) -> (conn, start step, counts)
        self._proj_counter = itertools.count()
        self._engine = None
        self._base_step = 0
        self._rollout_ready = False
        self._edited_mid_rollout = False  # connections written while a rollout was running

    # -- creation --------------------------------------------------------
    def create(self, model_cls, size=1, *, params=None, positions=None, **kw) -> NodeView:
        if positions is not None:
            raise NotImplementedError('spatial layers are not part of the network-step engine')
        p = dict(params or {})
        p.update(kw)
        if _is_generator(model_cls):
            k = int(np.prod(size)) if isinstance(size, (tuple, list)) else int(size)
            if k > 1:
                return NodeView([_GenSegment(_GeneratorSpec(
                    model_cls, {key: _index_channel(v, i, k) for key, v in p.items()}))
                    for i in range(k)])
            return NodeView([_GenSegment(_GeneratorSpec(model_cls, p))])
        mod = model_cls(size, **p)
        if isinstance(mod, (spike_recorder, multimeter)):
            # the taps are keyed by id(device): the Simulator holds every device
            # so that no id is reused by a later one while the taps exist
            self._devices.append(mod)
            return NodeView([Selection(mod, np.arange(1))])
        if not isinstance(mod, Neuron):
            raise TypeError(f'{model_cls!r} is not a supported neuron model or device')
        self._pops.append(mod)
        self._invalidate()
        return NodeView.of(mod)

    # -- connection ------------------------------------------------------
    def connect(self, pre, post, *, rule=all_to_all, weight=None, delay=None, comm='dense',
                receptor_type=None, synapse=None, vt=None, allow_autapses=True,
                allow_multapses=True, seed=None):
        """Connect ``pre`` to ``post`` or register a recorder tap (simulator.py:606-683).

        A scalar ``delay``, or an array with one entry per pre neuron (axonal),
        is read as the reference's InputDelay reads it (delay.py:114-128):
        ``delay / dt`` in the network dtype, frames ``floor`` and ``floor + 1``
        mixed by the fraction (``comm='dense'`` and receptor inputs: graded
        parts ``w (1 - f)`` and ``w f``; ``comm='sparse'``: an event in each
        frame). ``delay`` may also be a per-edge array (one entry per realised
        edge, in the rule's edge order) or a :mod:`~brainpy.state_amd.dist`
        distribution sampled per edge, rounded to whole steps as the
        reference's per-edge delays are (``_kernel/delays.py:50-108``).
        """
        if comm not in ('dense', 'sparse'):
            raise ValueError(f"comm must be 'dense' or 'sparse', got {comm!r}")
        if len(pre.segments) == 1 and isinstance(pre.segments[0].population, multimeter):
            rec = pre.segments[0].population
            if len(post.segments) != 1:
                raise NotImplementedError('a voltmeter/multimeter records a single population segment')
            seg = post.segments[0]
            for name in rec.record_from:
                if name not in seg.population.recordables:
                    raise KeyError(f'recordable {name!r} is not available on '
                                   f'{type(seg.population).__name__}')
            self._analog_taps[id(rec)] = (seg.population, np.asarray(seg.indices), tuple(rec.record_from))
            self._invalidate()
            return None
        if len(post.segments) == 1 and isinstance(post.segments[0].population, spike_recorder):
            if len(pre.segments) != 1:
                raise NotImplementedError('recording a multi-segment view requires one recorder per segment')
            seg = pre.segments[0]
            if seg.population is None:
                raise NotImplementedError('recording a generator is not supported')
            self._taps[id(post.segments[0].population)] = (seg.population, np.asarray(seg.indices))
            self._invalidate()
            return None
        seg_weights = self._segment_weights(weight, len(pre.segments))
        projs = []
        for pre_seg, w_seg in zip(pre.segments, seg_weights):
            for post_seg in post.segments:
                pr = self._connect_pair(pre_seg, post_seg, rule, w_seg, delay, allow_autapses,
                                        allow_multapses, seed, receptor_type, synapse, comm)
                if pr is not None:
                    projs.append(pr)
        self._invalidate()
        if not projs:
            return None
        return projs[0] if len(projs) == 1 else projs

    @staticmethod
    def _segment_weights(weight, n_seg):
        if n_seg > 1 and weight is not None and np.ndim(getattr(weight, 'value', weight)) == 1 \
                and len(np.atleast_1d(getattr(weight, 'value', weight))) == n_seg:
            return [weight[i] for i in range(n_seg)]
        return [weight] * n_seg

    @staticmethod
    def _derive_seed(base, ordinal: int) -> int:
        """Simulator._derive_seed (simulator.py:926-935)."""
        b = 0 if base is None else int(base)
        return (b * 1_000_003 + ordinal + 1) & 0x7FFFFFFF

    def _receptor(self, post_pop, receptor_type):
        if isinstance(post_pop, aeif_cond_alpha):
            if receptor_type is None:
                raise ValueError('aeif_cond_alpha is a two-receptor model: connect with '
                                 'receptor_type=1 (g_ex) or 2 (g_in)')
            k = int(receptor_type)
            if not 1 <= k <= 2:
                raise ValueError(f'receptor_type {k} out of range [1, 2]')
            return k - 1
        if receptor_type is not None:
            raise NotImplementedError(f'{type(post_pop).__name__} has no receptor ports')
        return 0

    def _connect_pair(self, pre_seg, post_seg, rule, weight, delay, allow_autapses,
                      allow_multapses, seed, receptor_type, synapse, comm='dense'):
        ordinal = next(self._proj_counter)
        post_pop = post_seg.population
        if post_pop is None or not isinstance(post_pop, Neuron):
            raise TypeError('connect target must be a neuron population')
        if isinstance(pre_seg, _GenSegment) and getattr(pre_seg.spec.model_cls, '_injects_current', False):
            if synapse is not None:
                raise NotImplementedError('a current generator has no synapse')
            if receptor_type is not None:
                raise ValueError(f'receptor_type={receptor_type} was given for current injection into '
                                 f'{type(post_pop).__name__}, which has no current receptors.')
            post_idx = np.asarray(post_seg.indices, dtype=np.int32)
            gen = pre_seg.spec.model_cls(len(post_idx), **dict(pre_seg.spec.params))
            if isinstance(weight, u.Quantity):
                raise TypeError('a current generator connection weight is a plain scale factor')
            w = None if weight is None else np.asarray(weight, dtype=np.float64)
            if w is not None and np.ndim(w) != 0:
                raise NotImplementedError('current generator connections take a scalar weight')
            rec = _CurChannel(ordinal, gen, post_pop, post_idx, None if w is None else float(w))
            self._records.append(rec)
            return rec
        receptor = self._receptor(post_pop, receptor_type)
        wdim = _weight_dim(post_pop)
        post_idx = np.asarray(post_seg.indices, dtype=np.int32)
        if isinstance(pre_seg, _GenSegment):
            spec = pre_seg.spec
            if synapse is not None:
                raise NotImplementedError('plastic synapses from generators are not supported')
            params = dict(spec.params)
            params['rng_seed'] = self._derive_seed(params.get('rng_seed'), ordinal)
            gen = spec.model_cls(len(post_idx), **params)
            w = 1.0 if weight is None else u.magnitude(weight, wdim)
            if np.ndim(w) != 0:
                raise NotImplementedError('generator connections take a scalar weight')
            lo, frac = input_delay_steps(delay, self.dt, self.dtype)
            if np.ndim(lo) != 0:
                raise NotImplementedError('generator connections take a scalar delay')
            rec = _GenChannel(ordinal, gen, post_pop, post_idx, float(w), lo, receptor, params['rng_seed'])
            rec.delay_frac = float(frac)
            self._records.append(rec)
            return rec
        pre_pop = pre_seg.population
        if not isinstance(pre_pop, Neuron):
            raise TypeError('connect source must be a neuron population or a generator')
        pre_sel = np.asarray(pre_seg.indices, dtype=np.int32)
        n_pre, n_post = len(pre_sel), len(post_idx)
        dseed = self._derive_seed(seed, ordinal)
        dev = None
        if isinstance(rule, _OneToOne):
            spec = rule.sample(n_pre, n_post, seed=dseed, pre_is_post=False,
                               allow_autapses=True, allow_multapses=True)
        else:
            if not isinstance(rule, ConnRule):
                raise TypeError(f'rule must be a connection rule, got {rule!r}')
            if self.connectivity == 'device' and rule.random:
                dev = rule.sample_device(n_pre, n_post, seed=dseed, pre_is_post=(pre_pop is post_pop),
                                         allow_autapses=allow_autapses, allow_multapses=allow_multapses,
                                         pre_map=pre_sel, post_map=post_idx)
            else:
                spec = rule.sample(n_pre, n_post, seed=dseed, pre_is_post=(pre_pop is post_pop),
                                   allow_autapses=allow_autapses, allow_multapses=allow_multapses)
        if dev is None:
            pre_l = _map_local(pre_sel, spec.pre_idx)
            post_l = _map_local(post_idx, spec.post_idx)
            E = int(spec.n_edges)
        else:
            pre_l = post_l = None            # population-local already, in device memory
            E = dev.n_edges
        # weights (resolve_param, connectivity.py:385-405)
        if synapse is not None:
            if not isinstance(synapse, stdp_synapse):
                raise NotImplementedError(f'synapse model {type(synapse).__name__} is not supported')
            w0 = synapse.weight if weight is None else weight
            w = u.magnitude(w0, wdim) if not isinstance(w0, Distribution) else w0
        else:
            w = 1.0 if weight is None else weight
        if isinstance(w, Distribution):
            wv = w.sample((E,), dseed + 7919)
            w = np.asarray(u.magnitude(wv, wdim), dtype=np.float64)
        else:
            w = np.asarray(u.magnitude(w, wdim), dtype=np.float64)
        if w.ndim == 0 or w.shape == (1,):
            w = float(w.reshape(()))
        elif w.shape != (E,):
            raise ValueError(f'parameter array shape {w.shape} does not match target {(E,)}')
        # delays: per edge (a distribution, or an array with one entry per
        # edge) are whole steps by NEST's rounding (the reference's per-edge
        # path, _kernel/delays.py:50-108); homogeneous or per pre neuron
        # (axonal) delays are read as InputDelay reads them, lerping between
        # the two bracketing frames when delay/dt is not integral
        if synapse is not None and delay is None:
            delay = synapse.delay
        frac = None
        if isinstance(delay, Distribution):
            d = nest_delay_steps(delay.sample((E,), dseed + 104729), self.dt)
        else:
            dm = None if delay is None else np.asarray(u.magnitude(delay, 'time'), dtype=np.float64)
            # an array as long as the pre population is axonal, as the
            # reference's InputDelay over the pre-population vector reads it
            # (event_proj.py:196-197), also when it happens to equal E
            if dm is not None and dm.ndim == 1 and dm.shape[0] == population_size(pre_pop):
                if pre_l is None:
                    pre_l, post_l = dev.host()
                lo, fr = input_delay_steps(dm, self.dt, self.dtype)
                d, frac = lo[pre_l], fr[pre_l]       # axonal: one delay per pre neuron
            elif dm is not None and dm.ndim >= 1:
                if dm.ndim != 1 or dm.shape[0] != E:
                    raise ValueError('delay array must have one entry per pre neuron or per edge')
                d = nest_delay_steps(dm, self.dt)
            else:
                d, frac = input_delay_steps(delay, self.dt, self.dtype)
            if frac is not None and not np.any(frac != 0):
                frac = None
        graded = comm == 'dense' or isinstance(post_pop, aeif_cond_alpha)
        if frac is not None:
            steps = np.asarray(u.magnitude(delay, 'time'), dtype=np.float64) / self.dt
            if synapse is not None:
                raise NotImplementedError(
                    f'a plastic projection needs a delay of whole steps; delay / dt = {np.ravel(steps)[0]!r} '
                    'is fractional in the network dtype, where the reference lerps the delayed pre '
                    'spikes (delay.py:114-128)')
            if not graded and pre_pop.ref_count(self.dt) < 1:
                raise NotImplementedError(
                    "comm='sparse' with a fractional delay needs a presynaptic refractory period of at "
                    'least one step (an event is the OR of the two bracketing frames, delay.py:114-128 '
                    'with BinaryArray, and a neuron that can fire on consecutive steps would put a '
                    'spike in both)')
            if dev is not None:                  # the two-frame edges are built from host arrays
                if pre_l is None:
                    pre_l, post_l = dev.host()
                dev = None
        if synapse is not None:
            if pre_l is None:                # plastic projections keep host edge arrays
                pre_l, post_l = dev.host()
            pr = PlasticProjection(ordinal, pre_pop, post_pop, pre_l, post_l,
                                   np.broadcast_to(np.asarray(w, np.float64), (E,)).copy(), d,
                                   receptor, rule=synapse)
        else:
            pr = StaticProjection(ordinal, pre_pop, post_pop, pre_l, post_l, w, d, receptor,
                                  device_edges=dev)
            pr.delay_frac = frac
            pr.graded = graded
        self._records.append(pr)
        return pr

    def record_weight(self, proj):
        if not isinstance(proj, PlasticProjection):
            raise TypeError('record_weight requires a plastic projection handle from '
                            f'connect(..., synapse=spec); got {type(proj).__name__}')
        self._weight_taps[id(proj)] = proj
        return proj

    # -- descriptor ----------------------------------------------------------
    def _invalidate(self):
        if self._engine is not None:
            self._engine.close()
        self._engine = None
        self._rollout_ready = False

    def build_desc(self, *, host_edges=False):
        """Flatten the network into the C-ABI descriptor (``bs_net_desc``).
        Edges sampled on the device are passed as device pointers unless
        ``host_edges`` (a descriptor any host reader can consume)."""
        return _NetBuilder(self, host_edges=host_edges).build()

    # -- run ---------------------------------------------------------------
    def _ensure_engine(self):
        if self._engine is None:
            L.require_device()
            self._engine = _Engine(self.build_desc())
            for conn, step0, counts in self._drives.values():
                self._engine.set_drive(self._engine.desc.chan_index[id(conn)], step0, counts)
        return self._engine

    def set_poisson_drive(self, conn, counts, *, start_step=0):
        """Inject the counts of a connected poisson_generator instead of its
        random draws: ``conn`` is the handle ``connect(generator, target)``
        returned, ``counts[t][e]`` the count of generator element ``e`` (one
        per target neuron, in the target's order) at generator step
        ``start_step + t`` (the step whose update emits it; it arrives after
        the connection's delay). The generator's start/stop window still
        applies. For runs against the reference's own Poisson stream
        (``jax.random.poisson``, poisson_generator.py:501-576): feed it the
        reference's counts. ``counts=None`` removes the injection. This is a
        parity hook with no counterpart in the reference."""
        if not isinstance(conn, _GenChannel):
            raise TypeError('set_poisson_drive takes the handle of a poisson_generator connection')
        if counts is None:
            self._drives.pop(id(conn), None)
            if self._engine is not None:
                self._engine.set_drive(self._engine.desc.chan_index[id(conn)], 0, np.zeros((0, 1), np.int32))
            return
        c = np.ascontiguousarray(counts, dtype=np.int64)
        if c.ndim != 2 or c.shape[1] != len(conn.post_idx):
            raise ValueError(f'counts must be [n_steps][{len(conn.post_idx)}] (one column per target neuron)')
        if c.size and (c.min() < 0 or c.max() > 65535):
            raise ValueError('counts must be in [0, 65535]')
        if int(start_step) < 0:
            raise ValueError('start_step must be >= 0')
        c = c.astype(np.int32)
        self._drives[id(conn)] = (conn, int(start_step), c)
        if self._engine is not None:
            self._engine.set_drive(self._engine.desc.chan_index[id(conn)], int(start_step), c)

    def reset_rollout(self, *, dt=None):
        eng = self._ensure_engine()
        eng.reset()
        self._base_step = 0
        self._rollout_ready = True
        self._edited_mid_rollout = False

    def simulate(self, duration, *, dt=None) -> SimulationResult:
        self.reset_rollout()
        return self._run_window(duration)

    def cont(self, duration, *, dt=None) -> SimulationResult:
        if self._edited_mid_rollout:
            raise RuntimeError('connections were edited (SynapseCollection.set) after this rollout '
                               'started; the engine builds its delivery tables from them, so start a '
                               'new rollout with simulate() (or reset_rollout())')
        if not self._rollout_ready:
            self.reset_rollout()
        return self._run_window(duration)

    def save_state(self) -> bytes:
        """Snapshot of the running rollout: every piece of engine state the
        next step depends on (neurons, refractory counters, in-flight spikes
        and arrival rings, plastic traces and weights) and the step counter,
        which positions every random stream. ``load_state`` then ``cont``
        continues from it bit for bit. (The reference keeps this state inside
        the Simulator between ``cont`` calls, simulator.py:1822-1839; here it
        can also be stored, restored and moved to another Simulator built
        the same way.)"""
        if not self._rollout_ready:
            self.reset_rollout()
        blob = self._engine.save_state()
        return int(self._base_step).to_bytes(8, 'little', signed=True) + blob

    def load_state(self, blob: bytes):
        """Restore a snapshot of ``save_state`` (of a network built the same
        way; another network's snapshot raises EngineError)."""
        if self._edited_mid_rollout or self._engine is None:
            self._engine = None
            self._edited_mid_rollout = False
        eng = self._ensure_engine()
        eng.load_state(blob[8:])
        self._base_step = int.from_bytes(blob[:8], 'little', signed=True)
        self._rollout_ready = True

    def n_steps_of(self, duration) -> int:
        T = u.scalar(duration, 'time')
        return int(math.ceil(T / self.dt))   # len(arange(0, T, dt))

    def get_connections(self, source=None, target=None, synapse=None):
        """Realised synapses matching the filters, as a ``SynapseCollection``
        (NEST ``GetConnections``; simulator.py:832-866): ``source`` / ``target``
        are NodeViews matched by population and local index, ``synapse`` a model
        name ('static_synapse', 'stdp_synapse')."""
        from .connections import collect_connections
        return collect_connections(self, source, target, synapse)

    def _connections_changed(self):
        if self._rollout_ready:
            self._edited_mid_rollout = True
        self._invalidate()
        self._rollout_ready = False

    def get_weights(self, proj):
        """Current weights of a plastic projection, in the reference's CSR
        (stable sorted-by-pre) edge order (event_plastic.py:293-296)."""
        if not isinstance(proj, PlasticProjection):
            raise TypeError('get_weights needs a plastic projection handle')
        eng = self._ensure_engine()
        return eng.weights(eng.desc.proj_index[id(proj)], proj.n_edges)

    def _run_window(self, duration) -> SimulationResult:
        eng = self._engine
        T = u.scalar(duration, 'time')
        n = self.n_steps_of(duration)
        weights = {}
        taps = list(self._weight_taps.items())
        # weight taps read the post-update weights every step (simulator.py:1868),
        # copied on the device after each step and brought back once per chunk
        eng.set_weight_taps([eng.desc.proj_index[id(p)] for _, p in taps])
        vm, (gids, steps) = eng.run(n)
        for k, (rid, p) in enumerate(taps):
            unit = {'voltage': u.mV, 'current': u.pA, 'conductance': u.nS}[_weight_dim(p.post_pop)]
            weights[rid] = u.Quantity(eng.weight_trace(k, n, p.n_edges), unit)
        res = self._result(eng.desc, n, T, vm, gids, steps)
        res._weights = weights
        return res

    def _result(self, desc, n, T, vm, gids, steps) -> SimulationResult:
        step0 = self._base_step
        recs = {}
        for rid, (pop, idx) in self._taps.items():
            gid0 = desc.gid0[id(pop)]
            spk = np.zeros((n, len(idx)), dtype=self.dtype)
            sel = (gids >= gid0) & (gids < gid0 + pop.n)
            lg = gids[sel] - gid0
            st = steps[sel] - step0
            cols = {}
            for c, li in enumerate(idx):
                cols.setdefault(int(li), []).append(c)
            if len(cols) == len(idx):
                colmap = np.full(pop.n, -1, dtype=np.int64)
                colmap[np.asarray(idx)] = np.arange(len(idx))
                cc = colmap[lg]
                ok = cc >= 0
                spk[st[ok], cc[ok]] = 1
            else:
                for a, b in zip(lg, st):
                    for c in cols.get(int(a), ()):
                        spk[b, c] = 1
            recs[rid] = spk
        traces = {}
        for rid, (pop, idx, names) in self._analog_taps.items():
            for name in names:
                cols = desc.tap_cols[(rid, name)]
                traces[f'{rid}|{name}'] = u.Quantity(vm[:, cols], _trace_unit(pop, name))
        times = u.Quantity((step0 + np.arange(n)) * self.dt, u.ms)
        self._base_step += n
        return SimulationResult(recs, T, self.dt, traces=traces, times=times)

    def __del__(self):
        try:
            if self._engine is not None:
                self._engine.close()
        except Exception:
            pass


def _trace_unit(pop, name):
    return {'V_m': u.mV, 'g_ex': u.nS, 'g_in': u.nS, 'w': u.pA}.get(name, u.pA)


def _index_channel(val, i, k):
    v = getattr(val, 'value', val)
    if np.ndim(v) == 1 and len(v) == k:
        return val[i] if isinstance(val, u.Quantity) else v[i]
    return val


# ------------------------------------------------------------- descriptor
class _TwoFrameEdges:
    """The edges a projection with fractional InputDelay reads delivers
    through (delay.py:114-128): every edge whose delay/dt has a fraction
    delivers from frame ``lo`` and from frame ``lo + 1``. With
    ``comm='sparse'`` (BinaryArray, a binary event matmul) both frames carry
    the whole weight, as two events; with ``comm='dense'`` and receptor scatter
    (graded inputs) they carry ``w * (1 - frac)`` and ``w * frac``, each one
    product in the network dtype, summed exactly as per-edge weights."""
    plastic = False
    device_edges = None

    def __init__(self, r, tdtype):
        pre = np.asarray(r.pre_idx, np.int32)
        post = np.asarray(r.post_idx, np.int32)
        E = pre.shape[0]
        d = np.broadcast_to(np.asarray(r.delay, np.int64), (E,))
        fr = np.broadcast_to(np.asarray(r.delay_frac, tdtype), (E,))
        lerp = fr != 0
        w = np.broadcast_to(np.asarray(r.weight, np.float64), (E,))
        if r.graded:
            wt = w.astype(tdtype)
            one = tdtype(1.0)
            w_lo = np.where(lerp, (wt * (one - fr)).astype(tdtype), wt).astype(np.float64)
            w_hi = (wt * fr).astype(tdtype).astype(np.float64)[lerp]
            self.weight = np.concatenate([w_lo, w_hi])
        else:
            self.weight = float(r.weight) if r.homogeneous else np.concatenate([w, w[lerp]])
        self.pre_idx = np.concatenate([pre, pre[lerp]])
        self.post_idx = np.concatenate([post, post[lerp]])
        self.delay = np.concatenate([d, d[lerp] + 1]).astype(np.int32)
        self.receptor = r.receptor
        self.pre_pop, self.post_pop = r.pre_pop, r.post_pop

    @property
    def n_edges(self):
        return int(self.pre_idx.shape[0])

    @property
    def homogeneous(self):
        return np.ndim(self.weight) == 0


class _Desc:
    """ctypes descriptor plus the NumPy arrays it points into (kept alive)."""

    def __init__(self):
        self.keep = []
        self.c = None
        self.gid0 = {}
        self.pop_index = {}
        self.tap_cols = {}
        self.n_taps = 0
        self.proj_index = {}
        self.chan_index = {}             # id(generator connection) -> Poisson channel
        self.device_edges = False        # some projection passes device pointers

    def arr(self, a, ctype):
        a = np.ascontiguousarray(a)
        self.keep.append(a)
        return a.ctypes.data_as(C.POINTER(ctype))

    def dptr(self, t, ctype):
        """A device tensor's address as a ctypes pointer (the tensor is kept)."""
        self.keep.append(t)
        return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(ctype))


class _NetBuilder:
    """``remote``: indices of populations whose neurons live on another shard
    (no channels, taps or updates here; their spikes are injected), and then
    every local population keeps its spikes (they are the exchange)."""

    def __init__(self, sim: Simulator, host_edges=False, remote=()):
        self.sim = sim
        self.host_edges = host_edges
        self.remote = frozenset(remote)

    def build(self) -> _Desc:
        sim = self.sim
        dt = sim.dt
        D = _Desc()
        pops = sim._pops
        npop = len(pops)
        if npop == 0:
            raise ValueError('the network has no neuron population')
        for p, pop in enumerate(pops):
            D.pop_index[id(pop)] = p
        gid = 0
        for pop in pops:
            D.gid0[id(pop)] = gid
            gid += pop.n
        # channels per post population, in registration (ordinal) order
        per_pop = [[] for _ in range(npop)]
        for r in sim._records:
            q = D.pop_index[id(r.post_pop)]
            if q in self.remote:              # a remote population's inputs are its own shard's
                continue
            if isinstance(r, _CurChannel):
                pieces = r.gen.plateaus(dt)
                if isinstance(r.gen, step_current_generator):
                    if pieces:                # the whole schedule in one channel
                        per_pop[q].append((r, ('sched', pieces)))
                else:                         # dc / ac: one current channel
                    per_pop[q].extend((r, piece) for piece in pieces)
            else:
                per_pop[q].append(r)
        chans = (L.Channel * max(1, sum(len(x) for x in per_pop)))()
        cpops = (L.Population * npop)()
        tables, table_off = [], []
        projs = []
        c = 0
        tdtype = np.float64 if sim.dtype == np.float64 else np.float32
        # voltmeter taps: per population, concatenated in registration order
        vm_lists = [[] for _ in range(npop)]
        tap_order = []
        for rid, (pop, idx, names) in sim._analog_taps.items():
            if D.pop_index[id(pop)] in self.remote:
                continue
            for name in names:
                if pop.recordables[name] != 0:
                    raise NotImplementedError('only V_m is recorded by the engine')
                p = D.pop_index[id(pop)]
                start = len(vm_lists[p])
                vm_lists[p].extend(int(i) for i in idx)
                tap_order.append((rid, name, p, start, len(idx)))
        # global tap column = concatenation over populations
        pop_tap_base = np.cumsum([0] + [len(v) for v in vm_lists])
        for rid, name, p, start, cnt in tap_order:
            D.tap_cols[(rid, name)] = np.arange(pop_tap_base[p] + start, pop_tap_base[p] + start + cnt)
        D.n_taps = int(pop_tap_base[-1])
        rec_pops = {D.pop_index[id(pop)] for (pop, _) in sim._taps.values()}
        if self.remote:
            rec_pops = set(range(npop)) - set(self.remote)
        off = 0
        for p, pop in enumerate(pops):
            P = cpops[p]
            P.model = pop.model_id
            P.n = pop.n
            blk = pop.param_block(dt)
            for k in range(L.BS_NPARAM):
                P.params[k] = float(blk[k])
            P.state0 = D.arr(pop.state0(dt).astype(np.float64), C.c_double)
            P.ch_begin = c
            P.ch_count = len(per_pop[p])
            P.n_vm = len(vm_lists[p])
            P.vm_idx = D.arr(np.asarray(vm_lists[p] or [0], dtype=np.int32), C.c_int32)
            P.record_spikes = 1 if p in rec_pops else 0
            P.remote = 1 if p in self.remote else 0
            for r in per_pop[p]:
                ch = chans[c]
                piece = None
                if isinstance(r, tuple):
                    r, piece = r
                ch.receptor = r.receptor
                if piece is not None and piece[0] == 'sched':
                    # step_current_generator: its plateaus as one schedule,
                    # each value device.update() * weight (simulator.py:1738-1740)
                    pcs = piece[1]
                    ch.kind = L.BS_CH_SCHEDULE
                    ch.weight = 1.0
                    se = np.array([[n0, min(n1, np.iinfo(np.int64).max)] for _, n0, n1 in pcs], dtype=np.int64)
                    sv = np.array([a if r.weight is None else a * r.weight for a, _, _ in pcs], dtype=np.float64)
                    ch.sched = D.arr(se.reshape(-1), C.c_int64)
                    ch.sched_val = D.arr(sv, C.c_double)
                    ch.sched_len = len(pcs)
                    ch.t_min_step = int(se[0, 0])
                    ch.t_max_step = int(se[-1, 1])
                    if len(r.post_idx) == pop.n and np.array_equal(r.post_idx, np.arange(pop.n)):
                        ch.gen_index = None
                    else:
                        gi = np.full(pop.n, -1, dtype=np.int32)
                        gi[r.post_idx] = np.arange(len(r.post_idx), dtype=np.int32)
                        ch.gen_index = D.arr(gi, C.c_int32)
                    table_off.append(-1)
                elif piece is not None:
                    amp, n0, n1 = piece
                    if amp == 'ac':   # the sinusoid, evaluated on the device
                        ch.kind = L.BS_CH_AC
                        ch.weight = 1.0 if r.weight is None else r.weight
                        for k, v in enumerate(r.gen.ac_params()):
                            ch.ac[k] = v
                    else:
                        ch.kind = L.BS_CH_CURRENT
                        # device.update() * weight (simulator.py:1738-1740), one product
                        ch.weight = amp if r.weight is None else amp * r.weight
                    ch.t_min_step = int(n0)
                    ch.t_max_step = int(min(n1, np.iinfo(np.int64).max))
                    if len(r.post_idx) == pop.n and np.array_equal(r.post_idx, np.arange(pop.n)):
                        ch.gen_index = None
                    else:
                        gi = np.full(pop.n, -1, dtype=np.int32)
                        gi[r.post_idx] = np.arange(len(r.post_idx), dtype=np.int32)
                        ch.gen_index = D.arr(gi, C.c_int32)
                    table_off.append(-1)
                elif isinstance(r, _GenChannel):
                    g = r.gen
                    D.chan_index[id(r)] = c
                    ch.kind = L.BS_CH_POISSON
                    ch.weight = r.weight
                    ch.lam = g.lam(dt)
                    tmin, tmax = g.step_window(dt)
                    ch.t_min_step = int(tmin)
                    ch.t_max_step = int(min(tmax, np.iinfo(np.int64).max))
                    ch.seed = int(r.seed) & 0xFFFFFFFF
                    ch.delay_steps = int(r.delay)
                    ch.delay_frac = float(r.delay_frac)
                    if len(r.post_idx) == pop.n and np.array_equal(r.post_idx, np.arange(pop.n)):
                        ch.gen_index = None
                    else:
                        gi = np.full(pop.n, -1, dtype=np.int32)
                        gi[r.post_idx] = np.arange(len(r.post_idx), dtype=np.int32)
                        ch.gen_index = D.arr(gi, C.c_int32)
                    table_off.append(-1)
                else:
                    r_id = id(r)
                    if getattr(r, 'delay_frac', None) is not None:
                        r = _TwoFrameEdges(r, tdtype)
                    pre_p = D.pop_index[id(r.pre_pop)]
                    E = r.n_edges
                    dev = None if (self.host_edges or r.plastic) else getattr(r, 'device_edges', None)
                    if r.homogeneous and not r.plastic:
                        ch.kind = L.BS_CH_COUNT
                        ch.weight = float(r.weight)
                        if not E:
                            kmax = 0
                        elif dev is not None:
                            kmax = dev.max_indegree
                        else:
                            kmax = int(np.bincount(r.post_idx, minlength=pop.n).max())
                        if kmax > 65535:
                            raise NotImplementedError('more than 65535 edges into one neuron '
                                                      'from one projection')
                        ch.max_count = kmax
                        t = seq_sum_table(r.weight, kmax, tdtype)
                        table_off.append(off)
                        tables.append(t)
                        off += len(t)
                    else:
                        ch.kind = L.BS_CH_FIXED
                        wabs = np.abs(np.asarray(r.weight, np.float64))
                        if r.plastic:
                            wabs = np.maximum(wabs, abs(r.rule.Wmax))
                        m = float(np.bincount(r.post_idx, weights=wabs, minlength=pop.n).max()) if E else 0.0
                        ch.fixed_shift = int(min(1000, max(0, 61 - math.ceil(math.log2(m)) - 1))) if m > 0 else 0
                        table_off.append(-1)
                    pj = L.Projection()
                    pj.pre_pop = pre_p
                    pj.post_pop = p
                    pj.channel = c
                    pj.plastic = 1 if r.plastic else 0
                    pj.n_edges = E
                    if dev is not None:
                        # edges stay in device memory; per-edge delays and
                        # weights follow them there (bs_projection.on_device)
                        import torch
                        pj.on_device = 1
                        pj.pre = D.dptr(dev.pre, C.c_int32)
                        pj.post = D.dptr(dev.post, C.c_int32)
                        up = lambda a, dt_: torch.from_numpy(np.ascontiguousarray(a, dtype=dt_)).cuda()  # noqa: E731
                        if np.ndim(r.delay) == 0:
                            pj.delay = None
                            pj.delay_steps = int(r.delay)
                        else:
                            pj.delay = D.dptr(up(r.delay, np.int32), C.c_int32)
                            pj.delay_steps = 0
                        pj.weight = None if np.ndim(r.weight) == 0 else D.dptr(up(r.weight, np.float64), C.c_double)
                        D.device_edges = True
                    else:
                        pj.pre = D.arr(np.asarray(r.pre_idx, dtype=np.int32), C.c_int32)
                        pj.post = D.arr(np.asarray(r.post_idx, dtype=np.int32), C.c_int32)
                        if np.ndim(r.delay) == 0:
                            pj.delay = None
                            pj.delay_steps = int(r.delay)
                        else:
                            pj.delay = D.arr(np.asarray(r.delay, np.int32), C.c_int32)
                            pj.delay_steps = 0
                        if np.ndim(r.weight) != 0:
                            pj.weight = D.arr(np.asarray(r.weight, np.float64), C.c_double)
                        else:
                            pj.weight = None
                    if r.plastic:
                        s = r.rule
                        pj.stdp.tau_plus, pj.stdp.tau_minus = s.tau_plus, s.tau_minus
                        pj.stdp.lambda_, pj.stdp.alpha = s.lambda_, s.alpha
                        pj.stdp.mu_plus, pj.stdp.mu_minus, pj.stdp.Wmax = s.mu_plus, s.mu_minus, s.Wmax
                        pj.stdp.decay_plus = float(np.exp(-dt / s.tau_plus))
                        pj.stdp.decay_minus = float(np.exp(-dt / s.tau_minus))
                    D.proj_index[r_id] = len(projs)
                    projs.append(pj)
                    D.keep.append(pj)
                c += 1
        cprojs = (L.Projection * max(1, len(projs)))(*projs) if projs else (L.Projection * 1)()
        tab = np.concatenate(tables) if tables else np.zeros(1)
        d = L.NetDesc()
        d.abi_version = L.BS_ABI_VERSION
        d.dtype = L.BS_F64 if sim.dtype == np.float64 else L.BS_F32
        d.dt_ms = dt
        d.n_pops = npop
        d.n_channels = c
        d.n_projs = len(projs)
        d.max_window = sim.max_window
        d.tile_size = sim.tile_size
        d.plastic_records = getattr(sim, 'plastic_records', 0)
        d.pops = cpops
        d.channels = chans
        d.projs = cprojs
        d.tables = D.arr(tab, C.c_double)
        d.table_off = D.arr(np.asarray(table_off or [-1], dtype=np.int64), C.c_int64)
        d.n_tables = len(tab)
        D.keep += [cpops, chans, cprojs]
        D.tables_np = tab
        D.table_off_np = np.asarray(table_off or [-1], dtype=np.int64)
        D.c = d
        D.ntot = gid
        if D.device_edges:
            import torch
            torch.cuda.synchronize()     # the uploads above are complete before the engine reads them
        return D


class _Engine:
    """Owns one ``bs_net`` handle."""

    def __init__(self, desc: _Desc):
        self.desc = desc
        self.lib = L.lib()
        h = C.c_void_p()
        L.check(self.lib.bs_net_create(C.byref(desc.c), C.byref(h)))
        self.h = h

    def reset(self):
        L.check(self.lib.bs_net_reset(self.h))

    def set_timing(self, on):
        L.check(self.lib.bs_net_set_timing(self.h, 1 if on else 0))

    def run(self, n_steps):
        nt = self.desc.n_taps
        vm = np.zeros((n_steps, max(nt, 0)), dtype=np.float64)
        ptr = vm.ctypes.data_as(C.POINTER(C.c_double)) if nt > 0 else None
        L.check(self.lib.bs_net_run(self.h, int(n_steps), ptr))
        cnt = C.c_int64()
        L.check(self.lib.bs_net_spike_count(self.h, C.byref(cnt)))
        gid = np.zeros(max(1, cnt.value), dtype=np.int32)
        st = np.zeros(max(1, cnt.value), dtype=np.int32)
        L.check(self.lib.bs_net_copy_spikes(self.h, gid.ctypes.data_as(C.POINTER(C.c_int32)),
                                            st.ctypes.data_as(C.POINTER(C.c_int32)), len(gid)))
        return vm, (gid[:cnt.value], st[:cnt.value])

    def window(self) -> int:
        """Steps per window launch."""
        w = C.c_int32()
        L.check(self.lib.bs_net_window(self.h, C.byref(w)))
        return int(w.value)

    def inject(self, gid, step):
        """A peer shard's spikes of the window just run (global gids)."""
        g = np.ascontiguousarray(gid, dtype=np.int32)
        st = np.ascontiguousarray(step, dtype=np.int32)
        L.check(self.lib.bs_net_inject_spikes(self.h, len(g), g.ctypes.data_as(C.POINTER(C.c_int32)),
                                              st.ctypes.data_as(C.POINTER(C.c_int32))))

    def timing(self):
        vals = [C.c_double() for _ in range(3)] + [C.c_int64() for _ in range(2)]
        L.check(self.lib.bs_net_last_timing(self.h, *[C.byref(v) for v in vals]))
        return dict(deliver_ms=vals[0].value, update_ms=vals[1].value, total_ms=vals[2].value,
                    n_deliver=vals[3].value, n_update=vals[4].value)

    def bytes(self):
        a, b, n = C.c_double(), C.c_double(), C.c_int64()
        L.check(self.lib.bs_net_last_bytes(self.h, C.byref(a), C.byref(b), C.byref(n)))
        return dict(deliver_bytes=a.value, update_bytes=b.value, n_events=n.value)

    def set_drive(self, channel, step0, counts):
        c = np.ascontiguousarray(counts, dtype=np.int32)
        L.check(self.lib.bs_net_set_drive(self.h, int(channel), int(step0), c.shape[0], c.shape[1],
                                          c.ctypes.data_as(C.POINTER(C.c_int32))))

    def set_weight_taps(self, proj_indices):
        arr = np.ascontiguousarray(proj_indices, dtype=np.int32)
        L.check(self.lib.bs_net_set_weight_taps(self.h, len(arr), arr.ctypes.data_as(C.POINTER(C.c_int32))))

    def weight_trace(self, k, n_steps, n_edges):
        out = np.zeros((n_steps, n_edges), dtype=np.float64)
        if out.size:
            L.check(self.lib.bs_net_copy_weight_trace(self.h, k, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def set_state(self, pop_index, var, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        L.check(self.lib.bs_net_set_state(self.h, pop_index, var, v.ctypes.data_as(C.POINTER(C.c_double))))

    def save_state(self) -> bytes:
        """Snapshot of the whole engine state (bs_net_save_state)."""
        n = C.c_int64()
        L.check(self.lib.bs_net_state_bytes(self.h, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        L.check(self.lib.bs_net_save_state(self.h, buf, n.value))
        return buf.raw

    def load_state(self, blob: bytes):
        buf = C.create_string_buffer(bytes(blob), len(blob))
        L.check(self.lib.bs_net_load_state(self.h, buf, len(blob)))

    def run_device(self, n_steps, stream=None):
        """Queue ``n_steps`` on a torch/HIP stream (the current one by
        default) with device outputs: ``(raster, vm, status)`` torch tensors,
        raster uint32 bits [n_steps][ceil(N / 32)] of every neuron, vm f64
        [n_steps][taps], status int32 [1] (0 = OK). Nothing is synchronised."""
        import torch
        st = torch.cuda.current_stream() if stream is None else stream
        words = (int(self.desc.ntot) + 31) // 32
        with torch.cuda.stream(st):
            raster = torch.empty((n_steps, words), dtype=torch.int32, device='cuda')
            vm = torch.empty((n_steps, max(self.desc.n_taps, 0)), dtype=torch.float64, device='cuda')
            status = torch.empty(1, dtype=torch.int32, device='cuda')
        vm_ptr = C.c_void_p(vm.data_ptr()) if self.desc.n_taps > 0 and n_steps > 0 else None
        L.check(self.lib.bs_net_run_device(self.h, int(n_steps), C.c_void_p(st.cuda_stream),
                                           C.c_void_p(raster.data_ptr()) if n_steps > 0 else None, vm_ptr,
                                           C.c_void_p(status.data_ptr())))
        return raster, vm, status

    def state(self, pop_index, var, n):
        out = np.zeros(max(1, n))
        L.check(self.lib.bs_net_get_state(self.h, pop_index, var,
                                          out.ctypes.data_as(C.POINTER(C.c_double))))
        return out[:n]

    def weights(self, proj_index, n):
        out = np.zeros(max(1, n))
        L.check(self.lib.bs_net_get_weights(self.h, proj_index,
                                            out.ctypes.data_as(C.POINTER(C.c_double))))
        return out[:n]

    def close(self):
        if getattr(self, 'h', None) is not None and self.h.value:
            self.lib.bs_net_destroy(self.h)
            self.h = C.c_void_p()
