"""NodeView: NEST NodeCollection-style views over (population, local indices).

Same algebra as the reference's ``_nest_network/nodeview.py:20-56``: ``a + b``
concatenates segments, ``a[sl]`` slices a single-segment view.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ['NodeView']


def _flat_size(module) -> int:
    sz = getattr(module, 'in_size', None)
    if sz is None:
        sz = module.varshape
    if isinstance(sz, tuple):
        n = 1
        for s in sz:
            n *= int(s)
        return n
    return int(sz)


@dataclass(frozen=True, eq=False)
class _Segment:
    """A reference into one population: ``(population, 1-D local indices)``."""
    population: object
    indices: np.ndarray


class NodeView:
    """A view over one or more slices of populations / devices."""
    __module__ = 'brainpy.state_amd'

    def __init__(self, segments):
        self._segments = list(segments)

    @classmethod
    def of(cls, population) -> 'NodeView':
        return cls([_Segment(population, np.arange(_flat_size(population)))])

    @property
    def segments(self):
        return self._segments

    @property
    def size(self) -> int:
        return int(sum(int(s.indices.shape[0]) for s in self._segments))

    def __add__(self, other):
        if not isinstance(other, NodeView):
            return NotImplemented
        return NodeView(self._segments + other._segments)

    def __getitem__(self, item) -> 'NodeView':
        if len(self._segments) != 1:
            raise NotImplementedError('slicing is supported on single-segment views only')
        seg = self._segments[0]
        idx = np.asarray(seg.indices)[item]
        idx = np.atleast_1d(idx)
        return NodeView([_Segment(seg.population, idx)])

    def __len__(self) -> int:
        return self.size
