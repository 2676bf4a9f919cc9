"""NodeView: ordered selections of neurons or devices across populations.

A view is a sequence of selections. Each selection names one population of
this package (or a device) and the local indices taken from it, in order; the
view's neurons are the selections' neurons one after the other. ``connect``
realises one projection per (pre selection, post selection) pair, so the
selections are what the Simulator reads (``view.segments``).

The public algebra is the reference's NodeCollection-style interface
(``_nest_network/nodeview.py:20-56``): ``a + b`` concatenates, ``view[i]``
selects, ``len(view)`` / ``view.size`` count. Indexing here works on the
concatenated positions of any view (integers, slices, boolean masks and index
arrays, as NumPy reads them), so a multi-population view can be indexed too;
the result keeps the selections' populations.
"""
from __future__ import annotations

import numpy as np

__all__ = ['NodeView', 'Selection', 'population_size']


def population_size(population) -> int:
    """Number of neurons (or device channels) of a population spec: every
    model and device of this package stores it as the integer ``in_size``."""
    return int(population.in_size)


class Selection:
    """One population and the 1-D int64 array of its local indices."""

    def __init__(self, population, indices):
        self.population = population
        self.indices = np.asarray(indices, dtype=np.int64).reshape(-1)

    def __len__(self) -> int:
        return int(self.indices.shape[0])

    def __repr__(self) -> str:
        name = type(self.population).__name__
        return f'Selection({name}, {len(self)} indices)'


class NodeView:
    """An ordered view over selections of one or more populations / devices."""
    __module__ = 'brainpy.state_amd'

    def __init__(self, segments=()):
        self._parts = tuple(segments)

    @classmethod
    def of(cls, population) -> 'NodeView':
        """The whole population, in index order."""
        return cls((Selection(population, np.arange(population_size(population))),))

    @property
    def segments(self):
        """The selections, in view order (a new list each call)."""
        return list(self._parts)

    @property
    def size(self) -> int:
        return sum(len(p) for p in self._parts)

    def __len__(self) -> int:
        return self.size

    def __add__(self, other):
        if not isinstance(other, NodeView):
            return NotImplemented
        return NodeView(self._parts + other._parts)

    def __getitem__(self, item) -> 'NodeView':
        lens = np.array([len(p) for p in self._parts], dtype=np.int64)
        ends = np.cumsum(lens)
        pos = np.atleast_1d(np.arange(int(ends[-1]) if len(ends) else 0)[item])
        if pos.size == 0:
            if not self._parts:
                return NodeView()
            return NodeView((Selection(self._parts[0].population, pos),))
        part = np.searchsorted(ends, pos, side='right')
        starts = ends - lens
        # consecutive positions of one selection stay one selection
        cut = np.flatnonzero(np.diff(part)) + 1
        out = []
        for run in np.split(np.arange(pos.size), cut):
            k = int(part[run[0]])
            src = self._parts[k]
            out.append(Selection(src.population, src.indices[pos[run] - starts[k]]))
        return NodeView(out)

    def __repr__(self) -> str:
        return f'NodeView({len(self._parts)} selections, {self.size} nodes)'
