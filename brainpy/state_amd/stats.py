"""Spike-train statistics for comparing runs (float32 against float64 mode).

The definitions follow the reference's validation statistics
(``_kernel/validation/stats.py:37-60``): the mean rate of ``n`` neurons over a
duration, and the mean over neurons with at least ``min_spikes`` spikes of the
coefficient of variation of their inter-spike intervals. The comparison bands
are the reference's distributional category (``tolerance_conventions.py``: 5 %
on the mean rate and ISI CV over >= 4 s).
"""
from __future__ import annotations

import numpy as np

__all__ = ['rate', 'cv_isi', 'first_divergence', 'compare_runs']


def rate(steps, n, n_steps, dt_ms):
    """Mean rate in Hz of ``n`` neurons with ``len(steps)`` spikes in ``n_steps`` steps."""
    if n <= 0 or n_steps <= 0:
        raise ValueError('n and n_steps must be positive')
    return float(len(steps)) / n / (n_steps * dt_ms * 1e-3)


def cv_isi(gids, steps, n, min_spikes=3):
    """Mean ISI coefficient of variation over neurons with >= ``min_spikes`` spikes."""
    g = np.asarray(gids, np.int64)
    t = np.asarray(steps, np.float64)
    order = np.lexsort((t, g))
    g, t = g[order], t[order]
    lo = np.searchsorted(g, np.arange(n))
    hi = np.searchsorted(g, np.arange(n), side='right')
    cvs = []
    for a, b in zip(lo, hi):
        if b - a >= min_spikes:
            isi = np.diff(t[a:b])
            m = isi.mean()
            if m > 0:
                cvs.append(isi.std() / m)
    return float(np.mean(cvs)) if cvs else float('nan')


def first_divergence(g1, s1, g2, s2):
    """First step at which two spike lists (sorted by (step, gid)) differ, or None."""
    n = min(len(g1), len(g2))
    a = np.asarray(s1[:n], np.int64) << 32 | np.asarray(g1[:n], np.int64)
    b = np.asarray(s2[:n], np.int64) << 32 | np.asarray(g2[:n], np.int64)
    d = np.nonzero(a != b)[0]
    if len(d):
        return int(min(s1[d[0]], s2[d[0]]))
    if len(g1) != len(g2):
        return int(s1[n]) if len(g1) > n else int(s2[n])
    return None


def compare_runs(ref, test, n, n_steps, dt_ms):
    """Report of a test run against a reference run: each is
    ``(gids, steps, vm)`` with ``vm`` the [n_steps, taps] voltmeter samples."""
    g1, s1, v1 = ref
    g2, s2, v2 = test
    div = first_divergence(g1, s1, g2, s2)
    upto = n_steps if div is None else div
    err = np.abs(np.asarray(v2[:upto], np.float64) - np.asarray(v1[:upto], np.float64))
    r1, r2 = rate(s1, n, n_steps, dt_ms), rate(s2, n, n_steps, dt_ms)
    c1, c2 = cv_isi(g1, s1, n), cv_isi(g2, s2, n)
    return {'first_divergence_step': div,
            'vm_max_abs_err_mV_before_divergence': float(err.max()) if err.size else 0.0,
            'rate_ref_hz': r1, 'rate_hz': r2, 'rate_rel_dev': abs(r2 - r1) / r1 if r1 else float('nan'),
            'cv_isi_ref': c1, 'cv_isi': c2, 'cv_isi_rel_dev': abs(c2 - c1) / c1 if c1 else float('nan')}
