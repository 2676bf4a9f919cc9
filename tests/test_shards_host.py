"""CPU checks of the shard descriptors: each shard keeps the channels, taps and
updates of its own populations and holds the others as remote populations
whose spikes it receives; the oracle refuses remote populations."""
import numpy as np
import pytest

import oracle
from brainpy.state_amd import units as u
from brainpy.state_amd.networks import brunel_delta
from brainpy.state_amd.simulator import _NetBuilder


def test_shard_descriptors_split_channels_and_taps():
    sim, h = brunel_delta(order=20, n_vm=3)
    full = sim.build_desc()
    d0 = _NetBuilder(sim, remote=[1]).build()       # E local, I remote
    d1 = _NetBuilder(sim, remote=[0]).build()
    P0, P1 = d0.c.pops, d1.c.pops
    assert (P0[0].remote, P0[1].remote) == (0, 1) and (P1[0].remote, P1[1].remote) == (1, 0)
    # the channels of a population live on its own shard only, in the same order
    assert P0[1].ch_count == 0 and P1[0].ch_count == 0
    assert P0[0].ch_count == full.c.pops[0].ch_count and P1[1].ch_count == full.c.pops[1].ch_count
    kinds_full = [full.c.channels[full.c.pops[0].ch_begin + k].kind for k in range(full.c.pops[0].ch_count)]
    kinds_0 = [d0.c.channels[P0[0].ch_begin + k].kind for k in range(P0[0].ch_count)]
    assert kinds_full == kinds_0
    # projections: into local populations only, from local and remote ones
    posts0 = {d0.c.projs[q].post_pop for q in range(d0.c.n_projs)}
    pres0 = {d0.c.projs[q].pre_pop for q in range(d0.c.n_projs)}
    assert posts0 == {0} and pres0 == {0, 1}
    # every local population keeps its spikes (the exchange); taps stay with their owner
    assert P0[0].record_spikes == 1 and P0[1].record_spikes == 0
    assert d0.n_taps == 3 and d1.n_taps == 0
    with pytest.raises(RuntimeError, match='remote'):
        oracle.OracleNet(d0)


def test_shards_reject_bad_groups():
    from brainpy.state_amd import Shards
    sim, h = brunel_delta(order=10, n_vm=0)
    with pytest.raises(ValueError, match='whole populations'):
        Shards(sim, [[h['ne'][:5]], [h['ni']]])
    with pytest.raises(ValueError, match='one shard'):
        Shards(sim, [[h['ne']]])
    with pytest.raises(ValueError, match='two shards'):
        Shards(sim, [[h['ne']], [h['ne'], h['ni']]])
