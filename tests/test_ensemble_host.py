"""CPU checks of the ensemble descriptor merge (no device needed)."""
import numpy as np

from brainpy.state_amd.ensemble import merge_descs
from brainpy.state_amd.networks import brunel_exp, brunel_exp_ensemble


def test_merge_offsets_populations_channels_projections_tables():
    a, _ = brunel_exp(300, seed=1)
    b, _ = brunel_exp(150, seed=3, g=4.0)
    da, db = a.build_desc(), b.build_desc()
    m = merge_descs([da, db], tile_size=128, max_window=8)
    c = m.c
    assert c.n_pops == da.c.n_pops + db.c.n_pops == 4
    assert c.n_channels == da.c.n_channels + db.c.n_channels
    assert c.n_projs == da.c.n_projs + db.c.n_projs
    assert c.tile_size == 128 and c.max_window == 8
    # second network: indices shifted by the first one's counts
    pa = da.c.n_pops
    assert c.pops[pa].ch_begin == db.c.pops[0].ch_begin + da.c.n_channels
    r = da.c.n_projs
    assert c.projs[r].pre_pop == db.c.projs[0].pre_pop + pa
    assert c.projs[r].post_pop == db.c.projs[0].post_pop + pa
    assert c.projs[r].channel == db.c.projs[0].channel + da.c.n_channels
    # pointers still reach the same edge arrays
    assert c.projs[r].pre[5] == db.c.projs[0].pre[5]
    # tables: offsets of the second network moved past the first's tables
    for k in range(db.c.n_channels):
        o = db.table_off_np[k]
        mo = m.table_off_np[da.c.n_channels + k]
        if o >= 0:
            assert mo == o + len(da.tables_np)
            assert m.tables_np[mo] == db.tables_np[o]
        else:
            assert mo == -1
    assert m.ntot == da.ntot + db.ntot
    assert [i['gid0'] for i in m.instances] == [0, da.ntot]


def test_ensemble_builder_draws_uneven_networks():
    sims, p = brunel_exp_ensemble(5, n_min=200, n_max=800, seed=2)
    assert len(sims) == 5
    assert np.all((p['N'] >= 200) & (p['N'] <= 800)) and len(set(p['N'].tolist())) == 5
    assert np.all((p['g'] >= 3) & (p['g'] <= 7)) and np.all((p['eta'] >= 0.9) & (p['eta'] <= 4))
    assert sims[0].dtype == np.float32
