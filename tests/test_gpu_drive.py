"""GPU parity of injected Poisson drive (bs_net_set_drive): the engine and the
oracle fed the same counts give the same raster and V_m, on channels drawn
ahead by producer blocks (each population's first Poisson channel), a channel
drawn in the inputs phase and a channel with a fractional delay (two frames);
removing the injection returns to the channels' own streams."""
import numpy as np
import pytest

import oracle
from test_drive import _net

pytestmark = pytest.mark.gpu


def _compare(sim, desc, drives, n_steps):
    eng = sim._ensure_engine()
    eng.reset()
    vm_g, (g_g, s_g) = eng.run(n_steps)
    o = oracle.OracleNet(desc)
    for ch, step0, counts in drives:
        o.set_drive(ch, step0, counts)
    vm_o, g_o, s_o = o.run(n_steps)
    assert len(g_g) == len(g_o), f'spike count differs: gpu {len(g_g)} oracle {len(g_o)}'
    assert np.array_equal(g_g, g_o) and np.array_equal(s_g, s_o)
    np.testing.assert_allclose(vm_g, vm_o, rtol=1e-12, atol=0)
    return g_o


def test_injected_drive_equals_the_oracle():
    sim, conns = _net()
    rng = np.random.default_rng(7)
    plan = [(conns[0], 30, rng.poisson(2.5, (300, 80))), (conns[1], 0, rng.poisson(1.0, (500, 20))),
            (conns[2], 90, rng.poisson(4.0, (200, 40))), (conns[3], 10, rng.poisson(0.5, (400, 30)))]
    for conn, a, c in plan:
        sim.set_poisson_drive(conn, c, start_step=a)
    desc = sim.build_desc()
    drives = [(desc.chan_index[id(conn)], a, np.asarray(c, np.int32)) for conn, a, c in plan]
    g_inj = _compare(sim, desc, drives, 600)
    # without the injection: the channels' own streams again
    for conn, _, _ in plan:
        sim.set_poisson_drive(conn, None)
    g_own = _compare(sim, desc, [], 600)
    assert len(g_inj) != len(g_own)
