"""Regenerate the golden fixtures in this folder from the CPU oracle.

    python tests/golden/make_golden.py

Fixtures (inputs are the seeded networks named below; outputs from the C++
oracle, which is pinned by the known-answer tests in tests/test_oracle_kat.py):

* brunel1000_delta_f64.npz -- Brunel-1000 iaf_psc_delta (networks.brunel_delta,
  order 200, delay 1.5 ms, f64): spike list (gid, step) of the first 2000 steps
  and the V_m of neurons 0..9.
* brunel_exp_mixed_f32.npz -- a 300 + 75 neuron iaf_psc_exp network with
  Poisson drive, f32 (networks.brunel_exp(375, seed=3)): spikes of 1500 steps.
* threefry_kat.json -- Random123 Threefry-2x32-20 published known answers.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from brainpy.state_amd.networks import brunel_delta, brunel_exp  # noqa: E402


def main():
    sim, h = brunel_delta(order=200, n_vm=10, dtype='float64')
    vm, g, s = oracle.OracleNet(sim.build_desc()).run(2000)
    np.savez_compressed(os.path.join(HERE, 'brunel1000_delta_f64.npz'), gid=g.astype(np.int32),
                        step=s.astype(np.int32), vm=vm)
    sim, h = brunel_exp(375, seed=3)
    vm, g, s = oracle.OracleNet(sim.build_desc()).run(1500)
    np.savez_compressed(os.path.join(HERE, 'brunel_exp_mixed_f32.npz'), gid=g.astype(np.int32),
                        step=s.astype(np.int32))
    kat = [  # key0, key1, ctr0, ctr1 -> out0, out1 (Random123 kat_vectors, threefry2x32_20)
        [0x00000000, 0x00000000, 0x00000000, 0x00000000, 0x6b200159, 0x99ba4efe],
        [0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff, 0x1cb996fc, 0xbb002be7],
        [0x13198a2e, 0x03707344, 0x243f6a88, 0x85a308d3, 0xc4923a9c, 0x483df7a0],
    ]
    with open(os.path.join(HERE, 'threefry_kat.json'), 'w') as f:
        json.dump({'threefry2x32_20': kat}, f, indent=1)
    print('wrote fixtures to', HERE)


if __name__ == '__main__':
    main()
