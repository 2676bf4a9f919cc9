"""CPU check of the compiled gfx950 kernels: no kernel of the engine uses
private (scratch) memory. A runtime index into a register array or struct
silently moves it to scratch, which costs a read and a write through memory
per thread and launch (on the window kernel that was ~4 MB of extra writes per
window); the compiler's resource-usage remarks catch it without a GPU."""
import os
import shutil
from concurrent.futures import ThreadPoolExecutor

import pytest

from brainpy.state_amd import build as B

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')


@pytest.mark.skipif(not (os.path.exists(HIPCC) or shutil.which('hipcc')), reason='hipcc not installed')
def test_no_kernel_uses_scratch():
    with ThreadPoolExecutor(len(B.SRC)) as ex:
        usage = {}
        for r in ex.map(B.resource_usage, B.SRC):
            usage.update(r)
    names = ' '.join(usage)
    for k in ('k_window', 'k_pois_fill', 'k_bptt_fwd', 'k_bptt_bwd',
              'k_bptt_wgrad'):
        assert k in names, k
    bad = {k: v['ScratchSize'] for k, v in usage.items() if v.get('ScratchSize', 0) != 0}
    assert not bad, f'kernels with scratch memory: {bad}'
    # the window kernel keeps two waves per SIMD (VGPR budget 256), and its
    # float instantiations for static networks six (the ensemble's
    # latency-bound blocks lose ~10 % per wave fewer)
    for k, v in usage.items():
        if 'k_window' in k:
            assert v['Occupancy'] >= 2, (k, v)
        if 'k_windowIfLb0E' in k:
            assert v['Occupancy'] >= 6, (k, v)
