"""Compile ``libbrainspike_hip.so`` for gfx950 with hipcc (in-tree)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = [os.path.join(HERE, 'csrc', f) for f in ('bs_net.hip', 'bs_bptt.hip', 'bs_conn.hip')]
# every file the library is compiled from (sources and the headers they include)
DEPS = SRC + [os.path.join(HERE, 'csrc', f) for f in ('bs_math.h', 'bs_models.h')] + [
    os.path.join(HERE, '..', '..', 'include', 'brainspike_hip.h')]
OUT = os.path.join(HERE, 'libbrainspike_hip.so')


def source_hash() -> str:
    """SHA-256 over the library's source files (name and bytes, in DEPS
    order); compiled into the library and checked by the loader."""
    import hashlib
    h = hashlib.sha256()
    for f in DEPS:
        h.update(os.path.basename(f).encode() + b'\0')
        with open(f, 'rb') as fh:
            h.update(fh.read())
    return h.hexdigest()

# -ffp-contract=off: every arithmetic op is one IEEE-rounded op (no FMA
# contraction), which the FP64 bit-exact parity with the CPU oracle relies on.
# -packed-fp32-ops off: no packed FP32 instructions are generated.
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared', '-ffp-contract=off',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']
# the HIP runtime's directory, for the loader (no other ROCm library is linked)
LIBS = ['-L/opt/rocm/lib', '-Wl,-rpath,/opt/rocm/lib']


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    tmp = OUT + f'.tmp{os.getpid()}'
    cmd = [hipcc] + FLAGS + [f'-DBS_SOURCE_HASH="{source_hash()}"', '-o', tmp] + SRC + LIBS
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('hipcc failed:\n' + res.stdout + res.stderr)
    os.replace(tmp, OUT)
    if verbose:
        print('built', OUT)
    return OUT


def resource_usage(src: str) -> dict:
    """Per-kernel register and private-memory use of one source file for
    gfx950, from the compiler's resource-usage remarks (device-only compile,
    nothing written): ``{mangled name: {'VGPRs': n, 'ScratchSize': bytes per
    lane, 'Occupancy': waves per SIMD, ...}}``. Kernels of library headers
    (rocprim) are left out."""
    import re
    import tempfile
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    with tempfile.TemporaryDirectory() as td:
        cmd = [hipcc] + [f for f in FLAGS if f != '-shared'] + [
            '--cuda-device-only', '-c', src, '-o', os.path.join(td, 'k.o'),
            '-Rpass-analysis=kernel-resource-usage']
        res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('hipcc failed:\n' + res.stderr[-4000:])
    out, cur = {}, None
    base = os.path.basename(src)
    for line in res.stderr.splitlines():
        if 'remark:' not in line:
            continue
        m = re.search(r'remark: Function Name: (\S+)', line)
        if m:
            cur = m.group(1) if os.path.basename(line.split(':', 1)[0]) == base else None
            if cur:
                out[cur] = {}
            continue
        m = re.search(r'remark:\s+([A-Za-z ]+?)(?: \[[^]]*\])?: (\d+)', line)
        if cur and m:
            out[cur][m.group(1).replace(' ', '')] = int(m.group(2))
    return out


if __name__ == '__main__':
    build(verbose=True)
