"""Benchmark: synaptic events/s and real-time factor of Brunel-11250 on one MI355X.

Workload (one GPU): Brunel balanced network, NE = 9000 + NI = 2250
iaf_psc_delta neurons, fixed in-degree CE = 900 / CI = 225 (12,656,250
synapses), per-edge delays uniform over the integer multiples of dt in
1.0-3.0 ms, Poisson drive at 2 nu_th (lambda = 2 per step), dt = 0.1 ms, spikes
recorded from every neuron and V_m from 10, float64.

A bench "step" is one simulated window of ``--window-ms`` (default 1000 ms =
10,000 network steps), the reference benchmark's timed unit
(benchmarks/brunel_simulator_baseline.py:186-343). The network first settles
for ``--settle-ms`` (200 ms) from rest, then runs ``--warmup`` untimed windows,
then ``--steps`` timed windows, with the network resident in HBM.

The JSON line carries:
  value        synaptic events/s: sum over the timed windows' spikes of the
               spiking neuron's out-degree, over the timed wall time (which
               includes sorting and copying the spike records to the host)
  rtf          wall time / simulated time
  roofline     the window kernel's algorithmic bytes per launch (delivery:
               8 B per gathered spike + one edge word per edge, 2 B (4 B when
               a network's delay * stride + lane needs 32 bits); update: state and
               refractory counter read and written once per window; 8 B per
               spike record and V_m sample) over its mean HIP-event duration,
               on a steady-state window after the timed ones; ``traffic`` is
               HBM bytes per launch of the same kernel over a steady-state
               window, from rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by
               a plain device copy
  cpu_baseline the CPU oracle (C++ restatement, one thread) on one simulated
               second of the same network after the same settle, pinned to
               one core and unpinned, with the host's CPU model and affinity
  bptt         the differentiation workload (recurrent iaf_psc_exp, 4096
               neurons, 10 % trainable recurrent and input weights, 200 Poisson
               inputs, batch 64, 1000 steps, checkpoint 100): gradient steps/s
               of the device adjoint; the same episode differentiated by torch
               autograd in float32 on the same GPU (``autodiff_same_gpu``, the
               framework reverse mode jax.grad stands for in the reference);
               and the CPU reference (torch float64 autograd) on a bounded
               sample

The run exits with status 1 when the timed windows deliver no events.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20, help='timed windows of --window-ms each')
    ap.add_argument('--warmup', type=int, default=5, help='untimed windows after the settle')
    ap.add_argument('--window-ms', type=float, default=1000.0)
    ap.add_argument('--settle-ms', type=float, default=200.0)
    ap.add_argument('--dtype', default='float64', choices=['float64', 'float32'])
    ap.add_argument('--cpu-ms', type=float, default=1000.0, help='simulated ms timed on the CPU oracle')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-leg', choices=['pinned', 'unpinned'], default=None,
                    help=argparse.SUPPRESS)
    ap.add_argument('--profile-only', action='store_true',
                    help='run the settle, warm-up and timed windows only (for rocprofv3)')
    ap.add_argument('--bptt-steps', type=int, default=5, help='timed gradient steps (0: skip)')
    ap.add_argument('--bptt-warmup', type=int, default=1)
    ap.add_argument('--bptt-autodiff-steps', type=int, default=None,
                    help='timed gradient steps of the torch.autograd restatement on the same GPU '
                         '(0: skip; default 3, and 0 with --bptt-only)')
    ap.add_argument('--no-pmc', action='store_true',
                    help='skip the rocprofv3 FETCH_SIZE/WRITE_SIZE passes behind roofline.traffic')
    ap.add_argument('--ensemble', type=int, default=0, metavar='M',
                    help='instead: the batched-ensemble workload with M networks (256 = full size)')
    ap.add_argument('--ensemble-steps', type=int, default=5000)
    ap.add_argument('--ens-tile', type=int, default=0,
                    help='ensemble launch shape without autotuning (with --ens-window)')
    ap.add_argument('--ens-window', type=int, default=0)
    ap.add_argument('--connectivity', default='device', choices=['host', 'device'],
                    help='where the ensemble samples its random edges (the networks differ; with '
                         'device edges an engine builds in well under a second, and the autotuner '
                         'times the full ensemble instead of a 16-network sample)')
    ap.add_argument('--build-bench', action='store_true',
                    help='instead: network + engine build times of Brunel-11250 and the '
                         '256-network ensemble with host and with device connectivity')
    ap.add_argument('--aeif-stdp', type=int, default=0, metavar='STEPS',
                    help='instead: the conductance-based aeif_cond_alpha + stdp_synapse network '
                         '(50,000 neurons, p = 0.01) for STEPS steps')
    ap.add_argument('--fp32-report', type=int, default=0, metavar='STEPS',
                    help='instead: Brunel-11250 in float32 against float64 over STEPS steps')
    ap.add_argument('--aeif-wnoise', type=float, default=1.0,
                    help='Poisson drive weight in nS (1.0 as specified leaves the network silent)')
    ap.add_argument('--bptt-only', action='store_true',
                    help='run only the BPTT leg (for rocprofv3)')
    ap.add_argument('--direct', action='store_true',
                    help='launch every window directly instead of replaying the captured graph')
    args = ap.parse_args()
    if args.cpu_leg:
        print(json.dumps(run_cpu_leg(args)), flush=True)
        return
    if args.gpus != 1:
        print(f'bench.py: this build runs on exactly one GPU; --gpus {args.gpus} is not supported '
              '(multi-GPU sharding is design-only, see DESIGN.md)', file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from brainpy.state_amd.networks import brunel_delta

    if args.build_bench:
        print(json.dumps(run_build_bench(args)), flush=True)
        return
    if args.ensemble:
        print(json.dumps(run_ensemble(args)), flush=True)
        return
    if args.aeif_stdp:
        print(json.dumps(run_aeif_stdp(args)), flush=True)
        return
    if args.fp32_report:
        print(json.dumps(run_fp32_report(args)), flush=True)
        return
    if args.bptt_autodiff_steps is None:
        args.bptt_autodiff_steps = 0 if args.bptt_only else 3
    if args.bptt_only:
        args.no_cpu = True
        print(json.dumps({'bptt': run_bptt(args)}), flush=True)
        return

    t0 = time.time()
    sim, h = brunel_delta(order=2250, delay=('uniform', 1.0, 3.0), n_vm=10, dtype=args.dtype)
    desc = sim.build_desc()
    eng = sim._ensure_engine()
    log(f'built Brunel-11250 in {time.time() - t0:.1f}s (window {eng.window()} steps per launch)')
    win = int(round(args.window_ms / sim.dt))
    settle = int(round(args.settle_ms / sim.dt))
    eng.reset()
    if args.direct:
        eng.set_timing(True)
    eng.run(settle)
    for _ in range(args.warmup):
        eng.run(win)
    log(f'settle {settle} + warm-up {args.warmup} x {win} steps done')

    # ---- timed windows (bs_net_run returns after its stream has drained).
    # The interpreter's objects from the imports and the build are collected
    # and frozen first: a full collection over them (tens of ms with torch
    # loaded) would otherwise land inside whichever timed window the
    # allocation count happens to trigger it in.
    import gc
    gc.collect()
    gc.freeze()
    gc2_before = gc.get_stats()[2]['collections']
    torch.cuda.synchronize()
    walls, n_events, n_spikes = [], 0, 0
    t1 = time.perf_counter()
    for _ in range(args.steps):
        tw = time.perf_counter()
        _, (gids, steps) = eng.run(win)
        walls.append(time.perf_counter() - tw)
        n_events += eng.bytes()['n_events']
        n_spikes += len(gids)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    gc.unfreeze()
    log(f'full collections during the timed windows: {gc.get_stats()[2]["collections"] - gc2_before}')
    sim_s = args.steps * win * sim.dt / 1000.0
    value = n_events / wall if wall > 0 else 0.0
    rate = n_spikes / 11250 / sim_s if sim_s > 0 else 0.0
    log(f'timed {args.steps} x {win} steps: {wall:.3f}s (median window {np.median(walls) * 1e3:.2f} ms), '
        f'{n_spikes} spikes ({rate:.1f} Hz), {n_events} events, {value:.3e} ev/s, RTF {wall / max(sim_s, 1e-12):.4f}')
    if args.profile_only:
        return

    # ---- window kernel: HIP events around every launch on the engine's
    # stream, over one more steady-state window
    eng.set_timing(True)
    eng.run(win)
    tm = eng.timing()
    by = eng.bytes()
    eng.set_timing(False)
    n_launch = max(1, tm['n_update'])
    kms = tm['update_ms'] / n_launch
    del_b, upd_b = by['deliver_bytes'] / n_launch, by['update_bytes'] / n_launch
    kbytes = del_b + upd_b
    achieved = kbytes / (kms * 1e-3) / 1e9
    log(f'k_window: {kms * 1e3:.2f} us/launch over {n_launch} launches, {kbytes / 1e6:.3f} MB/launch '
        f'(delivery {del_b / 1e6:.3f} MB = {del_b / (kms * 1e-3) / 1e9:.1f} GB/s, update '
        f'{upd_b / 1e6:.3f} MB = {upd_b / (kms * 1e-3) / 1e9:.1f} GB/s)')
    traffic, tnote = (None, 'skipped (--no-pmc)') if args.no_pmc else pmc_traffic('k_window', n_last=n_launch)
    log(f'k_window HBM traffic per launch: {traffic} ({tnote})')
    roofline = {'bound': 'hbm', 'kernel': 'k_window', 'achieved': round(achieved, 2),
                'peak': 8000.0, 'unit': 'GB/s', 'frac': round(achieved / 8000.0, 5),
                'traffic': traffic, 'traffic_note': tnote, 'bytes_per_launch': round(kbytes),
                'delivery_bytes_per_launch': round(del_b), 'update_bytes_per_launch': round(upd_b),
                'delivery_gbs': round(del_b / (kms * 1e-3) / 1e9, 2),
                'update_gbs': round(upd_b / (kms * 1e-3) / 1e9, 2),
                'us_per_launch': round(kms * 1e3, 3), 'launches': n_launch,
                'launch_fraction_of_wall': round(tm['update_ms'] / max(1e-9, tm['total_ms']), 4)}

    # ---- BPTT leg, before the CPU legs, with the network engine released
    n_synapses = int(sum(desc.c.projs[k].n_edges for k in range(4)))
    eng.close()
    sim._engine = None
    bptt = run_bptt(args) if args.bptt_steps > 0 else None

    # ---- CPU baseline: the oracle on one simulated second of the same
    # network, in child processes (pinned to one core, then unpinned)
    cpu = None
    if not args.no_cpu:
        import subprocess
        legs = {}
        for leg in ('pinned', 'unpinned'):
            cmd = [sys.executable, os.path.join(ROOT, 'bench.py'), '--cpu-leg', leg,
                   '--cpu-ms', str(args.cpu_ms), '--settle-ms', str(args.settle_ms), '--dtype', args.dtype]
            res = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
            if res.returncode != 0:
                log(f'cpu leg {leg} failed:\n{res.stderr[-2000:]}')
                continue
            legs[leg] = json.loads(res.stdout.strip().splitlines()[-1])
            log(f'cpu oracle ({leg}): {legs[leg]["events_per_s"]:.3e} ev/s, RTF {legs[leg]["rtf"]:.3f}')
        if 'pinned' in legs:
            p = legs['pinned']
            cpu = {'value': p['events_per_s'], 'unit': 'events/s', 'cores': 1, 'kind': 'port',
                   'sample': f'simulated ms {args.settle_ms:.0f}-{args.settle_ms + args.cpu_ms:.0f} of the same '
                             f'network (after the same settle from rest), {p["spikes"]} spikes, '
                             f'{p["wall_s"]:.1f}s wall; C++ oracle, one thread, pinned to core {p["core"]}',
                   'rtf': p['rtf'], 'cpu_model': p['cpu_model'], 'affinity_cpus': p['affinity'],
                   'os_cpu_count': p['cpu_count'], 'loadavg': p['loadavg'],
                   'unpinned': legs.get('unpinned')}

    ms_step = wall / max(1, args.steps) * 1e3
    out = {
        'metric': 'synaptic events/s (Brunel-11250, 1 GPU)',
        'value': round(value, 1),
        'unit': 'events/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms_step, 4),
        'step_unit': f'one simulated window of {args.window_ms:g} ms ({win} network steps of dt = {sim.dt} ms)',
        'ms_per_step_median': round(float(np.median(walls)) * 1e3, 4),
        'ms_per_window': [round(w * 1e3, 3) for w in walls],
        'ms_per_network_step': round(ms_step / win, 6),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64' if args.dtype == 'float64' else 'f32',
        'data': 'synthetic (seeded Brunel network, on-device Poisson drive)',
        'config': {'workload': 'brunel-11250 iaf_psc_delta, CE=900 CI=225, per-edge delays '
                               '1.0-3.0 ms, dt=0.1 ms', 'n_neurons': 11250,
                   'n_synapses': n_synapses, 'settle_ms': args.settle_ms,
                   'sim_ms_timed': args.steps * win * sim.dt, 'parallelism': 'single-gpu'},
        'rtf': round(wall / sim_s, 6),
        'firing_rate_hz': round(rate, 2),
        'n_events': int(n_events),
        'roofline': roofline,
        'cpu_baseline': cpu,
        'bptt': bptt,
    }
    print(json.dumps(out), flush=True)
    if n_events == 0:
        log('bench: the timed windows delivered no synaptic events; the measurement is invalid')
        sys.exit(1)


def run_cpu_leg(args):
    """One CPU-oracle leg (child process): pinned to the first core of the
    affinity set, or unpinned; settle from rest, then time ``--cpu-ms``."""
    allowed = sorted(os.sched_getaffinity(0))
    core = None
    if args.cpu_leg == 'pinned':
        core = allowed[0]
        os.sched_setaffinity(0, {core})
    model = ''
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    import oracle
    from brainpy.state_amd.networks import brunel_delta
    sim, h = brunel_delta(order=2250, delay=('uniform', 1.0, 3.0), n_vm=10, dtype=args.dtype)
    desc = sim.build_desc()
    import numpy as np
    rowlen = np.zeros(desc.ntot, dtype=np.int64)
    for pj in desc.c.projs[:desc.c.n_projs]:
        gid0 = sum(desc.c.pops[k].n for k in range(pj.pre_pop))
        pre = np.ctypeslib.as_array(pj.pre, shape=(pj.n_edges,))
        rowlen += np.bincount(pre + gid0, minlength=desc.ntot)
    orc = oracle.OracleNet(desc)
    orc.run(int(round(args.settle_ms / sim.dt)))
    n = int(round(args.cpu_ms / sim.dt))
    la0 = os.getloadavg()
    t0 = time.perf_counter()
    _, g, _ = orc.run(n)
    wall = time.perf_counter() - t0
    la1 = os.getloadavg()
    ev = int(rowlen[g].sum())
    return {'leg': args.cpu_leg, 'core': core, 'events_per_s': round(ev / wall, 1), 'events': ev,
            'spikes': int(len(g)), 'wall_s': round(wall, 3), 'rtf': round(wall / (n * sim.dt * 1e-3), 3),
            'cores_used': 1, 'affinity': len(allowed), 'cpu_count': os.cpu_count(), 'cpu_model': model,
            'loadavg': [round(la0[0], 2), round(la1[0], 2)]}


def pmc_traffic(kernel, n_last, prof_args=None):
    """HBM bytes per launch of ``kernel`` from two rocprofv3 counter passes
    (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) over a profile-only run of
    the same network (settle, one warm-up window, one timed window), each
    scaled by what the counter reads for a 256 MiB device copy. Only the last
    ``n_last`` dispatches (the timed steady-state window) are averaged.
    Returns (bytes or None, note)."""
    import glob
    import shutil
    import sqlite3
    import subprocess
    import tempfile
    exe = shutil.which('rocprofv3')
    if exe is None:
        return None, 'rocprofv3 not found'
    tmp = tempfile.mkdtemp(prefix='bench_pmc_', dir='/tmp')
    env = dict(os.environ, TMPDIR='/tmp')
    prof = [os.path.join(ROOT, 'bench.py')] + (prof_args or [
        '--steps', '1', '--warmup', '1', '--profile-only', '--no-cpu', '--bptt-steps', '0', '--no-pmc'])
    copy = ['-c', "import torch; x = torch.ones(1 << 28, dtype=torch.uint8, device='cuda'); "
                  "y = torch.empty_like(x); [y.copy_(x) for _ in range(4)]; torch.cuda.synchronize()"]

    def values(counter, argv, tag, name):
        log(f'rocprofv3 --pmc {counter} pass ({tag})')
        out = os.path.join(tmp, tag)
        cmd = [exe, '--pmc', counter, '-d', out, '-o', tag, '--', sys.executable, '-u'] + argv
        try:
            subprocess.run(cmd, env=env, cwd='/tmp', capture_output=True, timeout=150, check=False)
        except subprocess.TimeoutExpired:
            return None
        dbs = glob.glob(os.path.join(out, '**', '*.db'), recursive=True)
        if not dbs:
            return None
        con = sqlite3.connect(dbs[0])
        rows = con.execute('select value from counters_collection where kernel_name like ? '
                           'order by dispatch_id', (f'%{name}%',)).fetchall()
        con.close()
        return [r[0] for r in rows] or None

    fetch = values('FETCH_SIZE', prof, 'fetch', kernel)
    write = values('WRITE_SIZE', prof, 'write', kernel) if fetch else None
    cf = values('FETCH_SIZE', copy, 'copy_fetch', 'copyBuffer') if write else None
    cw = values('WRITE_SIZE', copy, 'copy_write', 'copyBuffer') if cf else None
    shutil.rmtree(tmp, ignore_errors=True)
    if not cw:
        return None, 'rocprofv3 counter passes gave no rows'
    copy_kib = (1 << 28) / 1024.0
    ff, fw = copy_kib / (sum(cf) / len(cf)), copy_kib / (sum(cw) / len(cw))
    f = fetch[-n_last:]
    w = write[-n_last:]
    kib = sum(f) / len(f) * ff + sum(w) / len(w) * fw
    return round(kib * 1024.0), (f'FETCH_SIZE x {ff:.3f} + WRITE_SIZE x {fw:.3f} (copy-kernel '
                                 f'scale), mean of the last {len(f)} dispatches (one steady-state window)')


def run_build_bench(args):
    """Seconds to sample the network (Simulator.connect) and to create the
    engine (bs_net_create: gather segments, tables, state), for Brunel-11250
    (12.7M edges, per-edge delays) and the 256-network ensemble (1.76e9
    edges), with the random rules drawn on the host (NumPy PCG64, segments by
    a host counting sort) and on the device (Threefry samplers, segments by a
    device radix sort). The two modes draw different networks of the same
    distribution."""
    import torch
    from brainpy.state_amd import Ensemble
    from brainpy.state_amd.networks import brunel_delta, brunel_exp_ensemble
    out = {'metric': 'network build seconds (sampling + engine creation)', 'unit': 's',
           'higher_is_better': False, 'dtype': 'int32 edges', 'cases': {}}
    for mode in ('host', 'device'):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim, h = brunel_delta(order=2250, delay=('uniform', 1.0, 3.0), n_vm=10, connectivity=mode)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        eng = sim._ensure_engine()
        t2 = time.perf_counter()
        E = sum(int(p.n_edges) for p in eng.desc.c.projs[:eng.desc.c.n_projs])
        eng.close()
        sim._engine = None
        out['cases'][f'brunel11250/{mode}'] = {'edges': E, 'sample_s': round(t1 - t0, 3),
                                               'engine_s': round(t2 - t1, 3),
                                               'edges_per_s': round(E / (t2 - t0), 1)}
        log(f'brunel-11250 {mode}: {E} edges, sampled in {t1 - t0:.2f}s, engine {t2 - t1:.2f}s')
        del sim, h
    for mode in ('host', 'device'):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sims, prm = brunel_exp_ensemble(256, seed=0, connectivity=mode)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ens = Ensemble(sims, tile_size=128)
        t2 = time.perf_counter()
        E = sum(int(p.n_edges) for p in ens._engine.desc.c.projs[:ens._engine.desc.c.n_projs])
        ens.reset()
        _, (g, _) = ens.run_raw(500)
        ens.close()
        out['cases'][f'ensemble256/{mode}'] = {'edges': E, 'sample_s': round(t1 - t0, 3),
                                               'engine_s': round(t2 - t1, 3),
                                               'edges_per_s': round(E / (t2 - t0), 1),
                                               'spikes_500_steps': int(len(g))}
        log(f'ensemble-256 {mode}: {E} edges, sampled in {t1 - t0:.2f}s, engine {t2 - t1:.2f}s, '
            f'{len(g)} spikes in 500 steps')
        del sims, ens
    return out


def run_ensemble(args):
    """Independent Brunel iaf_psc_exp networks (N log-uniform 500-20000, 10 %
    in-degree, g ~ U[3,7], nu_ext/nu_thr ~ U[0.9,4], tau_syn 0.5 ms, float32)
    advanced together; the launch shape comes from the autotuner (timed on
    the whole ensemble when its edges are sampled on the device, else on a
    16-network sample); the sample is also run one by one for comparison."""
    import numpy as np
    from brainpy.state_amd import Ensemble
    from brainpy.state_amd.networks import brunel_exp_ensemble
    M, steps = args.ensemble, args.ensemble_steps
    t0 = time.time()
    sims, prm = brunel_exp_ensemble(M, seed=0, connectivity=args.connectivity)
    syn = int(sum(0.1 * n * n for n in prm['N']))
    log(f'ensemble: {M} networks, {int(prm["N"].sum())} neurons, ~{syn:.3e} synapses, '
        f'built in {time.time() - t0:.1f}s')
    rng = np.random.default_rng(1)
    sample = sorted(rng.choice(M, size=min(16, M), replace=False).tolist())
    if args.ens_tile:
        best, table, tune_on = (args.ens_tile, args.ens_window), {}, []
    else:
        tune_on = sims if args.connectivity == 'device' else [sims[k] for k in sample]
        best, table = Ensemble.autotune(tune_on, steps=300)
        log(f'autotuned on {len(tune_on)} networks')
        log('autotune (tile_size, max_window) -> ms/step: ' +
            ', '.join(f'{k}: {v * 1e3:.3f}' for k, v in sorted(table.items())))
    t1 = time.time()
    ens = Ensemble(sims, tile_size=best[0], max_window=best[1])
    log(f'engine created in {time.time() - t1:.1f}s (tile {best[0]}, window bound {best[1]})')
    if args.profile_only:
        ens.reset()
        ens.run_raw(200)
        _, (g_p, _) = ens.run_raw(steps)
        log(f'profile run: {steps} steps, {len(g_p)} spikes')
        ens.close()
        return {'profile_only': True}
    ens.reset()
    ens.run_raw(200)
    t2 = time.perf_counter()
    vm, (g_all, st_all) = ens.run_raw(steps)
    wall = time.perf_counter() - t2
    by = ens.engine_bytes()
    parts = ens.split(vm, g_all, st_all)
    n_sp = len(g_all)
    counts_ens = {k: len(parts[k][1]) for k in sample}
    # window kernel: HIP events around each launch over 320 steady-state steps
    ens._engine.set_timing(True)
    ens.run_raw(320)
    tm, by2 = ens._engine.timing(), ens.engine_bytes()
    ens._engine.set_timing(False)
    n_launch = max(1, tm['n_update'])
    kms = tm['update_ms'] / n_launch
    del_b, upd_b = by2['deliver_bytes'] / n_launch, by2['update_bytes'] / n_launch
    achieved = (del_b + upd_b) / (kms * 1e-3) / 1e9
    ens.close()
    traffic, tnote = (None, 'skipped (--no-pmc)') if args.no_pmc else pmc_traffic(
        'k_window', n_launch, ['--ensemble', str(M), '--ensemble-steps', '320', '--profile-only',
                               '--ens-tile', str(best[0]), '--ens-window', str(best[1]),
                               '--connectivity', args.connectivity, '--no-pmc'])
    roofline = {'bound': 'hbm', 'kernel': 'k_window', 'achieved': round(achieved, 2), 'peak': 8000.0,
                'unit': 'GB/s', 'frac': round(achieved / 8000.0, 5), 'traffic': traffic, 'traffic_note': tnote,
                'bytes_per_launch': round(del_b + upd_b), 'delivery_bytes_per_launch': round(del_b),
                'update_bytes_per_launch': round(upd_b), 'us_per_launch': round(kms * 1e3, 3),
                'launches': n_launch}
    log(f'k_window (ensemble): {kms * 1e3:.1f} us/launch, {(del_b + upd_b) / 1e6:.2f} MB/launch, '
        f'{achieved:.1f} GB/s; traffic {traffic} ({tnote})')
    log(f'ensemble: {steps} steps in {wall:.3f}s, {n_sp} spikes, {by["n_events"]:.3e} events')
    # the sample networks one by one (plain runs), same steps after the same warm-up
    seq = 0.0
    same = True
    for k in sample:
        eng = sims[k]._ensure_engine()
        eng.reset()
        eng.run(200)
        t3 = time.perf_counter()
        _, (g, _) = eng.run(steps)
        seq += time.perf_counter() - t3
        same &= len(g) == counts_ens[k]
        eng.close()
        sims[k]._engine = None
    t4 = time.time()
    ens_s = Ensemble([sims[k] for k in sample], tile_size=best[0], max_window=best[1])
    ens_s.reset()
    ens_s.run_raw(200)
    t5 = time.perf_counter()
    ens_s.run_raw(steps)
    ens_sample = time.perf_counter() - t5
    ens_s.close()
    log(f'sample of {len(sample)}: sequential plain runs {seq:.3f}s, one ensemble {ens_sample:.3f}s, '
        f'spike counts identical: {same}')
    return {'metric': 'synaptic events/s (batched ensemble, 1 GPU)', 'value': round(by['n_events'] / wall, 1),
            'unit': 'events/s', 'steps': steps, 'ms_per_step': round(wall / steps * 1e3, 4),
            'rtf': round(wall / (steps * sims[0].dt * 1e-3), 4), 'dtype': 'f32',
            'config': {'workload': 'brunel iaf_psc_exp ensemble', 'networks': M, 'connectivity': args.connectivity,
                       'neurons': int(prm['N'].sum()), 'synapses_approx': syn, 'tile_size': best[0],
                       'max_window': best[1]},
            'roofline': roofline,
            'autotuned_on_networks': len(tune_on),
            'autotune_ms_per_step': {f'{k[0]}/{k[1]}': round(v * 1e3, 4) for k, v in table.items()},
            'sample': {'networks': len(sample), 'sequential_s': round(seq, 4),
                       'ensemble_s': round(ens_sample, 4), 'spike_counts_identical': bool(same)}}


def run_fp32_report(args):
    """Brunel-11250 (per-edge delays) in float32 against float64 (the float64
    engine is bit-exact with the float64 oracle): V_m error before the first
    raster divergence, and the deviation of the mean rate and ISI CV."""
    from brainpy.state_amd import stats
    from brainpy.state_amd.networks import brunel_delta
    n = args.fp32_report
    runs = {}
    for dt in ('float64', 'float32'):
        sim, h = brunel_delta(order=2250, delay=('uniform', 1.0, 3.0), n_vm=10, dtype=dt)
        eng = sim._ensure_engine()
        eng.reset()
        t0 = time.perf_counter()
        vm, (g, st) = eng.run(n)
        runs[dt] = (g, st, vm)
        log(f'{dt}: {n} steps in {time.perf_counter() - t0:.3f}s, {len(g)} spikes')
        eng.close()
        sim._engine = None
    rep = stats.compare_runs(runs['float64'], runs['float32'], 11250, n, 0.1)
    log(json.dumps(rep))
    return {'fp32_report': rep, 'steps': n, 'config': {'workload': 'brunel-11250 iaf_psc_delta, '
                                                                   'per-edge delays 1.0-3.0 ms'}}


def run_aeif_stdp(args):
    """aeif_cond_alpha (adaptive RKF45) 40,000 E + 10,000 I, p = 0.01 (~25M
    synapses), stdp_synapse (additive) on E->E (~16M plastic edges), delays
    1-5 ms per edge, Poisson 8 kHz at --aeif-wnoise nS, float64."""
    import numpy as np
    from brainpy.state_amd.networks import aeif_stdp
    t0 = time.time()
    sim, h = aeif_stdp(n=50000, w_noise=args.aeif_wnoise, n_vm=0)
    desc = sim.build_desc()
    E = sum(desc.c.projs[k].n_edges for k in range(desc.c.n_projs))
    eng = sim._ensure_engine()
    log(f'aeif+stdp: {E} synapses, built in {time.time() - t0:.1f}s')
    eng.reset()
    t1 = time.perf_counter()
    _, (g, st) = eng.run(args.aeif_stdp)
    wall = time.perf_counter() - t1
    by = eng.bytes()
    w = sim.get_weights(h['pee'])
    sim_s = args.aeif_stdp * sim.dt * 1e-3
    log(f'aeif+stdp: {args.aeif_stdp} steps in {wall:.2f}s, {len(g)} spikes '
        f'({len(g) / 50000 / sim_s:.2f} Hz), weights mean {w.mean():.4f} [{w.min():.4f}, {w.max():.4f}]')
    return {'metric': 'synaptic events/s (aeif_cond_alpha + stdp_synapse, 1 GPU)',
            'value': round(by['n_events'] / wall, 1), 'unit': 'events/s', 'steps': args.aeif_stdp,
            'ms_per_step': round(wall / args.aeif_stdp * 1e3, 4), 'rtf': round(wall / sim_s, 4),
            'dtype': 'f64', 'rate_hz': round(len(g) / 50000 / sim_s, 3),
            'config': {'workload': 'aeif_cond_alpha 40000 E + 10000 I, p=0.01, stdp_synapse E->E, '
                                   'delays 1-5 ms, Poisson 8 kHz', 'w_noise_nS': args.aeif_wnoise,
                       'synapses': int(E)},
            'weights': {'mean': float(w.mean()), 'min': float(w.min()), 'max': float(w.max())}}


def run_bptt(args):
    """Gradient steps/s of the device BPTT on the differentiation workload."""
    import numpy as np
    import torch
    from brainpy.state_amd.bptt import BPTT, bptt_problem
    net, w_in, w_rec, cnt, target = bptt_problem()
    B, T = cnt.shape[1], cnt.shape[0]
    bp = BPTT(net, batch=B, n_steps=T, checkpoint=100)
    dev = [torch.tensor(a).cuda() for a in (w_in, w_rec, cnt, target)]
    for _ in range(args.bptt_warmup):
        bp.loss_and_grad(*dev)
    t0 = time.perf_counter()
    for _ in range(args.bptt_steps):
        loss, g_in, g_rec = bp.loss_and_grad(*dev)
    wall = (time.perf_counter() - t0) / args.bptt_steps
    _, counts = bp.spikes()
    log(f'bptt: {wall * 1e3:.2f} ms per gradient step ({1 / wall:.2f}/s), loss {loss:.3f}, '
        f'mean rate {counts.mean() / (T * net.dt * 1e-3):.1f} Hz')
    autodiff = None
    if args.bptt_autodiff_steps > 0:
        autodiff = run_autodiff(net, dev, args.bptt_autodiff_steps, loss, g_in, g_rec)
        autodiff['speedup_of_fused_adjoint'] = round(autodiff['ms_per_grad_step'] / (wall * 1e3), 2)
        log(f'bptt torch.autograd on the same GPU: {autodiff["ms_per_grad_step"]:.1f} ms per gradient step '
            f'(fused adjoint {autodiff["speedup_of_fused_adjoint"]:.1f}x faster), loss {autodiff["loss"]:.3f}, '
            f'gradient difference {autodiff["grad_rel_diff_rec"]:.2e} (recurrent), '
            f'{autodiff["grad_rel_diff_in"]:.2e} (input)')
    res = {'metric': 'BPTT gradient steps/s', 'value': round(1.0 / wall, 3), 'unit': 'grad-steps/s',
           'ms_per_grad_step': round(wall * 1e3, 3), 'steps': args.bptt_steps,
           'warmup': args.bptt_warmup, 'dtype': 'f32 (transposed delivery bf16x3 split, f32 accumulate)',
           'config': {'workload': 'recurrent iaf_psc_exp 3277 E + 819 I, 10% trainable recurrent and '
                                  '200x4096 10% input weights, Poisson 5-50 Hz inputs, rate_mse to '
                                  'U[2,20] Hz, fast-sigmoid surrogate 5/mV, delay 1 ms',
                      'batch': B, 'n_steps': T, 'checkpoint': 100, 'dt_ms': net.dt},
           'vs_baseline': None, 'autodiff_same_gpu': autodiff, 'cpu_baseline': None}
    if not args.no_cpu:
        from oracle import bptt_ref as R
        Bs, Ts = 2, 60
        c_s = np.ascontiguousarray(cnt[:Ts, :Bs])
        t1 = time.perf_counter()
        fwd = R.forward_f32(net, w_in, w_rec, c_s, target)
        R.grad_autograd(net, w_in, w_rec, c_s, target, fwd)
        cw = time.perf_counter() - t1
        per_step = cw * (B / Bs) * (T / Ts)
        res['cpu_baseline'] = {
            'value': round(1.0 / per_step, 6), 'unit': 'grad-steps/s', 'cores': torch.get_num_threads(),
            'kind': 'port',
            'sample': f'{Bs} trials x {Ts} steps of the same network (NumPy float32 forward + torch '
                      f'float64 autograd), {cw:.1f}s wall, scaled linearly to {B} x {T}'}
        log(f'bptt cpu reference: {cw:.1f}s for {Bs}x{Ts}, ~{per_step:.0f}s per full gradient step')
    return res


def run_autodiff(net, dev, n_timed, loss_fused, g_in_fused, g_rec_fused):
    """The differentiation workload's episode differentiated by a framework
    instead of the fused adjoint: plain torch autograd in float32 on the same
    GPU (JAX, whose jax.grad the reference uses, is not installed here). The
    same network, inputs, update (_kernel/models/iaf_psc_exp.py:136-151), spike
    rule with the fast-sigmoid straight-through surrogate (spiking.py:65-79),
    detached reset and rate_mse loss (bptt.py:170-191); the forward computes
    its own spikes; every step's activations are kept (no checkpointing: the
    fastest form of framework reverse mode). Dense matrix products stand for
    the event matmuls. The network is padded with 64 disconnected neurons
    (zero weights in and out, outside the per-class products and the loss),
    so that no operand slice of those products ends at the end of its
    allocation. Returns ms per gradient step,
    its loss and the relative difference of its gradients from the fused
    adjoint's (its float32 sums run in another order, so near-threshold spikes
    may differ)."""
    import torch
    w_in, w_rec, cnt, tgt = dev
    k = net.constants()
    T, B, _ = cnt.shape
    N, nE = net.n, net.n_exc
    d, d_in = net.delay_steps, net.delay_in_steps
    f32 = torch.float32
    alpha = k['surrogate_alpha']
    span = k['V_th'] - k['V_reset']
    reset_to = k['ref_steps'] + 1 if k['ref_steps'] > 0 else 0
    T_s = T * net.dt * 1e-3
    cntf = cnt.to(f32)
    NP = N + 64  # padded neuron count

    def episode(Win, Wrec):
        V = torch.full((B, NP), k['E_L'], dtype=f32, device=Win.device)
        ie = torch.zeros_like(V)
        ii = torch.zeros_like(V)
        r = torch.zeros((B, NP), dtype=torch.int32, device=Win.device)
        spikes = []
        for n in range(T):
            xe = cntf[n - d_in] @ Win if n >= d_in else torch.zeros_like(V)
            xi = torch.zeros_like(V)
            if n >= d:
                s = spikes[n - d]
                xe = xe + s[:, :nE] @ Wrec[:nE]
                xi = xi + s[:, nE:N] @ Wrec[nE:N]
            ref = r > 0
            y = V - k['E_L']
            y_new = y * k['P22'] + ie * k['P21ex'] + ii * k['P21in'] + k['I_e'] * k['P20']
            y = torch.where(ref, y, y_new)
            ie = ie * k['P11ex'] + xe
            ii = ii * k['P11in'] + xi
            V = y + k['E_L']
            m = (V - k['V_th']) / span
            hard = (~ref) & (m.detach() >= 0)
            g = m / (1.0 + alpha * m.abs())
            spikes.append(hard.to(f32) + (g - g.detach()) * (~ref).to(f32))
            V = torch.where(hard, torch.full_like(V, k['V_reset']), V)
            r = torch.clamp(torch.where(hard, torch.full_like(r, reset_to), r) - 1, min=0)
        rate = torch.stack(spikes).sum(0)[:, :N] / T_s
        return ((rate - tgt[None, :]) ** 2).mean()

    def grad_step():
        Win = torch.zeros((w_in.shape[0], NP), dtype=f32, device=w_in.device)
        Win[:, :N] = w_in
        Wrec = torch.zeros((NP, NP), dtype=f32, device=w_rec.device)
        Wrec[:N, :N] = w_rec
        Win.requires_grad_(True)
        Wrec.requires_grad_(True)
        L = episode(Win, Wrec)
        gi, gr = torch.autograd.grad(L, [Win, Wrec])
        return float(L.detach()), gi[:, :N], gr[:N, :N]

    grad_step()  # warm-up (library initialisation, allocator)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n_timed):
        L, gi, gr = grad_step()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / n_timed

    def rel(a, b):
        b = torch.as_tensor(b, device=a.device, dtype=f32)
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b).clamp_min(1e-30))

    return {'metric': 'torch.autograd gradient steps/s, same GPU', 'value': round(1.0 / wall, 3),
            'unit': 'grad-steps/s', 'ms_per_grad_step': round(wall * 1e3, 2), 'steps': n_timed,
            'dtype': 'f32', 'loss': L, 'loss_fused_adjoint': float(loss_fused),
            'grad_rel_diff_rec': rel(gr, g_rec_fused), 'grad_rel_diff_in': rel(gi, g_in_fused),
            'form': 'plain reverse mode over all 1000 steps, dense matrix products, no checkpointing'}


if __name__ == '__main__':
    main()
