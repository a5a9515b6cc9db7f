"""Throughput of the LSRK time-stepping path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rupture]

A step is one LSRK(5,4) time step (five RK stages) of the whole two-block
problem.  The default workload is the curvilinear rupture problem
(2 x 201x101x101 nodes, h = 100 m, slip-weakening fault, random topography),
FP32 volume fields with the FP64 fault; the FP64 path is timed beside it.

Metric: M grid-pt RK stages/s = nodes x 5 x steps / seconds / 1e6 (owned
nodes, device-synchronised, no output).  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'waveqlab3d-gh_amd'))

HBM_PEAK_GBS = 8000.0
# launch variants of the interior volume kernel (wql_api.hip): family, tile
VARIANTS = [('fused', 32, 8), ('hold', 32, 8), ('fused', 64, 8), ('hold', 64, 8), ('fused', 32, 16),
            ('hold', 32, 16), ('split', 32, 8)]


def per_node(prob):
    """the upwind / upwind-DRP operators run on the operator kernels (the
    tiled k_volume_opt and the per-node k_volume_op)"""
    return prob.fd_type != 'traditional'


def kernel_name(prob, precision, variant):
    """template name of the main volume kernel a launch variant selects"""
    if per_node(prob):
        # the operator path: the tiled k_volume_opt (upwind / DRP operators)
        # and k_volume_op on the remaining nodes, timed together
        real = "float" if precision == "fp32" else "double"
        return f'k_volume_opt<{real}> + k_volume_op<{real}>'
    fam, tx, ty = VARIANTS[variant]
    an = prob.anelastic is not None
    if an:
        # the anelastic response runs on 32 x 8 tiles, fused or held
        fam, tx, ty = ('hold', 32, 8) if fam == 'hold' else ('fused', 32, 8)
    curv = 'true' if any(b.mesh != 'cartesian' for b in prob.blocks) else 'false'
    het = 'true' if any(b.material is not None or b.medium is not None for b in prob.blocks) else 'false'
    real = 'float' if precision == 'fp32' else 'double'
    if fam == 'split':
        return f'k_volume_split<{real}, {tx}, {ty}, {curv}, {het}, 0>'  # interior (not s-closure) instantiation
    hold = 'true' if fam == 'hold' else 'false'  # lead plane held in registers
    return f'k_volume_fused<{real}, {tx}, {ty}, {curv}, {het}, {hold}, {"true" if an else "false"}>'


def measured_traffic(workload, shape, kname):
    """HBM bytes per launch of the kernel from the committed FETCH_SIZE /
    WRITE_SIZE counter summary (copy-calibrated), if it was taken at this
    workload and launch shape"""
    try:
        with open(os.path.join(ROOT, 'profiles', 'traffic.json')) as f:
            allj = json.load(f)
    except (OSError, ValueError):
        return None
    for entry in allj.values():
        if entry.get('workload') != workload or entry.get('tile_variant') != shape[0] or \
                entry.get('zchunk') != shape[1] or entry.get('tile_order', 0) != shape[2]:
            continue
        k = entry['kernels'].get(kname)
        if k:
            return k['read_bytes'] + k['write_bytes']
    return None


def main_kernel_nodes(prob):
    """nodes handled by the main volume kernel per launch (both blocks)"""
    if per_node(prob):
        return sum(b.n[0] * b.n[1] * b.n[2] for b in prob.blocks)
    tot = 0
    for b in prob.blocks:
        nq, nr, ns = b.n
        lo = lambda on: max(b.npml, 6) if on else 0
        x0, x1 = lo(b.pml[0]), nq - lo(b.pml[1])
        y0, y1 = lo(b.pml[2]), nr - lo(b.pml[3])
        z0 = max(b.npml, 6) if b.pml[4] else 6
        z1 = ns - (max(b.npml, 6) if b.pml[5] else 6)
        tot += max(0, x1 - x0) * max(0, y1 - y0) * max(0, z1 - z0)
    return tot


def bytes_per_node(prob, precision):
    """algorithmic HBM bytes per node and RK stage of the volume update,
    averaged over the five stages of a step: F in 9 R, DF 9 R + 9 W, F out
    9 W, except that the first stage (A = 0) reads no DF; + 10 (metrics + J)
    if curvilinear; + 3 (lambda, mu, rho) if heterogeneous; with the
    anelastic response + 98 (memory variables 24 R + 24 W, their rates 24 R
    + 24 W, Q^-1 2 R), less the rates' read at the first stage"""
    vals = 36 - 9 / 5
    if prob.anelastic is not None:
        vals += 98 - 24 / 5
    if any(b.mesh != 'cartesian' for b in prob.blocks):
        vals += 10
    if any(b.material is not None or b.medium is not None for b in prob.blocks):
        vals += 3
    return vals * (4 if precision == 'fp32' else 8)


def run(W, precision, steps, warmup, tile, zchunk, graph, order=None):
    import torch
    from wql import native
    from wql.solver import Simulation
    prob = W['problem']
    sim = Simulation(prob, precision=precision, tile=tile, zchunk=zchunk, coords=W['coords'], graph=graph)
    if order is not None:
        native.check(native.lib().wql_set_tile_order(sim.ctx, *order))
    if W.get('F0') is not None:
        for ib in range(2):
            sim.set_fields(ib, W['F0'][ib])
    if W.get('init_state'):
        sim.init_state()
    sim.step(warmup)
    sim.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    # per-kernel timing of the main volume kernel (HIP events on the library stream)
    sim.profile(True)
    sim.step(2)
    sim.sync()
    kms = sim.volume_kernel_ms()
    sim.profile(False)
    bytes_dev = sim.device_bytes()
    shape = sim.tile
    sim.close()
    return sec, kms, bytes_dev, shape


def fp32_error(W, steps):
    """FP32-volume against FP64 on the same workload and initial state after
    `steps` steps: relative L2 of block 1's fields, and of slip, rupture time
    (nodes ruptured in FP64) and peak slip rate on block 1's fault copy
    (None for a fault that does not slip, such as a locked interface)"""
    import numpy as np
    from wql.solver import Simulation
    out = {}
    for prec in ('fp64', 'fp32'):
        with Simulation(W['problem'], precision=prec, coords=W['coords']) as sim:
            if W.get('F0') is not None:
                for ib in range(2):
                    sim.set_fields(ib, W['F0'][ib])
            if W.get('init_state'):
                sim.init_state()
            sim.step(steps)
            sim.sync()
            f = sim.fault(0)
            out[prec] = dict(slip=f['S'][3], trup=f['trup'][0], psr=sim.fault_peak(0)[0], F=sim.fields(0))
    a, b = out['fp64'], out['fp32']
    rel = lambda x, y: float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))  # noqa: E731
    rup = a['trup'] < 1e9
    both = rup & (b['trup'] < 1e9)
    slipping = float(np.abs(a['slip']).max()) > 1e-9
    return {'steps': steps, 'fields_block1': rel(b['F'], a['F']),
            'slip': rel(b['slip'], a['slip']) if slipping else None,
            'rupture_time': rel(b['trup'][both], a['trup'][both]) if both.any() else None,
            'peak_slip_rate': rel(b['psr'], a['psr']) if slipping else None,
            'ruptured_nodes_fp64': int(rup.sum()), 'ruptured_nodes_fp32': int((b['trup'] < 1e9).sum()),
            'max_slip_fp64': float(a['slip'].max())}


def serving(n_scen, steps, rate, max_active, n_seq, seed=5, fifo=True):
    """the serving stream: n_scen scenarios of the seeded size/friction
    distribution, Poisson arrivals at `rate`/s, `steps` steps each, FP32
    fields / FP64 friction through the continuous-batching server; beside
    it the same scenarios run one after another in plain FP64 (first n_seq)"""
    import numpy as np
    from wql import workloads
    from wql.serving import run_stream
    from wql.solver import Simulation
    scen = []
    for k in range(n_scen):
        sc = workloads.scenario(k, seed=seed)
        sc['steps'] = steps
        scen.append(sc)
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_scen))
    arrivals -= arrivals[0]
    # warm the library (first graph capture, code objects) outside the timing
    run_stream(scen[:1], np.zeros(1), max_active=1, precision='fp32')

    def progress(tag):
        def cb(n):
            if n % 16 == 0 or n == n_scen:
                print(f'[bench] serving {tag}: {n}/{n_scen} scenarios retired', file=sys.stderr, flush=True)
        return cb
    # admission shortest job first with aging; arrival order beside it
    pool = {}
    lat, wall, _ = run_stream(scen, arrivals, max_active=max_active, precision='fp32', policy='sjf', stats=pool,
                              progress=progress('sjf'))
    lat_f, wall_f, _ = (None, None, None) if not fifo else \
        run_stream(scen, arrivals, max_active=max_active, precision='fp32', policy='fifo', progress=progress('fifo'))
    pts = sum(workloads.points(s['problem']) for s in scen) * 5 * steps
    t0 = time.perf_counter()
    for k, sc in enumerate(scen[:n_seq]):
        with Simulation(sc['problem'], precision='fp64', coords=sc['coords']) as sim:
            sim.step(steps)
            sim.sync()
        if (k + 1) % 8 == 0:
            print(f'[bench] sequential fp64: {k + 1}/{n_seq}', file=sys.stderr, flush=True)
    seq = time.perf_counter() - t0
    seq_rate = n_seq / seq
    return {'scenarios': n_scen, 'steps_each': steps, 'arrival_rate_per_s': rate, 'max_active': max_active,
            'sizes_per_block': [list(v) for v in workloads.SERVING_SIZES], 'policy': 'sjf',
            'p50_latency_s': round(float(np.percentile(lat, 50)), 4),
            'p99_latency_s': round(float(np.percentile(lat, 99)), 4),
            'scenarios_per_s': round(n_scen / wall, 4), 'wall_s': round(wall, 3),
            'M_grid_pt_stages_per_s': round(pts / wall / 1e6, 2),
            'sequential_fp64': {'scenarios': n_seq, 'scenarios_per_s': round(seq_rate, 4)},
            'speedup_vs_sequential_fp64': round((n_scen / wall) / seq_rate, 3),
            'context_pool': pool,
            'fifo': None if lat_f is None else {'p50_latency_s': round(float(np.percentile(lat_f, 50)), 4),
                                                'p99_latency_s': round(float(np.percentile(lat_f, 99)), 4),
                                                'scenarios_per_s': round(n_scen / wall_f, 4)}}


def configure(a, scale):
    """the workload named on the command line at `scale`, with the operator,
    the anelastic response and the PML switched as asked"""
    from wql import workloads
    W = workloads.WORKLOADS[a.workload](scale=scale)
    prob = W['problem']
    if a.operator:
        prob.fd_type, order = a.operator.split(':')
        prob.order = int(order)
    if a.anelastic == 'off':
        prob.anelastic = None
    elif a.anelastic:
        prob.anelastic = tuple(float(v) for v in a.anelastic.split(','))
    if a.pml == 'off':
        for b in prob.blocks:
            b.pml, b.npml = (False,) * 6, 0
    return W


def cpu_baseline(W, min_seconds=12.0, max_steps=40):
    """the oracle (C, OpenMP over the host's cores) on full LSRK steps of the
    same workload at full size: steps until at least min_seconds of wall time"""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle as O
    from wql import workloads
    name = W['problem'].name
    sub = W
    prob = sub['problem']
    od = O.OracleDomain(prob, coords=sub['coords'])
    if sub.get('F0') is not None:
        for ib in range(2):
            od.blocks[ib]['F'][...] = sub['F0'][ib]
    if sub.get('init_state'):
        od.init_state()
    od.step(1)  # first touch of the arrays outside the timing
    nsteps = 0
    t0 = time.perf_counter()
    while nsteps < max_steps and (time.perf_counter() - t0) < min_seconds:
        od.step(1)
        nsteps += 1
    sec = (time.perf_counter() - t0) / nsteps
    nodes = workloads.points(prob)
    threads = int(os.environ.get('OMP_NUM_THREADS', '0') or 0) or os.cpu_count()
    return dict(value=nodes * 5 / sec / 1e6, unit='M grid-pt RK stages/s', cores=threads, kind='port',
                sample=f'oracle C restatement (OpenMP, {threads} threads), {nsteps} LSRK steps of the full {name} '
                       f'workload ({"x".join(str(v) for v in prob.blocks[0].n)} per block, {nodes} nodes), '
                       f'{sec * nsteps:.1f} s')


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=30)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--workload', default='rupture')
    ap.add_argument('--scale', type=float, default=1.0)
    ap.add_argument('--tile', type=int, default=None, help='main-kernel launch variant (default: tuned)')
    ap.add_argument('--zchunk', type=int, default=None)
    ap.add_argument('--order', default=None, help='tile dispatch order:group, e.g. 2:2')
    ap.add_argument('--graph', type=int, default=0)
    ap.add_argument('--operator', default=None,
                    help="spatial operator 'upwind:<order>' or 'upwind_drp:<order>' (default: the sixth-order "
                         "central one)")
    ap.add_argument('--anelastic', default=None,
                    help="anelastic response 'c,weight_exp,fref', e.g. 2,0.6,1, or 'off'")
    ap.add_argument('--pml', default=None, help="'off': the workload without its PML slabs")
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-fp64', action='store_true')
    ap.add_argument('--err-steps', type=int, default=None,
                    help="steps of the FP32-vs-FP64 slip comparison (default: the workload's own step count, "
                         "1500 for the rupture workload; 0: off)")
    ap.add_argument('--serve', type=int, default=64, help='scenarios in the serving stream (0: off)')
    ap.add_argument('--serve-steps', type=int, default=100)
    ap.add_argument('--serve-rate', type=float, default=20.0)
    ap.add_argument('--serve-active', type=int, default=8)
    ap.add_argument('--serve-seq', type=int, default=16)
    ap.add_argument('--serve-full', action='store_true',
                    help='the serving stream as stated: 256 scenarios x 1000 steps, Poisson arrivals at 20/s, '
                         'against 32 sequential plain FP64 runs (several minutes)')
    a = ap.parse_args()
    if a.gpus != 1:
        print(f'bench.py: the library runs one problem on one GPU; --gpus {a.gpus} is not supported '
              '(the two-block problem would shard one block per GPU, with the fault faces exchanged '
              'once per RK stage)', file=sys.stderr)
        sys.exit(2)
    from wql import workloads
    W = configure(a, a.scale)
    prob = W['problem']
    nodes = workloads.points(prob)
    res = {}
    for prec in (['fp32'] if a.no_fp64 else ['fp64', 'fp32']):
        order = tuple(int(v) for v in a.order.split(':')) if a.order else None
        sec, kms, dev, shape = run(W, prec, a.steps, a.warmup, a.tile, a.zchunk, a.graph, order)
        res[prec] = dict(sec=sec, kms=kms, dev=dev, shape=shape, mpts=nodes * 5 * a.steps / sec / 1e6)
    r32 = res['fp32']
    workload = f'{a.workload}: 2 blocks x {"x".join(str(v) for v in prob.blocks[0].n)}'
    if a.operator or prob.fd_type != 'traditional':
        workload += f', {prob.fd_type} order {prob.order}'
    if prob.anelastic is not None:
        workload += ', anelastic ' + ','.join(f'{v:g}' for v in prob.anelastic)
    if a.pml == 'off':
        workload += ', PML off'
    kname = kernel_name(prob, 'fp32', r32['shape'][0])
    traffic = measured_traffic(workload, r32['shape'], kname)
    bpn = bytes_per_node(prob, 'fp32')
    kn = main_kernel_nodes(prob)
    achieved = kn * bpn / (r32['kms'] * 1e-3) / 1e9 if r32['kms'] > 0 else None
    out = {
        'metric': 'M grid-pt RK stages/s FP32 vs FP64 @<=1e-5 slip err; p50/p99 scenario latency',
        'value': round(r32['mpts'], 2),
        'unit': 'M grid-pt RK stages/s',
        'n_gpus': 1,
        'steps': a.steps,
        'warmup': a.warmup,
        'ms_per_step': round(r32['sec'] / a.steps * 1e3, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32 volume / f64 fault',
        'data': 'synthetic (seeded geometry, topography and friction)',
        'config': {'workload': workload,
                   'nodes': nodes, 'tile_variant': r32['shape'][0], 'zchunk': r32['shape'][1],
                   'tile_order': r32['shape'][2], 'tile_group': r32['shape'][3]},
        'roofline': {'bound': 'hbm', 'achieved': round(achieved, 1) if achieved else None,
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(achieved / HBM_PEAK_GBS, 4) if achieved else None, 'traffic': traffic,
                     'kernel': kname, 'kernel_ms': round(r32['kms'], 5),
                     'algorithmic_bytes_per_launch': round(kn * bpn)},
    }
    if 'fp64' in res:
        out['fp64'] = {'value': round(res['fp64']['mpts'], 2),
                       'ms_per_step': round(res['fp64']['sec'] / a.steps * 1e3, 4),
                       'kernel_ms': round(res['fp64']['kms'], 5),
                       'speedup_fp32_over_fp64': round(r32['mpts'] / res['fp64']['mpts'], 3)}
    err_steps = W['steps'] if a.err_steps is None else a.err_steps
    if err_steps > 0 and not a.no_fp64:
        out['fp32_error'] = fp32_error(W, err_steps)
    if a.serve_full:
        out['serving'] = serving(256, 1000, 20.0, a.serve_active, 32, fifo=False)
    elif a.serve > 0:
        out['serving'] = serving(a.serve, a.serve_steps, a.serve_rate, a.serve_active, min(a.serve_seq, a.serve))
    if not a.no_cpu:
        out['cpu_baseline'] = cpu_baseline(W if a.scale == 1.0 else configure(a, 1.0))
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
