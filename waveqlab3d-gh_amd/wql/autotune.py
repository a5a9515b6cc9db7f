"""Launch-shape autotuner for the main volume kernel.

Sweeps the launch variants of the interior volume kernel
(wql_tile_variants(): fused march on 32x8 / 64x8 / 32x16 tiles re-reading or
holding the centre plane, and the role-split march) and the s-chunk length
for one problem shape, timing the kernel with HIP events on the library
stream, and caches the fastest choice per (block size, precision, grid kind)
in tune_cache.json beside this module.  Simulation picks the cached shape up
when none is given, and the serving loop registers it with the server for
every scenario of that size (Server.register_shape).

    python -m wql.autotune [--workloads rupture,rough,...] [--serving] [--precisions fp32,fp64] [--cache FILE]

tunes the benchmark workloads and the serving block sizes and writes the
cache.
"""
import json
import os
import sys

from . import native
from .solver import Simulation

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'tune_cache.json')
ZCHUNKS = (4, 6, 8, 12, 16, 24)


def key(prob, precision):
    b = prob.blocks[0]
    kind = ('curv' if b.mesh != 'cartesian' else 'cart') + ('_het' if b.material is not None or b.medium is not None else '') + \
           ('_pml' if any(b.pml) else '')
    return f'{"x".join(map(str, b.n))}:{precision}:{kind}'


def load_cache(path=None):
    try:
        with open(CACHE if path is None else path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(prob, precision, path=None):
    return load_cache(path).get(key(prob, precision))


def tune(prob, precision='fp32', coords=None, steps=2, variants=None, zchunks=ZCHUNKS, orders=((0, 1),),
         log=None, cache=CACHE):
    """times every (variant, zchunk, (order, group)) launch shape of the main
    volume kernel; returns ((variant, zchunk, order, group, ms), results) of
    the fastest and records it in the cache file `cache` (None: not kept)"""
    L = native.lib()
    variants = range(L.wql_tile_variants()) if variants is None else variants
    best = None
    results = []
    with Simulation(prob, precision=precision, coords=coords) as sim:
        for v in variants:
            for zc in zchunks:
                for order, group in orders:
                    native.check(L.wql_set_tile(sim.ctx, v, zc))
                    native.check(L.wql_set_tile_order(sim.ctx, order, group))
                    sim.step(1)
                    sim.profile(True)
                    sim.step(steps)
                    sim.sync()
                    ms = sim.volume_kernel_ms()
                    sim.profile(False)
                    results.append((v, zc, order, group, ms))
                    if log:
                        log(f'tile variant {v} zchunk {zc} order {order} group {group}: {ms:.4f} ms')
                    if best is None or ms < best[4]:
                        best = (v, zc, order, group, ms)
    if cache is not None:
        entries = load_cache(cache)
        entries[key(prob, precision)] = {'variant': best[0], 'zchunk': best[1], 'order': best[2], 'group': best[3],
                                         'ms': round(best[4], 5)}
        with open(cache, 'w') as f:
            json.dump(entries, f, indent=1, sort_keys=True)
    return best, results


def _main(argv):
    import argparse
    from . import workloads as Wk
    ap = argparse.ArgumentParser(prog='python -m wql.autotune')
    ap.add_argument('--workloads', default='plain,rupture,rough,plastic')
    ap.add_argument('--serving', action='store_true', help='also the serving block sizes')
    ap.add_argument('--precisions', default='fp32,fp64')
    ap.add_argument('--steps', type=int, default=2)
    ap.add_argument('--cache', default=CACHE, help='cache file to update (default: the package cache)')
    a = ap.parse_args(argv)
    jobs = []
    for w in [x for x in a.workloads.split(',') if x]:
        W = Wk.WORKLOADS[w](scale=1.0)
        jobs.append((w, W['problem'], W['coords']))
    if a.serving:
        seen = set()
        for k in range(64):
            sc = Wk.scenario(k)
            n = tuple(sc['problem'].blocks[0].n)
            if n not in seen:
                seen.add(n)
                jobs.append((f'serving {"x".join(map(str, n))}', sc['problem'], sc['coords']))
            if len(seen) == len(Wk.SERVING_SIZES):
                break
    for name, prob, coords in jobs:
        for prec in a.precisions.split(','):
            best, _ = tune(prob, prec, coords=coords, steps=a.steps, cache=a.cache)
            print(f'[tune] {name} {prec}: variant {best[0]} zchunk {best[1]} order {best[2]} group {best[3]} '
                  f'{best[4]:.4f} ms', flush=True)


if __name__ == '__main__':
    _main(sys.argv[1:])
