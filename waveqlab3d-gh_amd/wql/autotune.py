"""Launch-shape autotuner for the main volume kernel.

Sweeps the tile variants of k_volume_main (wql_tile_variants()) and the
s-chunk length for one problem shape, timing the kernel with HIP events on
the library stream, and caches the fastest choice per (shape, precision,
geometry kind) in a JSON file next to the library.
"""
import json
import os

from . import native
from .solver import Simulation

CACHE = os.path.join(os.path.dirname(native.LIB_PATH), 'tune_cache.json')
ZCHUNKS = (4, 8, 12, 16, 24, 32)


def key(prob, precision):
    b = prob.blocks[0]
    kind = ('curv' if b.mesh != 'cartesian' else 'cart') + ('_het' if b.material is not None else '') + \
           ('_pml' if any(b.pml) else '')
    return f'{"x".join(map(str, b.n))}:{precision}:{kind}'


def load_cache():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(prob, precision):
    return load_cache().get(key(prob, precision))


def tune(prob, precision='fp32', coords=None, steps=2, variants=None, zchunks=ZCHUNKS, orders=((0, 1),),
         log=None):
    """times every (variant, zchunk, (order, group)) launch shape of the main
    volume kernel; returns ((variant, zchunk, order, group, ms), results) of
    the fastest and caches it"""
    L = native.lib()
    variants = range(L.wql_tile_variants() // 2) if variants is None else variants
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
    cache = load_cache()
    cache[key(prob, precision)] = {'variant': best[0], 'zchunk': best[1], 'order': best[2], 'group': best[3],
                                   'ms': best[4]}
    try:
        with open(CACHE, 'w') as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    except OSError:
        pass
    return best, results
