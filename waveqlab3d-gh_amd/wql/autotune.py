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


def tune(prob, precision='fp32', coords=None, steps=2, variants=None, zchunks=ZCHUNKS, log=None):
    """returns (variant, zchunk, ms) of the fastest main-kernel launch shape"""
    variants = range(native.lib().wql_tile_variants()) if variants is None else variants
    best = None
    results = []
    with Simulation(prob, precision=precision, coords=coords) as sim:
        for v in variants:
            for zc in zchunks:
                native.check(native.lib().wql_set_tile(sim.ctx, v, zc))
                sim.step(1)
                sim.profile(True)
                sim.step(steps)
                sim.sync()
                ms = sim.volume_kernel_ms()
                sim.profile(False)
                results.append((v, zc, ms))
                if log:
                    log(f'tile variant {v} zchunk {zc}: {ms:.4f} ms')
                if best is None or ms < best[2]:
                    best = (v, zc, ms)
    cache = load_cache()
    cache[key(prob, precision)] = {'variant': best[0], 'zchunk': best[1], 'ms': best[2]}
    try:
        with open(CACHE, 'w') as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    except OSError:
        pass
    return best, results
