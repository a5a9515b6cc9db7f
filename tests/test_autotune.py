"""The committed autotuner cache (wql/tune_cache.json): one launch shape per
benchmark workload and serving block size in both precisions, each a shape
the library accepts; the cache key and lookup (CPU only)."""
import json

import pytest

from wql import autotune, native
from wql import workloads as Wk


def _cache():
    with open(autotune.CACHE) as f:
        return json.load(f)


def _serving_problems():
    seen = {}
    for k in range(64):
        sc = Wk.scenario(k)
        seen.setdefault(tuple(sc['problem'].blocks[0].n), sc['problem'])
    return list(seen.values())


@pytest.mark.parametrize('prec', ['fp32', 'fp64'])
def test_cache_covers_workloads_and_serving_sizes(prec):
    cache = _cache()
    probs = [Wk.WORKLOADS[w](scale=1.0)['problem'] for w in ('plain', 'rupture', 'rough', 'plastic')]
    probs += _serving_problems()
    assert len(probs) == 4 + len(Wk.SERVING_SIZES)
    nv = native.lib().wql_tile_variants()
    for p in probs:
        e = cache.get(autotune.key(p, prec))
        assert e is not None, autotune.key(p, prec)
        assert 0 <= e['variant'] < nv and e['zchunk'] >= 1 and e['ms'] > 0
        assert autotune.lookup(p, prec) == e


def test_key_distinguishes_grid_kind():
    r = Wk.WORKLOADS['rupture'](scale=0.2)['problem']
    pl = Wk.WORKLOADS['plastic'](scale=0.2)['problem']
    assert autotune.key(r, 'fp32').endswith(':curv')
    assert autotune.key(pl, 'fp32').endswith(':cart_het')
    assert autotune.key(r, 'fp32') != autotune.key(r, 'fp64')
