"""Serving loop (continuous batching): every scenario's result is the same
alone and in the mix.  Kernels are deterministic and scenarios share no
device state, so the bar is bit-for-bit equality of slip, rupture time and
peak slip rate."""
import sys

import numpy as np
import pytest

from wql import workloads as Wk
from wql.problem import Problem

pytestmark = pytest.mark.gpu


def _scenario(k):
    rng = np.random.default_rng([11, k])
    sizes = [(25, 21, 21), (41, 25, 23), (33, 31, 29), (49, 21, 25)]
    n = sizes[k % len(sizes)]
    h = 0.1
    blocks, coords = Wk._blocks(n, h, mesh='cartesian')
    Ly, Lz = (n[1] - 1) * h, (n[2] - 1) * h
    mu_s = rng.uniform(0.6, 0.7)
    f = Wk.sw_fault(mu_s, rng.uniform(0.45, 0.55), rng.uniform(0.2, 0.6), yc=0.5 * Ly, zc=0.0, depth=Ly, width=Lz)
    f.patches = [(0.5 * Ly, 0.0, 0.6, 0.6, 120.0 * mu_s - 70.0 + 0.5)]
    return dict(problem=Problem(blocks=blocks, fault=f, cfl=0.5, name=f's{k}'), coords=coords, steps=23 + 4 * k)


def _alone(sc, precision):
    from wql.solver import Simulation
    with Simulation(sc['problem'], precision=precision, coords=sc['coords']) as sim:
        sim.step(sc['steps'])
        sim.sync()
        f = sim.fault(0)
        return dict(S=f['S'], trup=f['trup'][0], psr=sim.fault_peak(0)[0], stats=sim.fault_stats(0))


@pytest.mark.parametrize('policy', ['fifo', 'sjf'])
@pytest.mark.parametrize('precision', ['fp32', 'fp64'])
def test_alone_equals_in_the_mix(precision, policy):
    from wql.serving import run_stream
    print(f'[gpu] serving mix {precision} {policy}', file=sys.stderr, flush=True)
    scen = [_scenario(k) for k in range(6)]
    arrivals = np.array([0.0, 0.0, 0.01, 0.02, 0.02, 0.05])
    pool = {}
    lat, wall, res = run_stream(scen, arrivals, max_active=3, steps_per_launch=7, precision=precision,
                                keep_results=True, policy=policy, stats=pool)
    assert len(res) == len(scen)
    # scenarios k and k + 4 share their blocks: the later one runs on the
    # earlier one's pooled context whenever that one retired first
    assert pool['created'] + pool['reused'] == len(scen) and pool['reused'] >= 1, pool
    assert (lat > 0).all()
    for k, sc in enumerate(scen):
        a = _alone(sc, precision)
        r = res[k]
        assert np.array_equal(r['S'], a['S']), k
        assert np.array_equal(r['trup'], a['trup']), k
        assert np.array_equal(r['psr'], a['psr']), k
        assert r['stats'] == a['stats'], k
        assert a['stats']['max_slip'] > 0.0, k  # the nucleation patch ruptures


def test_admission_respects_the_budget():
    """a budget that fits one scenario at a time still serves them all, one by one"""
    from wql.serving import Server
    from wql.solver import estimate_bytes
    print('[gpu] serving budget', file=sys.stderr, flush=True)
    scen = [_scenario(2) for _ in range(3)]
    need = estimate_bytes(scen[0]['problem'], 'fp32')
    with Server(max_active=3, hbm_budget=int(need * 1.5), steps_per_launch=5) as srv:
        ids = [srv.submit(s['problem'], s['coords'], s['steps'], 'fp32') for s in scen]
        done = []
        peak_active = 0
        while len(done) < len(ids):
            done += srv.poll()
            st = srv.status()
            peak_active = max(peak_active, st['active'])
            assert st['bytes_used'] <= st['budget']
        assert peak_active == 1
        for sid in ids:
            assert srv.result(sid)['stats']['max_slip'] > 0.0


def test_shortest_job_first_order():
    """with one active slot and every scenario queued before the first poll,
    the shortest-job-first policy runs them in increasing nodes x steps
    (submitted largest first, so arrival order would be the reverse)"""
    from wql.serving import Server
    print('[gpu] serving sjf order', file=sys.stderr, flush=True)
    scen = [_scenario(k) for k in (2, 1, 0)]
    work = [2 * np.prod(s['problem'].blocks[0].n) * s['steps'] for s in scen]
    assert work[0] > 1.4 * work[1] > 1.4 * 1.4 * work[2]
    with Server(max_active=1, steps_per_launch=5, policy='sjf') as srv:
        ids = [srv.submit(s['problem'], s['coords'], s['steps'], 'fp32') for s in scen]
        done = []
        while len(done) < len(ids):
            done += srv.poll()
        assert done == ids[::-1]
        for sid in ids:
            assert srv.result(sid)['stats']['max_slip'] > 0.0


def test_poll_reports_every_retirement_with_a_small_buffer():
    """poll with room for one id per call: every retired scenario is still
    reported, once, in a later call"""
    import ctypes as C
    from wql import native
    from wql.serving import Server
    print('[gpu] serving small poll buffer', file=sys.stderr, flush=True)
    scen = [_scenario(k) for k in (0, 0, 0, 1)]
    with Server(max_active=4, steps_per_launch=50) as srv:
        ids = [srv.submit(s['problem'], s['coords'], s['steps'], 'fp32') for s in scen]
        done = []
        one = (C.c_int * 1)()
        for _ in range(200000):
            n = native.lib().wql_server_poll(srv.srv, one, 1)
            assert n in (0, 1)
            if n:
                done.append(one[0])
            if len(done) == len(ids):
                break
        assert sorted(done) == sorted(ids)
        for sid in ids:
            assert srv.result(sid)['stats']['max_slip'] > 0.0


def test_pool_frees_least_recently_used_for_new_blocks():
    """a budget that holds one context at a time: a scenario with other
    blocks frees the pooled context, and the results are unchanged"""
    from wql.serving import Server
    from wql.solver import estimate_bytes
    print('[gpu] serving pool eviction', file=sys.stderr, flush=True)
    a, b = _scenario(0), _scenario(2)
    need = max(estimate_bytes(s['problem'], 'fp32') for s in (a, b))
    with Server(max_active=1, hbm_budget=int(need * 1.2), steps_per_launch=10) as srv:
        ids = [srv.submit(s['problem'], s['coords'], s['steps'], 'fp32') for s in (a, b, a, a)]
        done = []
        while len(done) < len(ids):
            done += srv.poll()
        st = srv.pool_stats()
        assert (st['created'], st['reused'], st['freed']) == (3, 1, 2), st
        res = [srv.result(sid) for sid in ids]
    ref = _alone(a, 'fp32')
    for r in (res[0], res[2], res[3]):
        assert np.array_equal(r['S'], ref['S'])


def test_tuned_shape_registered(tmp_path):
    """the autotuner's per-size launch shape reaches the server: it is
    registered at the first submission of that size, bad shapes are refused,
    and the scenario's result equals the same scenario run alone with that
    launch shape"""
    import ctypes as C
    import json

    from wql import autotune, native
    from wql.serving import Server, run_stream
    from wql.solver import Simulation
    print('[gpu] serving tuned shape', file=sys.stderr, flush=True)
    sc = _scenario(1)
    cache = tmp_path / 'tune.json'
    entry = {'variant': 6, 'zchunk': 4, 'order': 0, 'group': 1}
    cache.write_text(json.dumps({autotune.key(sc['problem'], 'fp32'): entry}))
    with Server(tune_cache=str(cache)) as srv:
        assert srv.register_shape(sc['problem'], 'fp32') == entry
        assert srv.register_shape(_scenario(0)['problem'], 'fp32') is None  # size not in the cache
        n = (C.c_int * 3)(*sc['problem'].blocks[0].n)
        assert native.lib().wql_server_set_shape(srv.srv, n, native.WQL_FP32, 1, 99, 4, 0, 1) == native.WQL_E_ARG
    _, _, res = run_stream([sc], np.zeros(1), max_active=1, steps_per_launch=7, precision='fp32',
                           keep_results=True, tune_cache=str(cache))
    with Simulation(sc['problem'], precision='fp32', coords=sc['coords'], tile=6, zchunk=4) as sim:
        sim.step(sc['steps'])
        sim.sync()
        f = sim.fault(0)
        assert np.array_equal(res[0]['S'], f['S'])
        assert np.array_equal(res[0]['trup'], f['trup'][0])
