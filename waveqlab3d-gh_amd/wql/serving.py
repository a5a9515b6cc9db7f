"""Continuous batching of independent rupture scenarios on one GPU.

`Server` wraps the library's serving loop (wql_server_* in include/wql.h):
scenarios are admitted by device footprint as memory and slots free up,
each keeps its state on the device and runs on its own stream with a
captured step graph, and finished scenarios are retired so the next can be
admitted.  `run_stream` drives a request stream with given arrival times
and reports per-scenario latency (arrival -> result on the host).
"""
import ctypes as C
import time

import numpy as np

from . import native
from .native import check
from .solver import block_descs, problem_desc


class Server:
    POLICIES = {'fifo': 0, 'sjf': 1}

    def __init__(self, max_active=8, hbm_budget=0, steps_per_launch=50, policy='fifo', tuned=True, tune_cache=None):
        """policy: 'fifo' admits queued scenarios in arrival order, 'sjf'
        shortest job first with aging (nodes x steps / (1 + seconds waited));
        tuned: scenarios whose block size the autotuner's cache (the file
        tune_cache, default the package's) holds run that launch shape"""
        self.srv = C.c_void_p()
        check(native.lib().wql_server_create(max_active, hbm_budget, steps_per_launch, C.byref(self.srv)))
        check(native.lib().wql_server_set_policy(self.srv, self.POLICIES[policy]))
        self.shape = {}
        self.tuned = tuned
        self.tune_cache = tune_cache
        self.registered = {}
        self._done = (C.c_int * 4096)()

    def register_shape(self, prob, precision):
        """hand the autotuner's launch shape for this block size to the
        server (once per size, precision and grid kind); returns it or None"""
        from . import autotune
        b = prob.blocks[0]
        cart = 1 if b.mesh == 'cartesian' else 0
        k = (tuple(b.n), precision, cart)
        if k not in self.registered:
            best = autotune.lookup(prob, precision, self.tune_cache)
            if best is not None:
                n = (C.c_int * 3)(*b.n)
                prec = native.WQL_FP32 if precision == 'fp32' else native.WQL_FP64
                check(native.lib().wql_server_set_shape(self.srv, n, prec, cart, best['variant'], best['zchunk'],
                                                        best.get('order', 0), best.get('group', 2)))
            self.registered[k] = best
        return self.registered[k]

    def close(self):
        if self.srv:
            native.lib().wql_server_destroy(self.srv)
            self.srv = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, prob, coords, nsteps, precision='fp32'):
        if self.tuned:
            self.register_shape(prob, precision)
        pd = problem_desc(prob, precision)
        arr, keep = block_descs(prob, coords)
        sid = C.c_int()
        check(native.lib().wql_server_submit(self.srv, C.byref(pd), arr, nsteps, C.byref(sid)))
        self.shape[sid.value] = (prob.blocks[0].n[1], prob.blocks[0].n[2])
        return sid.value

    def poll(self):
        """ids of scenarios retired by this call (never blocks on the device
        except to download a retired scenario's fault arrays)"""
        n = native.lib().wql_server_poll(self.srv, self._done, len(self._done))
        if n < 0:
            check(n)
        return [self._done[k] for k in range(n)]

    def result(self, sid):
        nr, ns = self.shape.pop(sid)
        S = np.empty((4, ns, nr))
        trup = np.empty((ns, nr))
        psr = np.empty((ns, nr))
        stats = np.empty(6)
        p = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        check(native.lib().wql_server_result(self.srv, sid, p(S), p(trup), p(psr), p(stats)))
        keys = ('max_slip', 'max_slip_rate', 'max_peak_slip_rate', 'ruptured_nodes', 'mean_slip_ruptured',
                'first_rupture_time')
        return dict(S=S, trup=trup, psr=psr, stats=dict(zip(keys, stats.tolist())))

    def pool_stats(self):
        out = (C.c_long * 4)()
        check(native.lib().wql_server_pool_stats(self.srv, out))
        return dict(created=out[0], reused=out[1], freed=out[2], pooled=out[3])

    def status(self):
        out = (C.c_long * 4)()
        check(native.lib().wql_server_status(self.srv, out))
        return dict(queued=out[0], active=out[1], bytes_used=out[2], budget=out[3])


def run_stream(scenarios, arrivals, max_active=8, steps_per_launch=50, precision='fp32', keep_results=False,
               idle_sleep=2e-4, policy='fifo', stats=None, progress=None, tuned=True, tune_cache=None):
    """Serve `scenarios` (list of dicts with problem, coords, steps) arriving
    at `arrivals` seconds after the start.  Returns per-scenario latency (s),
    the wall time from first arrival to last result, and optionally results;
    `stats` (a dict) receives the context pool's counters; `progress(n)`
    is called after each retired scenario with the number retired; tuned /
    tune_cache: as for Server."""
    order = np.argsort(arrivals)
    lat = np.zeros(len(scenarios))
    results = {}
    t_start = time.perf_counter()
    sub_t = {}
    nxt = 0
    ndone = 0
    with Server(max_active=max_active, steps_per_launch=steps_per_launch, policy=policy, tuned=tuned,
                tune_cache=tune_cache) as srv:
        while ndone < len(scenarios):
            now = time.perf_counter() - t_start
            while nxt < len(order) and arrivals[order[nxt]] <= now:
                k = int(order[nxt])
                sc = scenarios[k]
                sid = srv.submit(sc['problem'], sc['coords'], sc['steps'], precision)
                sub_t[sid] = (k, arrivals[k])
                nxt += 1
            finished = srv.poll()
            for sid in finished:
                k, ta = sub_t.pop(sid)
                r = srv.result(sid)
                lat[k] = time.perf_counter() - t_start - ta
                ndone += 1
                if progress is not None:
                    progress(ndone)
                if keep_results:
                    results[k] = r
            if not finished and (nxt >= len(order) or arrivals[order[nxt]] > now):
                time.sleep(idle_sleep)
        wall = time.perf_counter() - t_start - float(np.min(arrivals))
        if stats is not None:
            stats.update(srv.pool_stats())
    return lat, wall, results
