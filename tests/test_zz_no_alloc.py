"""wql_get_fields_noalloc (the allocation-free volume download) against
wql_get_fields through the C ABI: bit-identical fields on two calls in a
row, the second on a smaller block, in both precisions; and a block whose
fields exceed the fixed 256 MiB staging is refused with WQL_E_ARG before any
device work."""
import ctypes as C
import sys

import numpy as np
import pytest

import golden_cases as G
import gpu_helpers as H

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@pytest.mark.parametrize('prec', ['fp64', 'fp32'])
def test_noalloc_download_matches_get_fields(prec):
    from wql import native
    from wql.solver import Simulation
    L = native.lib()
    for case, n in (('curv_sw', (41, 33, 29)), ('cart_sw', (21, 17, 15))):
        _say(f'[gpu] no-alloc download {case} {n} {prec}')
        prob = G.problem(case)
        for b in prob.blocks:
            b.n = n
        with Simulation(prob, precision=prec) as sim:
            H.load_pattern_state(sim, prob)
            sim.step(2)
            sim.sync()
            for ib in range(2):
                want = sim.fields(ib)
                got = np.full_like(want, np.nan)
                native.check(L.wql_get_fields_noalloc(sim.ctx, ib, _dp(got)))
                assert np.array_equal(got, want), (case, ib)


def test_noalloc_download_refuses_oversized_block():
    from wql import native
    from wql.solver import Simulation
    L = native.lib()
    prob = G.problem('cart_sw')
    for b in prob.blocks:
        b.n = (161, 161, 161)   # 4.17 M nodes: 300 MB of FP64 fields
    _say('[gpu] no-alloc download refuses 161^3')
    with Simulation(prob, precision='fp32') as sim:
        out = np.empty((9, 161, 161, 161))
        assert L.wql_get_fields_noalloc(sim.ctx, 0, _dp(out)) == native.WQL_E_ARG
        assert b'allocation-free' in L.wql_last_error()
