"""The decks of the reference's PML and anelasticity study on the HIP path
against the oracle: the sixth-order central, upwind-6 and upwind-DRP-6
operators, each with and without the 12-node PML and with and without the
anelastic response (Q_s = 2 v_s), on two Cartesian blocks joined by a locked
interface and driven by an order-4 Cos-Bell moment-tensor source
(simulation/PML-12/input/*.in; the same h = 0.2 km, node counts scaled to
0.35).  30 LSRK steps; bars: FP64 relative L2 <= 1e-12 on the fields (and
the memory variables) of both blocks, FP32 <= 1e-5."""
import sys

import numpy as np
import pytest

import golden_cases as G
import oracle as O

pytestmark = pytest.mark.gpu

STEPS = 30
SCALE = 0.35
DECKS = [(fd, pml, an) for fd in ('traditional', 'upwind', 'upwind_drp') for pml in (True, False)
         for an in (True, False)]
IDS = [f'{fd}_6_pml-{"on" if pml else "off"}_{"anelastic-c2" if an else "elastic"}' for fd, pml, an in DECKS]
_oracle = {}


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def _workload(deck):
    from wql import workloads as Wk
    fd, pml, an = deck
    return Wk.pml12(scale=SCALE, fd_type=fd, pml=pml, anelastic=an)


def _run_oracle(deck):
    if deck not in _oracle:
        W = _workload(deck)
        od = O.OracleDomain(W['problem'], coords=W['coords'])
        od.step(STEPS)
        an = W['problem'].anelastic is not None
        _oracle[deck] = dict(F=[od.blocks[ib]['F'].copy() for ib in range(2)],
                             eta=[od.blocks[ib]['eta'].copy() if an else None for ib in range(2)], t=od.t)
    return _oracle[deck]


def test_workload_is_the_deck():
    """the restated decks: sizes, PML and source as the study's files give
    them, the source in block 1's list and inside block 1"""
    from wql import workloads as Wk
    W = _workload(DECKS[0])
    p = W['problem']
    assert [b.n for b in p.blocks] == [(27, 36, 53)] * 2
    assert p.blocks[0].npml == 12 and p.blocks[0].pml == (True, False, False, True, True, True)
    assert p.blocks[1].pml == (False, True, False, True, True, True)
    x = p.moment['lists']['U'][0]['x']
    assert all(p.blocks[0].a[d] < x[d] < p.blocks[0].b[d] for d in range(3))
    full = Wk.pml12()['problem']
    assert [b.n for b in full.blocks] == [(76, 101, 151)] * 2
    assert full.blocks[1].a == (15.0, 0.0, 0.0) and full.blocks[1].b == (30.0, 20.0, 30.0)
    assert full.nt() == 554


@pytest.mark.parametrize('prec', ['fp64', 'fp32'])
@pytest.mark.parametrize('deck', DECKS, ids=IDS)
def test_pml12_deck_vs_oracle(deck, prec):
    from wql.solver import Simulation
    _say(f'[gpu] pml12 {deck} {prec}')
    ref = _run_oracle(deck)
    W = _workload(deck)
    prob = W['problem']
    with Simulation(prob, precision=prec, coords=W['coords']) as sim:
        sim.step(STEPS)
        sim.sync()
        assert sim.t == pytest.approx(ref['t'], rel=1e-14)
        F = [sim.fields(ib) for ib in range(2)]
        eta = [sim.anelastic(ib) if prob.anelastic else None for ib in range(2)]
    tol = 1e-12 if prec == 'fp64' else 1e-5
    for ib in range(2):
        assert np.abs(ref['F'][ib]).max() > 0
        assert G.rel_l2(F[ib], ref['F'][ib]) <= tol, f'F block {ib + 1}'
        if eta[ib] is not None:
            assert G.rel_l2(eta[ib], ref['eta'][ib]) <= tol * 10, f'eta block {ib + 1}'
