"""The anelastic response (response = 'anelastic', &anelastic_list) on the
GPU through the C ABI: four GSLS mechanisms per stress component whose
memory variables ride in every volume kernel of the per-node path.  Rates
and memory-variable rates against the reference build's for the same
pattern state (the sixth-order operator on the curvilinear pair, with PML
slabs, and the DRP-6 operator with PML as in the reference's PML-12 runs);
the trajectories are in test_gpu_parity's golden trajectory test; FP32
against FP64."""
import sys

import numpy as np
import pytest

import golden_cases as G
import gpu_helpers as H

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def _sim(prob, **kw):
    from wql.solver import Simulation
    return Simulation(prob, **kw)


@pytest.mark.parametrize('case', G.AN_CASES)
def test_anelastic_rates_vs_golden(case):
    """DF with the volume, SAT and fault terms, and the memory-variable rates
    Deta, from the pattern state with patterned memory variables"""
    _say(f'[gpu] anelastic rhs {case}')
    z = G.load(case)
    prob = G.problem(case)
    assert prob.anelastic is not None
    with _sim(prob) as sim:
        H.load_pattern_state(sim, prob)
        sim.eval_rates(float(z['rhs_t']), what=1, stage=2)
        for ib in range(2):
            assert G.rel_l2(sim.rates(ib), z[f'b{ib + 1}_DFfull']) <= TOL
            assert G.rel_l2(sim.anelastic(ib, what=1), z[f'b{ib + 1}_Deta']) <= TOL
            # rates only: the memory variables themselves are unchanged
            assert np.array_equal(sim.anelastic(ib, what=0), G.pattern_state(prob)[ib]['eta'])


@pytest.mark.parametrize('case', ['an_curv', 'an_drp6_pml'])
def test_anelastic_fp32_tracks_fp64(case):
    """FP32 volume fields and memory variables within 1e-5 of FP64"""
    _say(f'[gpu] anelastic fp32 {case}')
    z = G.load(case)
    prob = G.problem(case)
    args = z['traj_args']
    out = {}
    for prec in ('fp64', 'fp32'):
        with _sim(prob, precision=prec) as sim:
            F0 = H.pulse_init(prob, args)
            for ib in range(2):
                sim.set_fields(ib, F0[ib])
            sim.init_state()
            sim.step(int(args[0]))
            sim.sync()
            out[prec] = [(sim.fields(ib), sim.anelastic(ib)) for ib in range(2)]
    for ib in range(2):
        assert G.rel_l2(out['fp32'][ib][0], out['fp64'][ib][0]) <= 1e-5
        assert G.rel_l2(out['fp32'][ib][1], out['fp64'][ib][1]) <= 1e-5
