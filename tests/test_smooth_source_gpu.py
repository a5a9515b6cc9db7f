"""The Gaussian-mollified moment-tensor sources (mollify_source = T,
set_moment_tensor_smooth, moment_tensor.f90:1400-1475) on the GPU against
the reference build: the rates of the pattern state with the volume terms
and with every term, and trajectories from rest driven by the sources alone
(the unified list on the curvilinear grid, where both blocks spread every
source, and the per-block lists on the Cartesian grid)."""
import sys

import numpy as np
import pytest

import golden_cases as G
import gpu_helpers as H

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize('case', G.SMOOTH_CASES)
def test_smooth_source_rhs_vs_golden(case):
    from wql.solver import Simulation
    _say(f'[gpu] smooth source rhs {case}')
    z = G.load(case)
    prob = G.problem(case)
    assert prob.mollify_source
    t = float(z['rhs_t'])
    with Simulation(prob) as sim:
        H.load_pattern_state(sim, prob)
        sim.eval_rates(t, what=0)
        for ib in range(2):
            assert G.rel_l2(sim.rates(ib), z[f'b{ib + 1}_DFvol']) <= TOL, f'DFvol block {ib + 1}'
        H.load_pattern_state(sim, prob)
        sim.eval_rates(t, what=1, stage=2)
        for ib in range(2):
            assert G.rel_l2(sim.rates(ib), z[f'b{ib + 1}_DFfull']) <= TOL, f'DFfull block {ib + 1}'


@pytest.mark.parametrize('prec', ['fp64', 'fp32'])
@pytest.mark.parametrize('case', G.SMOOTH_CASES)
def test_smooth_source_trajectory_vs_golden(case, prec):
    """FP64 at 1e-12, the FP32 volume within 1e-5"""
    from wql.solver import Simulation
    _say(f'[gpu] smooth source traj {case} {prec}')
    z = G.load(case)
    prob = G.problem(case)
    args = z['traj_args']
    tol = TOL if prec == 'fp64' else 1e-5
    with Simulation(prob, precision=prec) as sim:
        F0 = H.pulse_init(prob, args)
        for ib in range(2):
            sim.set_fields(ib, F0[ib])
        sim.init_state()
        sim.step(int(args[0]))
        sim.sync()
        assert sim.t == pytest.approx(float(z['t_end']), rel=1e-14)
        for ib in range(2):
            ref = z[f'b{ib + 1}_F']
            assert np.abs(ref).max() > 0
            assert G.rel_l2(sim.fields(ib), ref) <= tol, f'F block {ib + 1}'
