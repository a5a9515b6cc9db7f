"""The slip-weakening problems the reference keys by name -- TPV10, TPV29/30,
TPV31-34, SCITS2016, TPV36/37, wasatch100 (prestress, strength, cohesion and
forced rupture tabled per fault node by k_fault_tables; Interface_Condition.
f90:1268-2660, initial_stress_condition.f90:87-244) -- and init_material's
layered media (LOH1, OKLAHOMA, TPV31-33; material.f90:179-690) on the GPU
against the reference build: rates with the SAT and fault terms for the
pattern state, and trajectories from rest or a velocity pulse (TPV30 with its
plasticity and TPV29/30 background stress)."""
import sys

import numpy as np
import pytest

import golden_cases as G
import gpu_helpers as H

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize('case', G.KEY_RHS_CASES)
def test_keyed_rhs_vs_golden(case):
    from wql.solver import Simulation
    _say(f'[gpu] keyed rhs {case}')
    z = G.load(case)
    prob = G.problem(case)
    assert prob.fault.key != 0
    t = float(z['rhs_t'])
    with Simulation(prob) as sim:
        H.load_pattern_state(sim, prob)
        sim.eval_rates(t, what=1, stage=2)
        for ib in range(2):
            assert G.rel_l2(sim.rates(ib), z[f'b{ib + 1}_DFfull']) <= TOL, f'DF block {ib + 1}'
            DS, DW = sim.fault_rates(ib)
            ref = z[f'b{ib + 1}_DS']
            assert np.abs(DS - ref).max() <= TOL * max(1.0, np.abs(ref).max()), f'DS block {ib + 1}'


@pytest.mark.parametrize('prec', ['fp64', 'fp32'])
@pytest.mark.parametrize('case', G.KEY_TRAJ_CASES)
def test_keyed_trajectory_vs_golden(case, prec):
    """FP64 at 1e-12 (slip, state, slip rate and rupture time at 1e-10 of
    their scale); the FP32 volume with the FP64 fault within 1e-5"""
    from wql.solver import Simulation
    _say(f'[gpu] keyed traj {case} {prec}')
    z = G.load(case)
    prob = G.problem(case)
    args = z['traj_args']
    tol = TOL if prec == 'fp64' else 1e-5
    with Simulation(prob, precision=prec) as sim:
        assert sim.dt == pytest.approx(float(z['dt']), rel=1e-15)
        F0 = H.pulse_init(prob, args)
        for ib in range(2):
            sim.set_fields(ib, F0[ib])
        sim.init_state()
        sim.step(int(args[0]))
        sim.sync()
        assert sim.t == pytest.approx(float(z['t_end']), rel=1e-14)
        for ib in range(2):
            assert G.rel_l2(sim.fields(ib), z[f'b{ib + 1}_F']) <= tol, f'F block {ib + 1}'
            f = sim.fault(ib)
            for key in ('S', 'W', 'Svel'):
                ref = z[f'b{ib + 1}_{key}']
                assert np.abs(f[key] - ref).max() <= max(1e-10, tol) * max(1.0, np.abs(ref).max()), key
            if prec == 'fp64':
                ref = z[f'b{ib + 1}_trup']
                assert np.abs(f['trup'] - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max()), 'trup'


@pytest.mark.parametrize('case', G.MEDIUM_CASES)
def test_keyed_medium_on_the_device(case):
    """the library's view of the layered medium: the material the reference
    uses at every node"""
    from wql.solver import Simulation
    _say(f'[gpu] medium {case}')
    z = G.load(case)
    prob = G.problem(case)
    for ib, b in enumerate(prob.blocks):
        ref = z[f'b{ib + 1}_Mat']
        M = b.medium_material()
        if ref.ndim == 1:
            ref = np.broadcast_to(ref[:, None, None, None], M.shape)
        assert np.array_equal(M, ref), ib
    with Simulation(prob) as sim:
        sim.step(1)
        sim.sync()


@pytest.mark.parametrize('case', G.MATFILE_CASES)
def test_material_file_case(case, tmp_path):
    """a medium read from files (material_source = 'file'): rates of the
    pattern state with every term, and the FP64 trajectory, against the
    reference build"""
    from wql.solver import Simulation
    _say(f'[gpu] material file {case}')
    z = G.load(case)
    prob, _ = G.matfile_problem(case, str(tmp_path))
    with Simulation(prob) as sim:
        H.load_pattern_state(sim, prob)
        sim.eval_rates(float(z['rhs_t']), what=1, stage=2)
        for ib in range(2):
            assert G.rel_l2(sim.rates(ib), z[f'b{ib + 1}_DFfull']) <= TOL, f'DF block {ib + 1}'
    args = z['traj_args']
    with Simulation(prob) as sim:
        F0 = H.pulse_init(prob, args)
        for ib in range(2):
            sim.set_fields(ib, F0[ib])
        sim.init_state()
        sim.step(int(args[0]))
        sim.sync()
        for ib in range(2):
            assert G.rel_l2(sim.fields(ib), z[f'b{ib + 1}_F']) <= TOL, f'F block {ib + 1}'
            f = sim.fault(ib)
            ref = z[f'b{ib + 1}_S']
            assert np.abs(f['S'] - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max())
