"""The upwind and upwind-DRP operators (&problem_list fd_type / order) on the
GPU, through the C ABI: the per-node table kernel (k_volume_op) against the
reference build's rates and trajectories (seven operators), against the
oracle for all fifteen operators on a curvilinear rate-and-state problem
and a Cartesian one with PML slabs, and in FP32 against FP64."""
import copy
import sys

import numpy as np
import pytest

import golden_cases as G
import gpu_helpers as H
import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12
OPS = [('upwind', o) for o in range(2, 10)] + [('upwind_drp', o) for o in (3, 4, 5, 6, 7, 66, 679)]


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def _sim(prob, **kw):
    from wql.solver import Simulation
    return Simulation(prob, **kw)


@pytest.mark.parametrize('case', G.SBP_RHS_CASES)
def test_sbp_rhs_vs_golden(case):
    """volume rates, then rates with the SAT and fault terms, against the
    reference build's for the same pattern state"""
    _say(f'[gpu] sbp rhs {case}')
    z = G.load(case)
    prob = G.problem(case)
    assert prob.fd_type != 'traditional'
    t = float(z['rhs_t'])
    with _sim(prob) as sim:
        H.load_pattern_state(sim, prob)
        sim.eval_rates(t, what=0)
        for ib in range(2):
            assert G.rel_l2(sim.rates(ib), z[f'b{ib + 1}_DFvol']) <= TOL
            for s in range(6):
                key = f'b{ib + 1}_DQvol{s + 1}'
                if key in z:
                    assert G.rel_l2(sim.pml_rates(ib, s).reshape(9, -1), z[key].reshape(9, -1)) <= TOL, key
        H.load_pattern_state(sim, prob)
        sim.eval_rates(t, what=1, stage=2)
        for ib in range(2):
            assert G.rel_l2(sim.rates(ib), z[f'b{ib + 1}_DFfull']) <= TOL
            DS, DW = sim.fault_rates(ib)
            ref = z[f'b{ib + 1}_DS']
            assert np.abs(DS - ref).max() <= TOL * max(1.0, np.abs(ref).max())


def _problem(base, fam, order, n=None):
    prob = copy.deepcopy(G.problem(base))
    prob.fd_type, prob.order = fam, order
    if n is not None:
        for b in prob.blocks:
            b.n = n
    return prob


def _run_both(prob, steps, args, precision='fp64'):
    od = O.OracleDomain(prob)
    F0 = H.pulse_init(prob, args)
    with _sim(prob, precision=precision) as sim:
        for ib in range(2):
            sim.set_fields(ib, F0[ib])
            od.blocks[ib]['F'][...] = F0[ib]
        sim.init_state()
        od.init_state()
        sim.step(steps)
        sim.sync()
        od.step(steps)
        got = [(sim.fields(ib), sim.fault(ib)) for ib in range(2)]
    return got, od


@pytest.mark.parametrize('op', OPS, ids=[f'{f}{o}' for f, o in OPS])
def test_every_operator_curvilinear_rate_state_vs_oracle(op):
    """19x18x17 curvilinear blocks, rate and state (TPV101), a velocity
    pulse: 6 LSRK steps, fields and fault state against the oracle"""
    _say(f'[gpu] sbp oracle curv {op}')
    prob = _problem('up9_curv', *op)
    got, od = _run_both(prob, 6, np.array([6.0, -0.9, 7.4, 0.2, 0.5, 1.0]))
    for ib in range(2):
        F, f = got[ib]
        assert G.rel_l2(F, od.blocks[ib]['F']) <= TOL, ib
        for key in ('S', 'W', 'Svel'):
            ref = od.blocks[ib]['iface'][key]
            assert np.abs(f[key] - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max()), key


@pytest.mark.parametrize('op', OPS, ids=[f'{f}{o}' for f, o in OPS])
def test_every_operator_cartesian_pml_vs_oracle(op):
    """a locked Cartesian pair of 18x17x17 blocks with the PML slabs of the
    up6_pml case on five sides (npml 7, which widens the band of every
    operator of band < 7)"""
    _say(f'[gpu] sbp oracle pml {op}')
    prob = _problem('up6_pml', *op, n=(18, 17, 17))
    got, od = _run_both(prob, 5, np.array([5.0, -1.2, 7.4, 0.2, 0.5, 1.0]))
    for ib in range(2):
        assert G.rel_l2(got[ib][0], od.blocks[ib]['F']) <= TOL, ib


@pytest.mark.parametrize('op', [('upwind', 4), ('upwind', 9), ('upwind_drp', 66)],
                         ids=['upwind4', 'upwind9', 'upwind_drp66'])
def test_fp32_tracks_fp64(op):
    """FP32 volume fields with FP64 fault state stay within 1e-5 of FP64"""
    _say(f'[gpu] sbp fp32 {op}')
    prob = _problem('up9_curv', *op)
    args = np.array([20.0, -0.9, 7.4, 0.2, 0.5, 1.0])
    g64, _ = _run_both(prob, 20, args)
    g32, _ = _run_both(prob, 20, args, precision='fp32')
    for ib in range(2):
        assert G.rel_l2(g32[ib][0], g64[ib][0]) <= 1e-5
        s64, s32 = g64[ib][1]['S'], g32[ib][1]['S']
        assert np.abs(s32 - s64).max() <= 1e-5 * max(1e-12, np.abs(s64).max())


@pytest.mark.parametrize('op', [('upwind', 4), ('upwind_drp', 6), ('upwind_drp', 66)],
                         ids=['upwind4', 'upwind_drp6', 'upwind_drp66'])
def test_tiled_operator_kernel_larger_grid(op):
    """2 x 70x45x40 curvilinear slip-weakening (TPV28) blocks: the tiled
    operator kernel over several q/r tiles (partial last tiles) and s chunks,
    the per-node kernel on the s-band planes; 4 LSRK steps against the
    oracle"""
    _say(f'[gpu] sbp tiled {op}')
    prob = _problem('up4_curv', *op, n=(70, 45, 40))
    got, od = _run_both(prob, 4, np.array([4.0, -0.9, 7.4, 0.2, 0.5, 1.0]))
    for ib in range(2):
        assert G.rel_l2(got[ib][0], od.blocks[ib]['F']) <= TOL, ib
