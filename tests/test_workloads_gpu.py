"""Every benchmark workload builder on the HIP path against the oracle.

Each workload runs at reduced node counts (same h, physics and seeds) for 40
LSRK steps: the 2-block plain pulse, the curvilinear rupture with random
topography, the rough rate-and-state fault with PML (slip law, Newton solve),
the heterogeneous plastic medium with the TPV26-style fault, and two serving
scenarios.  Two changes let the short runs slip: the rough fault's
nucleation ramps up over 0.02 s instead of 1 s, and the plastic workload's
fault, only ~1 km deep here, has no cohesion and a 0.02 s forced weakening
(both fault terms are pinned at full size by the rw_plastic / tpv26 goldens).

Bars: FP64 relative L2 <= 1e-12 on the fields of both blocks and on slip,
rupture time and peak slip rate; FP32 volume with the FP64 fault <= 1e-5 on
slip, rupture time and peak slip rate (and on the fields for the locked
interface).
"""
import sys

import numpy as np
import pytest

import golden_cases as G
import oracle as O

pytestmark = pytest.mark.gpu

STEPS = 40
CASES = [('plain', 0.5), ('rupture', 0.2), ('rough', 0.15), ('plastic', 0.15), ('scenario0', 0.25),
         ('scenario3', 0.25)]
_oracle = {}


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def _workload(name, scale):
    from wql import workloads as Wk
    if name.startswith('scenario'):
        return Wk.scenario(int(name[len('scenario'):]), scale=scale)
    W = Wk.WORKLOADS[name](scale=scale)
    f = W['problem'].fault
    if name == 'rough':
        f.nucp = tuple(f.nucp[:4]) + (0.02,)
    if name == 'plastic':
        f.coh = (0.0, 0.0, 0.0)
        f.frc = tuple(f.frc[:4]) + (0.02,)
    return W


def _start(sim_or_od, W, is_oracle):
    for ib in range(2):
        if W.get('F0') is not None:
            if is_oracle:
                sim_or_od.blocks[ib]['F'][...] = W['F0'][ib]
            else:
                sim_or_od.set_fields(ib, W['F0'][ib])
    if W.get('init_state'):
        sim_or_od.init_state()


def _run_oracle(name, scale):
    if name in _oracle:
        return _oracle[name]
    W = _workload(name, scale)
    od = O.OracleDomain(W['problem'], coords=W['coords'])
    _start(od, W, True)
    psr = np.zeros(od.blocks[0]['iface']['trup'].shape)
    for _ in range(STEPS):
        od.step(1)
        sv = od.blocks[0]['iface']['Svel']  # stage-1 snapshot of the step
        psr = np.maximum(psr, np.sqrt(sv[1] ** 2 + sv[2] ** 2))
    out = dict(F=[od.blocks[ib]['F'].copy() for ib in range(2)],
               S=od.blocks[0]['iface']['S'].copy(), trup=od.blocks[0]['iface']['trup'].copy(), psr=psr,
               P=[od.blocks[ib]['P'].copy() for ib in range(2)], t=od.t)
    _oracle[name] = out
    return out


def _tuned(name, prec):
    """the launch shape the autotuner's cache holds for the full-size workload"""
    from wql import autotune
    from wql import workloads as Wk
    if name.startswith('scenario'):
        return None
    full = Wk.WORKLOADS[name](scale=1.0)['problem']
    return autotune.lookup(full, prec)


@pytest.mark.parametrize('prec', ['fp64', 'fp32'])
@pytest.mark.parametrize('shape', ['default', 'tuned'])
@pytest.mark.parametrize('name,scale', CASES)
def test_workload_vs_oracle(name, scale, shape, prec):
    """shape 'tuned': the reduced run on the launch shape the autotuner
    chose for the full-size workload"""
    from wql.solver import Simulation
    kw = {}
    if shape == 'tuned':
        e = _tuned(name, prec)
        if e is None:
            pytest.skip('no tuned shape for this workload')
        kw = dict(tile=e['variant'], zchunk=e['zchunk'], order=e.get('order', 0))
    _say(f'[gpu] workload {name} scale {scale} {prec} {shape} {kw}')
    ref = _run_oracle(name, scale)
    W = _workload(name, scale)
    prob = W['problem']
    with Simulation(prob, precision=prec, coords=W['coords'], **kw) as sim:
        _start(sim, W, False)
        sim.step(STEPS)
        sim.sync()
        assert sim.t == pytest.approx(ref['t'], rel=1e-14)
        F = [sim.fields(ib) for ib in range(2)]
        f = sim.fault(0)
        psr = sim.fault_peak(0)
        P = [sim.plastic(ib) for ib in range(2)] if prob.plastic else None
    slipping = np.abs(ref['S'][3]).max() > 1e-6
    rup = ref['trup'] < 1e9
    if prec == 'fp64':
        for ib in range(2):
            assert G.rel_l2(F[ib], ref['F'][ib]) <= 1e-12, f'F block {ib + 1}'
            if P is not None:
                assert G.rel_l2(P[ib], ref['P'][ib]) <= 1e-12, f'P block {ib + 1}'
        # slip, and a locked interface's round-off slip (~1e-20), on the
        # scale of a 1 m slip / 1 m/s slip rate at least
        assert np.abs(f['S'] - ref['S']).max() <= 1e-12 * max(1.0, np.abs(ref['S']).max())
        assert np.array_equal(f['trup'] < 1e9, rup)
        if rup.any():
            assert G.rel_l2(f['trup'][rup], ref['trup'][rup]) <= 1e-12
        assert np.abs(psr - ref['psr']).max() <= 1e-12 * max(1.0, ref['psr'].max())
        if slipping:
            assert G.rel_l2(f['S'][3], ref['S'][3]) <= 1e-12
            assert G.rel_l2(psr, ref['psr']) <= 1e-12
    else:
        if not slipping:
            for ib in range(2):
                assert G.rel_l2(F[ib], ref['F'][ib]) <= 1e-5, f'F block {ib + 1}'
        else:
            assert G.rel_l2(f['S'][3], ref['S'][3]) <= 1e-5, 'slip'
            assert G.rel_l2(psr, ref['psr']) <= 1e-5, 'peak slip rate'
            both = rup & (f['trup'] < 1e9)
            assert both.sum() >= 0.95 * rup.sum()
            assert G.rel_l2(f['trup'][both], ref['trup'][both]) <= 1e-5, 'rupture time'
    if name != 'plain':
        assert slipping, 'the test run should slip'
