"""The Fortran host (waveqlab3d-gh_amd/fhost): an ISO_C_BINDING program that
drives libwql.so the way a WaveQLab3D main program would."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, 'waveqlab3d-gh_amd', 'fhost', 'wql_run')


def _have_exe():
    return os.path.exists(EXE)


@pytest.mark.skipif(not _have_exe(), reason='Fortran host not built (no flang)')
def test_fortran_host_usage():
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert 'usage' in (r.stdout + r.stderr)


@pytest.mark.gpu
@pytest.mark.skipif(not _have_exe(), reason='Fortran host not built (no flang)')
@pytest.mark.parametrize('prec', ['fp64', 'fp32'])
def test_fortran_host_matches_python_host(prec):
    """the same two-block Cartesian rupture stepped from Fortran and from
    Python through the C ABI gives the same fault slip"""
    n, steps = (41, 31, 31), 20
    print(f'[gpu] fortran host {prec}', file=sys.stderr, flush=True)
    r = subprocess.run([EXE, str(n[0]), str(n[1]), str(n[2]), str(steps)] + (['fp32'] if prec == 'fp32' else []),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r't\s+(\S+)\s+max_slip\s+(\S+)\s+sum_slip\s+(\S+)', r.stdout)
    t_f, max_f, sum_f = (float(v) for v in m.groups())
    from wql import workloads as Wk
    from wql.problem import Problem
    from wql.solver import Simulation
    h = 0.1
    blocks, coords = Wk._blocks(n, h, mesh='cartesian')
    Ly, Lz = (n[1] - 1) * h, (n[2] - 1) * h
    prob = Problem(blocks=blocks, fault=Wk.sw_fault(depth=Ly, width=Lz), cfl=0.5, name='fhost')
    with Simulation(prob, precision=prec, coords=coords) as sim:
        sim.step(steps)
        sim.sync()
        S = sim.fault(0)['S'][3]
        assert sim.t == pytest.approx(t_f, rel=1e-14)
    assert max_f > 0
    assert S.max() == pytest.approx(max_f, rel=1e-12)
    assert S.sum() == pytest.approx(sum_f, rel=1e-12)
