"""Block grids generated on the device (wql_grid_block) against the host
formulation (wql/geometry.py, which reproduces the reference build's grid of
curv_sw to 1e-14, tests/test_oracle_golden.py): the uniform grid, the
'analytical_test_problem' bump profile of both golden blocks, planar faces
with a fault-surface and a free-surface offset field; and against the
reference build's grids of the TPV28, TPV10 and TPV36 fault profiles and of a
fault surface read from a file."""
import sys

import numpy as np
import pytest

import golden_cases as G

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize('case', ['cart_sw', 'curv_sw'])
def test_golden_grids(case):
    from wql.solver import device_grid
    _say(f'[gpu] grid {case}')
    z = G.load(case)
    prob = G.problem(case)
    for ib, b in enumerate(prob.blocks):
        X = device_grid(b)
        assert np.abs(X - b.coords()).max() <= 1e-14 * max(1.0, np.abs(X).max())
        if case == 'curv_sw':
            assert np.abs(X - z[f'b{ib + 1}_X']).max() <= 1e-13


def test_offset_fields():
    from wql import problem as P
    from wql.solver import device_grid
    _say('[gpu] grid offsets')
    rng = np.random.default_rng(9)
    n = (23, 19, 17)
    for side, bc in ((1, (1, 0, 2, 1, 1, 1)), (0, (0, 1, 2, 1, 1, 1))):
        b = P.Block(n=n, a=(-2.2, 0.0, -0.8), b=(0.0, 1.8, 0.8), bc=bc, mesh='curvilinear', profile='flat')
        b.fault_offset = 0.01 * rng.standard_normal((n[1], n[2]))
        b.top_offset = 0.02 * rng.standard_normal((n[0], n[2]))
        X = device_grid(b)
        ref = b.coords()
        assert np.abs(X - ref).max() <= 1e-14 * np.abs(ref).max(), side


@pytest.mark.parametrize('case', G.GRID_CASES)
def test_fault_profile_grids(case, tmp_path):
    from wql.solver import device_grid
    _say(f'[gpu] grid {case}')
    z = G.load(case)
    prob = G.grid_problem(case, str(tmp_path))
    for ib, b in enumerate(prob.blocks):
        X = device_grid(b)
        ref = z[f'b{ib + 1}_X']
        assert np.abs(X - ref).max() <= 1e-13 * np.abs(ref).max(), ib
