"""The metric structure that the FP32 path folds into constants, checked on
the oracle's FP64 metrics (grid.f90:396-469 restated) of the rupture grid:
x is affine in q and z affine in s, with topography only in y.  Then q_x and
s_z are constant and q_y, q_z, s_x, s_y vanish, each up to SBP round-off far
below the 1e-10 (relative to the metric scale) that the library allows, while
r_x, r_y, r_z and J vary by far more than that."""
import numpy as np

import oracle as O
from wql import workloads

NAMES = 'q_x r_x s_x q_y r_y s_y q_z r_z s_z J'.split()


def test_rupture_grid_metric_structure():
    W = workloads.WORKLOADS['rupture'](scale=0.25)
    od = O.OracleDomain(W['problem'], coords=W['coords'])
    for ib in range(2):
        G = od.blocks[ib]['metric']
        scale = np.abs(G[:9]).max()
        spread = {NAMES[c]: float(G[c].max() - G[c].min()) for c in range(10)}
        for c in ('q_x', 's_z', 'q_y', 'q_z', 's_x', 's_y'):
            assert spread[c] <= 1e-12 * scale, (ib, c, spread[c])
        for c in ('q_y', 'q_z', 's_x', 's_y'):
            assert np.abs(G[NAMES.index(c)]).max() <= 1e-12 * scale, (ib, c)
        # closed forms: x = q (nq-1) h and z = s (ns-1) h on the unit square
        h = 0.1
        nq, nr, ns = W['problem'].blocks[ib].n
        assert np.allclose(G[NAMES.index('q_x')], 1.0 / ((nq - 1) * h), rtol=1e-13, atol=0)
        assert np.allclose(G[NAMES.index('s_z')], 1.0 / ((ns - 1) * h), rtol=1e-13, atol=0)
        for c in ('r_x', 'r_y', 'r_z'):
            assert spread[c] > 1e-4 * scale, (ib, c, spread[c])
        assert spread['J'] > 1e-4 * np.abs(G[9]).max()
