"""The upwind and upwind-DRP operator tables (waveqlab3d-gh_amd/csrc/wql_sbp_ops.h,
read from JU_xJU_yJU_z6.f90:815-20922 by make_sbp_ops.py) against what the
mathematics of upwind summation-by-parts operators fixes, independently of
the reference-build goldens (which pin seven of the fifteen operators):

- consistency: every row maps constants to 0 and x to 1;
- the interior stencils are exact for polynomials up to the operator's order;
- the pair (Dp for velocities, Dm for stresses) satisfies the upwind SBP
  identity H Dp + (H Dm)^T = diag(-1, 0, ..., 0, 1) with a diagonal H, and
  H_00 = 1/tau0 (the SAT penalty, block.f90:196-255).

A dropped term, a wrong sign or index in reading the tables breaks one of
these.  The reference's own rows that break them are listed exactly: the
orders 8 and 9 print inconsistent closure rows (and drop the tap n-12 of
the stress row n-7 in r and s only), DRP 679 prints inconsistent closure
rows, DRP 66's interior routine uses rationals that do not sum to zero, and
two tau0 differ from 1/H_00 (upwind 6: 48/47 of it; DRP 66: a rounded
fraction).  The library follows the reference in all of them.
"""
import numpy as np
import pytest

import oracle as O

N = 40


def matrix(h, n=N):
    D = np.zeros((n, n))
    for p in range(n):
        if p < h.nl:
            D[p, :16] = list(h.left[p])
        elif p >= n - h.nr:
            for t in range(16):
                D[p, n - 1 - t] += h.right[n - 1 - p][t]
        else:
            for o in range(h.lo, h.hi + 1):
                D[p, p + o] = h.band[o - h.lo]
    return D


def ops():
    return [(op.family.decode(), op.order, op) for op in O.sbp_ops()]


def test_fifteen_operators_and_the_lookup():
    names = [(f, o) for f, o, _ in ops()]
    assert names == [('upwind', o) for o in range(2, 10)] + [('upwind_drp', o) for o in (3, 4, 5, 6, 7, 66, 679)]
    L = O.lib()
    for k, (f, o) in enumerate(names):
        assert L.ora_sbp_op(1 if f == 'upwind' else 2, o) == k
    assert L.ora_sbp_op(0, 5) == -1 and L.ora_sbp_op(1, 10) == -2 and L.ora_sbp_op(2, 2) == -2


# (family, order) -> rows (direction, 'v'|'s', row; negative rows count from
# the right face, -1 the last node) where the reference prints an
# inconsistent row
_V67 = {(d, 'v', r) for d in range(3) for r in (6, 7)}
INCONSISTENT = {
    ('upwind', 8): _V67 | {(d, 's', -8) for d in (1, 2)},
    ('upwind', 9): _V67 | {(d, 's', 7) for d in range(3)} | {(d, 's', -8) for d in (1, 2)},
    ('upwind_drp', 679): {(d, 'v', r) for d in range(3) for r in (3, 4, 5, 6, 7, -8, -7, -6, -5)}
                         | {(d, 's', r) for d in range(3) for r in (4, 5, 6, 7, -8, -7, -6, -5, -4)},
}


@pytest.mark.parametrize('k', range(15))
def test_rows_are_consistent(k):
    fam, order, op = ops()[k]
    bad = set()
    for d in range(3):
        for g, h in (('v', op.v[d]), ('s', op.s[d])):
            D = matrix(h)
            for p in range(N):
                if abs(D[p].sum()) > 1e-12 or abs(D[p] @ np.arange(N) - 1.0) > 1e-11:
                    bad.add((d, g, p if p < N // 2 else p - N))
    assert bad == INCONSISTENT.get((fam, order), set())


@pytest.mark.parametrize('k', range(15))
def test_interior_stencils_exact_to_the_order(k):
    fam, order, op = ops()[k]
    want = {('upwind_drp', 66): (0, 6), ('upwind_drp', 679): (6, 7)}.get((fam, order), (order, order))
    for d in range(3):
        for h in (op.v[d], op.s[d]):
            offs = np.arange(h.lo, h.hi + 1, dtype=float)
            for which, row, deg in (('exact', h.exact, want[0]), ('band', h.band, want[1])):
                c = np.array(row[:len(offs)])
                got = 0
                while got < 12 and abs((c * offs ** (got + 1)).sum() - (1.0 if got == 0 else 0.0)) \
                        < 1e-11 * np.abs(offs).max() ** (got + 1):
                    got += 1
                if (fam, order, which) == ('upwind_drp', 66, 'exact'):
                    # the interior routine's rationals sum to -6.0e-5 (v) / +6.0e-5 (s)
                    assert abs(abs(c.sum()) - 6.025068362847108e-05) < 1e-15
                else:
                    assert abs(c.sum()) < 1e-15 * len(c) * 4
                assert got == deg, (which, got)


@pytest.mark.parametrize('k', range(15))
def test_upwind_sbp_identity(k):
    fam, order, op = ops()[k]
    for d in range(3):
        Dp, Dm = matrix(op.v[d]), matrix(op.s[d])
        rows, rhs = [], []
        for a in range(N):
            for b in range(N):
                r = np.zeros(N)
                r[a] += Dp[a, b]
                r[b] += Dm[b, a]
                if r.any() or (a == b and a in (0, N - 1)):
                    rows.append(r)
                    rhs.append(-1.0 if a == b == 0 else 1.0 if a == b == N - 1 else 0.0)
        A, rhs = np.array(rows), np.array(rhs)
        H = np.linalg.lstsq(A, rhs, rcond=None)[0]
        resid = np.abs(A @ H - rhs).max()
        if (fam, order) in (('upwind', 8), ('upwind', 9)):
            assert resid > 1e-3      # the inconsistent closure rows break it
            continue
        assert resid < 1e-13
        assert (H > 0).all()
        ratio = H[0] * op.tau0
        want = {('upwind', 6): 47.0 / 48.0, ('upwind_drp', 66): 0.999999443703}.get((fam, order), 1.0)
        assert abs(ratio - want) < 1e-11, ratio


def test_planted_sign_error_is_caught():
    """the consistency and SBP checks reject a table with one flipped sign"""
    op = ops()[2][2]
    D = matrix(op.v[0])
    D[2, 3] = -D[2, 3]
    assert abs(D[2].sum()) > 1e-3


def test_undefined_operator_is_refused():
    """wql_create refuses an order the family does not define and blocks too
    small for the operator's closure rows, before any device call"""
    import copy
    import golden_cases as G
    from wql import native
    from wql.solver import Simulation
    prob = copy.deepcopy(G.problem('up4_curv'))
    prob.order = 10
    with pytest.raises(native.WqlError, match='no operator'):
        Simulation(prob)
    prob.order = 9   # 14x13x13 blocks: order 9 needs 17 nodes per direction
    with pytest.raises(native.WqlError, match='needs 17 nodes'):
        Simulation(prob)
