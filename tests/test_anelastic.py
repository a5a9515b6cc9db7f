"""The anelastic response without a GPU: the input-file keys (&problem_list
response = 'anelastic', &anelastic_list), the library's key table and its
argument checks.  The arithmetic (Q^-1, unrelaxed moduli, memory-variable
rates, the PML form) is pinned by the reference-build goldens an_curv,
an_pml and an_drp6_pml through the oracle (test_oracle_golden) and on the
GPU (test_anelastic_gpu, test_gpu_parity)."""
import copy
import ctypes as C

import pytest

import golden_cases as G


def test_anelastic_list_is_read():
    assert G.problem('an_curv').anelastic == (2.0, 0.6, 1.0)
    assert G.problem('an_pml').anelastic == (1.5, 0.0, 1.0)
    p = G.problem('an_drp6_pml')
    assert (p.anelastic, p.fd_type, p.order) == ((2.0, 0.6, 1.0), 'upwind_drp', 6)
    assert G.problem('cart_sw').anelastic is None


def test_problem_keys_anelastic_defaults():
    from wql import native
    out = native.ProblemDesc()
    out.precision, out.dt = native.WQL_FP64, 1e-3
    native.check(native.lib().wql_problem_keys(b'TPV5', b'locked', b'anelastic', b'default', None, C.byref(out)))
    assert (out.anelastic, out.an_c, out.an_weight_exp, out.an_fref, out.plastic) == (1, 1.0, 0.0, 1.0, 0)
    native.check(native.lib().wql_problem_keys(b'TPV5', b'locked', b'elastic', b'default', None, C.byref(out)))
    assert out.anelastic == 0


def test_anelastic_refusals():
    """weight_exp outside the reference's two weight sets, and anelastic with
    plasticity, are argument errors before any device call"""
    from wql import native
    from wql.solver import Simulation
    prob = copy.deepcopy(G.problem('an_curv'))
    prob.anelastic = (2.0, 0.3, 1.0)
    with pytest.raises(native.WqlError, match='weight set'):
        Simulation(prob)
    prob.anelastic = (2.0, 0.6, 1.0)
    prob.plastic = 1
    with pytest.raises(native.WqlError, match='plastic or anelastic'):
        Simulation(prob)
