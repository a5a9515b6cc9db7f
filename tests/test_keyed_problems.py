"""The problem names the reference keys its fault fields and media by, on the
host: the library's wql_problem_keys and the Python host's fault_params give
the same keyed law; init_material's layered media (wql_problem_material,
material.f90:179-690) equal the reference build's material at every node."""
import ctypes as C

import numpy as np
import pytest

import golden_cases as G
from wql import native
from wql.problem import fault_params, problem_material


@pytest.mark.parametrize('name', sorted(native.FAULT_KEYS))
def test_library_and_python_keys_agree(name):
    d = native.ProblemDesc()
    mbe = (C.c_double * 3)(0.0, 0.0, 0.0)
    native.check(native.lib().wql_problem_keys(name.encode(), b'slip-weakening_friction', b'elastic', b'default',
                                               mbe, C.byref(d)))
    f = fault_params(name, 'slip-weakening_friction')
    assert d.fault.key == f.key == native.FAULT_KEYS[name]
    assert d.fault.law == f.law == 1


def test_tpv30_plastic_background():
    d = native.ProblemDesc()
    mbe = (C.c_double * 3)(0.0, 0.0, 0.0)
    native.check(native.lib().wql_problem_keys(b'TPV30', b'slip-weakening_friction', b'plastic', b'scec', mbe,
                                               C.byref(d)))
    assert d.s0_kind == 2 and d.plastic == 1
    assert (d.pl_c, d.pl_nu, d.pl_Tv) == (1.18, 0.1680, 0.05)
    p = G.problem('tpv30_plastic')
    assert p.s0_kind == 2 and p.plastic == 1


@pytest.mark.parametrize('case', G.MEDIUM_CASES)
def test_medium_vs_reference(case):
    z = G.load(case)
    prob = G.problem(case)
    for ib, b in enumerate(prob.blocks):
        assert b.material is None and b.medium is not None
        M = b.medium_material()
        ref = z[f'b{ib + 1}_Mat']
        if ref.ndim == 1:
            ref = np.broadcast_to(ref[:, None, None, None], M.shape)
        assert np.array_equal(M, ref), ib
        assert np.array_equal(b.lame(), M)


def test_no_medium_for_other_problems():
    prob = G.problem('cart_sw')
    assert all(b.material is None and b.medium is None for b in prob.blocks)
    assert problem_material('TPV5', prob.blocks[0]) is None


def test_medium_layers_exercised():
    """the graded layers of each medium are reached by the golden grids"""
    for case, n_min in (('tpv31_sw', 5), ('tpv32_sw', 8), ('oklahoma_locked', 4), ('tpv33_sw', 2)):
        prob = G.problem(case)
        mus = np.unique(np.concatenate([b.medium_material()[1].ravel() for b in prob.blocks]))
        assert mus.size >= n_min, case


def test_oracle_refuses_keyed_faults():
    import oracle as O
    with pytest.raises(NotImplementedError):
        O.OracleDomain(G.problem('tpv31_sw'))
