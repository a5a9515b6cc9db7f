"""The library's problem-key table (wql_problem_keys, used by the Fortran
host) against the Python host's (wql/problem.py), which the reference-build
goldens pin: every field of the fault description and the plasticity
settings, for every key both support.  Runs on the CPU (no device call)."""
import ctypes as C

import numpy as np
import pytest

from wql import native
from wql import problem as P
from wql.solver import fault_struct

KEYS = [('TPV5', 'slip-weakening_friction'), ('TPV28', 'slip-weakening_friction'),
        ('TPV26', 'slip-weakening_friction'), ('TPV27', 'slip-weakening_friction'),
        ('TPV101', 'rate-and-state_friction'), ('TPV102', 'rate-and-state_friction'),
        ('rate-weakening', 'rate-and-state_friction'), ('TPV5', 'locked')]


def _lib_problem(problem, coupling, response='elastic', model='default', mbe=(0.0, 0.0, 0.0)):
    out = native.ProblemDesc()
    m = (C.c_double * 3)(*mbe)
    rc = native.lib().wql_problem_keys(problem.encode(), coupling.encode(), response.encode(), model.encode(), m,
                                       C.byref(out))
    return rc, out


def _fields(s):
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        if hasattr(v, '_length_'):
            v = np.array([list(x) if hasattr(x, '_length_') else x for x in v], dtype=float)
        out[name] = v
    return out


@pytest.mark.parametrize('problem,coupling', KEYS)
def test_fault_table_matches_python(problem, coupling):
    rc, got = _lib_problem(problem, coupling)
    assert rc == 0, native.lib().wql_last_error()
    want = fault_struct(P.fault_params(problem, coupling))
    g, w = _fields(got.fault), _fields(want)
    for k in w:
        if coupling == 'locked' and k not in ('law',):
            continue  # a locked fault reads nothing else
        assert np.array_equal(np.asarray(g[k]), np.asarray(w[k])), k


def test_plastic_settings():
    rc, p = _lib_problem('TPV26', 'slip-weakening_friction', 'plastic', 'scec')
    assert rc == 0 and p.plastic == 1 and p.s0_kind == 1 and (p.pl_c, p.pl_nu, p.pl_Tv) == (1.36, 0.1934, 0.03)
    rc, p = _lib_problem('rate-weakening', 'rate-and-state_friction', 'plastic', 'default', (0.35, 0.15, 0.05))
    assert rc == 0 and p.plastic == 2 and (p.pl_mu, p.pl_beta, p.pl_eta) == (0.35, 0.15, 0.05)
    rc, p = _lib_problem('TPV28', 'locked', 'plastic', 'scec')
    assert rc == 0 and list(p.fault.s0) == [-60.0, 0.0, 60.0, 0.0, 29.38, 0.0]


def test_unknown_keys_are_refused():
    assert _lib_problem('KDN', 'rate-and-state_friction')[0] == native.WQL_E_ARG
    assert b'prestress' in native.lib().wql_last_error()
    assert _lib_problem('TPV5', 'rate-and-state_friction')[0] == native.WQL_E_ARG
    assert _lib_problem('TPV5', 'slip-weakening_friction', 'viscous')[0] == native.WQL_E_ARG
