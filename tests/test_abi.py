"""The C-ABI library without a GPU: it loads, exports every function the
public header declares, and its host-only entry points behave (no device
calls here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import golden_cases as G
from wql import native, problem as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'wql.h')


def _declared():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    names = re.findall(r'^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(wql_\w+)\s*\(', text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_api():
    names = _declared()
    for must in ('wql_create', 'wql_destroy', 'wql_step', 'wql_set_fields', 'wql_get_fields', 'wql_set_fault',
                 'wql_get_fault', 'wql_eval_rates', 'wql_last_error', 'wql_abi_version'):
        assert must in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(native.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding binds exactly the header's functions
    assert sorted(native.EXPORTS) == _declared()


def test_exports_are_c_linkage():
    out = subprocess.run(['nm', '-D', '--defined-only', native.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for n in _declared():
        assert n in syms, n  # unmangled


def test_abi_version_matches_header():
    m = re.search(r'#define\s+WQL_ABI_VERSION\s+(\d+)', open(HEADER).read())
    assert native.lib().wql_abi_version() == int(m.group(1))


def test_tile_variants_and_bad_arguments():
    L = native.lib()
    assert L.wql_tile_variants() == 7
    # null context: argument errors, reported without touching a device
    assert L.wql_set_tile(None, 0, 4) == native.WQL_E_ARG
    assert L.wql_step(None, 1) == native.WQL_E_ARG
    assert b'null' in L.wql_last_error() or L.wql_last_error()


@pytest.mark.parametrize('case', ['cart_sw', 'curv_sw', 'pml_locked'])
def test_estimate_bytes_scales_with_the_problem(case):
    from wql.solver import estimate_bytes
    prob = G.problem(case)
    for b in prob.blocks:
        b.n = (100, 40, 37)
    e64 = estimate_bytes(prob, 'fp64')
    e32 = estimate_bytes(prob, 'fp32')
    assert e64 > e32 > 0
    big = G.problem(case)
    for b in big.blocks:
        b.n = (200, 80, 74)
    assert estimate_bytes(big, 'fp64') > 6 * e64


def test_namelist_reader():
    text = """
&problem_list
  name = 'x', t_final = 1.5d0, coupling = 'locked' ! comment
  ny = 13
/
"""
    d = P.read_namelist(text, 'problem_list')
    assert d['t_final'] == 1.5 and d['coupling'] == 'locked' and d['ny'] == 13
    for case in G.TRAJ_CASES:
        prob = G.problem(case)
        assert len(prob.blocks) == 2
        assert prob.blocks[0].n[1:] == prob.blocks[1].n[1:]
