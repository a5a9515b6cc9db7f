"""The Fortran host (waveqlab3d-gh_amd/fhost): the reference's main program
shape -- read the input file named on the command line with its own
&problem_list / &block_list namelists, `init_domain`, then
`time_step_RK(D, D%dt, RK, n)` per step -- over ISO_C_BINDING calls into
libwql.so.

* wql_main on the golden input files against the reference build's
  trajectories (FP64, relative L2 <= 1e-12 on the fields and the fault
  state), also with the block list read through the `&block_list = 'file'`
  indirection;
* wql_main on the fault-output input against the eight fault files the
  reference build wrote for it;
* wql_run (a problem built in Fortran without an input file) against the
  Python host.
"""
import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest

import golden_cases as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FHOST = os.path.join(ROOT, 'waveqlab3d-gh_amd', 'fhost')
MAIN = os.path.join(FHOST, 'wql_main')
RUN = os.path.join(FHOST, 'wql_run')


def _have(exe):
    return os.path.exists(exe)


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.skipif(not _have(MAIN), reason='Fortran host not built (no flang)')
def test_fortran_host_usage():
    for exe in (MAIN, RUN):
        r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
        assert 'usage' in (r.stdout + r.stderr)


def _check(inp):
    r = subprocess.run([MAIN, inp, 'check'], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines() if ln.strip()}
    return lines


@pytest.mark.skipif(not _have(MAIN), reason='Fortran host not built (no flang)')
@pytest.mark.parametrize('case', list(G.TRAJ_CASES))
def test_fortran_reader_matches_python_reader(case, tmp_path):
    """the Fortran host's namelist reader (native namelist I/O with the
    reference's variables) against the Python host's, directly and through
    the `&block_list = 'file'` indirection (CPU: no device call)"""
    from wql.problem import read_namelist
    inp = os.path.join(G.GOLDEN, 'inputs', f'{case}.in')
    text = open(inp).read()
    a = text.index('&block_list')
    b = text.index('/', text.index('faultsize', text.index('btp(2)%faultsize'))) + 1
    (tmp_path / 'blocks.nml').write_text(text[a:b] + '\n')
    (tmp_path / 'main.in').write_text(text[:a] + "&block_list = 'blocks.nml'\n" + text[b:])
    pl, bl = read_namelist(text, 'problem_list'), read_namelist(text, 'block_list')
    for path in (inp, str(tmp_path / 'main.in')):
        got = _check(path)
        assert got['problem'] == [pl['problem'], pl['coupling'], pl['response'], pl['type_of_mesh']]
        assert float(got['time'][0]) == pl['t_final'] and float(got['time'][1]) == pl['cfl']
        assert got['operator'] == [pl.get('fd_type', 'traditional'), str(pl.get('order', 5))]
        for i in (1, 2):
            v = got[f'block{i}']
            g = lambda k: bl[f'btp({i})%{k}']  # noqa: E731
            assert [int(x) for x in v[:3]] == g('nqrs')
            assert [float(x) for x in v[3:6]] == [float(x) for x in g('aqrs')]
            assert [float(x) for x in v[6:9]] == [float(x) for x in g('bqrs')]
            assert [float(x) for x in v[9:11]] == [float(g('lc')), float(g('rc'))]
            assert [int(x) for x in v[11:17]] == list(g('lqrs')) + list(g('rqrs'))
            assert [x == 'T' for x in v[17:23]] == list(g('pml_lqrs')) + list(g('pml_rqrs'))
            assert int(v[23]) == g('npml') and v[24] == g('profile_type')


def _read_dump(path, prob):
    a = np.fromfile(path, '<f8')
    out, o = {}, 0
    for ib, b in enumerate(prob.blocks):
        nq, nr, ns = b.n
        n = 9 * nq * nr * ns
        out[f'b{ib + 1}_F'] = a[o:o + n].reshape(9, ns, nr, nq)
        o += n
    for ib, b in enumerate(prob.blocks):
        nq, nr, ns = b.n
        for key, nf in (('S', 4), ('W', 1), ('trup', 1), ('Svel', 3)):
            n = nf * nr * ns
            out[f'b{ib + 1}_{key}'] = a[o:o + n].reshape(nf, ns, nr)
            o += n
    assert o == a.size
    return out


def _run_main(inp, steps, cwd, extra=()):
    os.makedirs(os.path.join(cwd, 'data'), exist_ok=True)
    dump = os.path.join(cwd, 'dump.bin')
    r = subprocess.run([MAIN, inp, str(steps), *extra, dump], capture_output=True, text=True, timeout=300, cwd=cwd)
    assert r.returncode == 0, r.stdout + r.stderr
    return dump, r.stdout


@pytest.mark.gpu
@pytest.mark.skipif(not _have(MAIN), reason='Fortran host not built (no flang)')
@pytest.mark.parametrize('case', ['cart_sw', 'curv_sw', 'curv_rs_long', 'plastic_sw', 'rw_plastic', 'tpv26_sw',
                                  'mt_cart', 'mt_curv', 'mt_smooth', 'up4_curv', 'drp66_cart', 'an_curv'])
def test_fortran_main_matches_reference(case, tmp_path):
    _say(f'[gpu] fortran main {case}')
    z = G.load(case)
    prob = G.problem(case)
    steps = int(z['traj_args'][0])
    inp = os.path.join(G.GOLDEN, 'inputs', f'{case}.in')
    dump, out = _run_main(inp, steps, str(tmp_path))
    m = re.search(r'steps\s+(\d+)\s+t\s+(\S+)\s+dt\s+(\S+)', out)
    assert int(m.group(1)) == steps
    assert float(m.group(2)) == pytest.approx(float(z['t_end']), rel=1e-14)
    assert float(m.group(3)) == pytest.approx(float(z['dt']), rel=1e-15)
    got = _read_dump(dump, prob)
    tol = G.TRAJ_TOL.get(case, 1e-12)
    for ib in (1, 2):
        assert G.rel_l2(got[f'b{ib}_F'], z[f'b{ib}_F']) <= tol, f'F block {ib}'
        for key in ('S', 'W', 'Svel', 'trup'):
            ref = z[f'b{ib}_{key}']
            assert np.abs(got[f'b{ib}_{key}'] - ref).max() <= max(1e-10, tol) * max(1.0, np.abs(ref).max()), key


@pytest.mark.gpu
@pytest.mark.skipif(not _have(MAIN), reason='Fortran host not built (no flang)')
def test_fortran_main_material_files(tmp_path):
    """material_source = 'file': the host reads each block's three files
    from the run directory (init_material_from_file)"""
    _say('[gpu] fortran main material files')
    z = G.load('matfile_sw')
    prob, inp = G.matfile_problem('matfile_sw', str(tmp_path))
    steps = int(z['traj_args'][0])
    dump, out = _run_main(inp, steps, str(tmp_path))
    got = _read_dump(dump, prob)
    for ib in (1, 2):
        assert G.rel_l2(got[f'b{ib}_F'], z[f'b{ib}_F']) <= 1e-12, f'F block {ib}'
        ref = z[f'b{ib}_S']
        assert np.abs(got[f'b{ib}_S'] - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max())


@pytest.mark.gpu
@pytest.mark.skipif(not _have(MAIN), reason='Fortran host not built (no flang)')
def test_fortran_main_block_list_file(tmp_path):
    """the same run with the block list moved to a second file named by
    `&block_list = 'blocks/curv_sw.nml'` (input_utils.f90:45-94)"""
    _say('[gpu] fortran main block-list file')
    z = G.load('curv_sw')
    prob = G.problem('curv_sw')
    text = open(os.path.join(G.GOLDEN, 'inputs', 'curv_sw.in')).read()
    a = text.index('&block_list')
    b = text.index('/', text.index('faultsize', text.index('btp(2)%faultsize'))) + 1
    os.makedirs(tmp_path / 'blocks')
    (tmp_path / 'blocks' / 'curv_sw.nml').write_text(text[a:b] + '\n')
    (tmp_path / 'main.in').write_text(text[:a] + "&block_list = 'blocks/curv_sw.nml'\n" + text[b:])
    dump, _ = _run_main(str(tmp_path / 'main.in'), 12, str(tmp_path))
    got = _read_dump(dump, prob)
    for ib in (1, 2):
        assert G.rel_l2(got[f'b{ib}_F'], z[f'b{ib}_F']) <= 1e-12, f'F block {ib}'


@pytest.mark.gpu
@pytest.mark.skipif(not _have(MAIN), reason='Fortran host not built (no flang)')
@pytest.mark.parametrize('case', ['fo_cart_sw', 'fo_curv_sw'])
def test_fortran_main_fault_output(case, tmp_path):
    """the eight fault output files of a full run (t_final, w_stride from the
    input) against the reference build's, as in test_fault_output.py"""
    from wql import fault_output as FO
    from wql.problem import from_namelist
    _say(f'[gpu] fortran main fault output {case}')
    fo = os.path.join(G.GOLDEN, 'fault_output')
    prob = from_namelist(os.path.join(fo, f'{case}.in'))
    _run_main(os.path.join(fo, f'{case}.in'), 0, str(tmp_path))
    nr, ns = prob.blocks[0].n[1:]
    for e in FO.EXTS:
        name = str(tmp_path / 'data' / case)
        got = FO.read(name, e, nr, ns).astype(np.float64)
        ref = FO.read(os.path.join(fo, case), e, nr, ns).astype(np.float64)
        assert got.shape == ref.shape, e
        scale = max(np.abs(ref).max(), 1e-30)
        assert (np.abs(got - ref) <= 2.4e-7 * np.abs(ref) + 1e-9 * scale).all(), e


@pytest.mark.gpu
@pytest.mark.skipif(not _have(RUN), reason='Fortran host not built (no flang)')
@pytest.mark.parametrize('prec', ['fp64', 'fp32'])
def test_fortran_run_matches_python_host(prec):
    """a two-block Cartesian rupture built in Fortran (no input file) and the
    same problem from Python give the same fault slip"""
    n, steps = (41, 31, 31), 20
    _say(f'[gpu] fortran run {prec}')
    r = subprocess.run([RUN, str(n[0]), str(n[1]), str(n[2]), str(steps)] + (['fp32'] if prec == 'fp32' else []),
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


@pytest.mark.gpu
@pytest.mark.skipif(not _have(MAIN), reason='Fortran host not built (no flang)')
@pytest.mark.parametrize('case', G.GRID_CASES)
def test_fortran_main_fault_profile_grids(case, tmp_path):
    """the Fortran host's grids of the TPV28, TPV10, TPV36 and file fault
    profiles (wql_grid_block) against the reference build's"""
    _say(f'[gpu] fortran grid {case}')
    z = G.load(case)
    prob = G.grid_problem(case, str(tmp_path))
    inp = os.path.join(str(tmp_path), f'{case}.in')
    if not os.path.exists(inp):
        inp = os.path.join(G.GOLDEN, 'inputs', f'{case}.in')
    dump = os.path.join(str(tmp_path), 'grid.bin')
    r = subprocess.run([MAIN, inp, 'grid', dump], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    a = np.fromfile(dump, '<f8')
    o = 0
    for ib, b in enumerate(prob.blocks):
        nq, nr, ns = b.n
        X = a[o:o + 3 * nq * nr * ns].reshape((nq, nr, ns, 3), order='F').transpose(3, 2, 1, 0)
        o += 3 * nq * nr * ns
        ref = z[f'b{ib + 1}_X']
        assert np.abs(X - ref).max() <= 1e-13 * np.abs(ref).max(), ib
    assert o == a.size
