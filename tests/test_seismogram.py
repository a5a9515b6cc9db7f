"""Station seismograms (seismogram.f90) against the reference build's files
for the same input (tests/golden/seismogram/, made by
tests/golden/make_fault_output.py): six stations per input (on a node,
between nodes, on the fault, on the free surface, inside block 2, off both
blocks), the list in the input file (curvilinear grid, files named by node)
or in an external file (Cartesian grid, files named by coordinates).

CPU: the &output_list reader, the station lists, the node search
(Find_Coordinates) against the nodes the reference's file names carry, and
the file names themselves.  GPU: the Python host's station files of the full
runs, and the Fortran host's (wql_main) for the same inputs, against the
reference's, value by value (FP64; the files print 10 decimals).
"""
import glob
import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest

import golden_cases as G
from wql import seismogram as SG
from wql.problem import from_namelist

SG_DIR = os.path.join(G.GOLDEN, 'seismogram')
CASES = ['sg_curv_sw', 'sg_cart_sw']
MAIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'waveqlab3d-gh_amd', 'fhost',
                    'wql_main')


def _ref_files(case):
    root = os.path.join(SG_DIR, case)
    return sorted(os.path.relpath(p, root) for p in glob.glob(os.path.join(root, '**', '*.dat'), recursive=True))


def _setup(case, tmp):
    """the case's input (and station file) in a fresh run directory"""
    src = os.path.join(SG_DIR, case)
    for fn in os.listdir(src):
        if os.path.isfile(os.path.join(src, fn)):
            shutil.copy(os.path.join(src, fn), os.path.join(tmp, fn))
    os.makedirs(os.path.join(tmp, 'data'), exist_ok=True)
    return os.path.join(tmp, f'{case}.in')


def _expected_names(case, inp):
    prob = from_namelist(inp)
    opts = SG.output_options(inp)
    with open(inp) as f:
        name = re.search(r"name\s*=\s*'([^']*)'", f.read()).group(1)
    out = []
    for ib, b in enumerate(prob.blocks):
        xyz = SG.station_coordinates(inp, opts, ib + 1)
        ijk = SG.find_stations(b.coords(), xyz)
        for s in range(len(xyz)):
            if (ijk[s] >= 0).all():
                out.append(os.path.normpath(SG.file_name(name, opts, ib + 1, ijk[s], xyz[s])))
    return sorted(out)


@pytest.mark.parametrize('case', CASES)
def test_station_nodes_and_names_match_reference(case, tmp_path):
    cwd = os.getcwd()
    try:
        inp = _setup(case, str(tmp_path))
        os.chdir(tmp_path)  # the external station file is named relative to the run directory
        got = _expected_names(case, inp)
    finally:
        os.chdir(cwd)
    assert got == _ref_files(case)


def test_output_options_and_lists():
    inp = os.path.join(SG_DIR, 'sg_cart_sw', 'sg_cart_sw.in')
    o = SG.output_options(inp)
    assert o['output_seismograms'] is True and o['station_xyz_index'] is True
    assert o['station_list'] == 'extfile' and o['station_file_directory'] == 'seis_out'
    o2 = SG.output_options(os.path.join(SG_DIR, 'sg_curv_sw', 'sg_curv_sw.in'))
    assert o2['station_list'] == 'infile' and o2['station_file_directory'] == 'seismogram'
    xyz = SG.station_coordinates(os.path.join(SG_DIR, 'sg_curv_sw', 'sg_curv_sw.in'), o2, 1)
    assert xyz.shape == (6, 3) and xyz[1].tolist() == [-0.8, 6.9, 0.37]
    assert SG.fmt_f(-0.25, 20, 3).strip() == '-0.250' and SG.fmt_f(12345.0, 15, 10) == '*' * 15


def test_find_stations_brute_force():
    """the search on a small skewed grid against a direct restatement: the
    nearest node within half the local diagonal, the last one on ties"""
    rng = np.random.default_rng(3)
    nq, nr, ns = 7, 6, 5
    i, j, k = np.meshgrid(np.arange(nq), np.arange(nr), np.arange(ns), indexing='ij')
    X = np.stack([0.5 * i + 0.05 * j, 0.4 * j + 0.03 * k ** 2, 0.3 * k + 0.02 * i]).transpose(0, 3, 2, 1).copy()
    # nodes, midpoints (ties between two nodes) and random points, some outside
    pts = [X[:, 2, 3, 4], 0.5 * (X[:, 1, 1, 1] + X[:, 1, 1, 2]), [10.0, 10.0, 10.0]]
    pts += list(rng.uniform([-0.5, -0.5, -0.5], [4.0, 2.5, 1.5], (20, 3)))
    xyz = np.array(pts, dtype=np.float64)
    got = SG.find_stations(X, xyz)
    for s, p in enumerate(xyz):
        best, md = (-1, -1, -1), 1.0e8
        for kk in range(ns):
            for jj in range(nr):
                for ii in range(nq):
                    hx = X[0, kk, jj, ii + 1] - X[0, kk, jj, ii] if ii < nq - 1 else X[0, kk, jj, nq - 1] - X[0, kk, jj, nq - 2]
                    hy = X[1, kk, jj + 1, ii] - X[1, kk, jj, ii] if jj < nr - 1 else X[1, kk, nr - 1, ii] - X[1, kk, nr - 2, ii]
                    hz = X[2, kk + 1, jj, ii] - X[2, kk, jj, ii] if kk < ns - 1 else X[2, ns - 1, jj, ii] - X[2, ns - 2, jj, ii]
                    v = X[:, kk, jj, ii] - p
                    dist = np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
                    if dist <= md and dist <= 0.5 * np.sqrt(hx * hx + hy * hy + hz * hz):
                        best, md = (ii, jj, kk), dist
        assert tuple(got[s]) == best, s
    assert tuple(got[0]) == (4, 3, 2) and tuple(got[2]) == (-1, -1, -1)
    assert tuple(got[1]) in ((1, 1, 1), (2, 1, 1))


def _compare(case, run_dir):
    files = _ref_files(case)
    assert files
    for rel in files:
        ref = SG.read(os.path.join(SG_DIR, case, rel))
        got_p = os.path.join(run_dir, rel)
        assert os.path.exists(got_p), rel
        got = SG.read(got_p)
        assert got.shape == ref.shape, rel
        assert np.abs(got - ref).max() <= 1.5e-10, rel
    got_all = sorted(os.path.relpath(p, run_dir) for p in glob.glob(os.path.join(run_dir, '**', '*.dat'), recursive=True))
    assert got_all == files


@pytest.mark.gpu
@pytest.mark.parametrize('case', CASES)
def test_python_host_seismograms_match_reference(case, tmp_path):
    from wql import fault_output as FO
    from wql.solver import Simulation
    print(f'[gpu] seismograms python {case}', file=sys.stderr, flush=True)
    cwd = os.getcwd()
    try:
        inp = _setup(case, str(tmp_path))
        os.chdir(tmp_path)
        prob = from_namelist(inp)
        coords = [b.coords() for b in prob.blocks]
        with Simulation(prob, coords=coords) as sim:
            with SG.Seismograms(sim, inp, coords) as st:
                FO.run(sim, prob.nt(), prob.w_stride, 'unused', stations=st, fault=False)
    finally:
        os.chdir(cwd)
    _compare(case, str(tmp_path))


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(MAIN), reason='Fortran host not built (no flang)')
@pytest.mark.parametrize('case', CASES)
def test_fortran_host_seismograms_match_reference(case, tmp_path):
    print(f'[gpu] seismograms fortran {case}', file=sys.stderr, flush=True)
    inp = _setup(case, str(tmp_path))
    r = subprocess.run([MAIN, inp, '0'], capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    _compare(case, str(tmp_path))
