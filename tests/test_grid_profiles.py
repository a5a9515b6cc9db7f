"""The fault profiles of the curvilinear grid generator on the host
(wql/geometry.py) against the reference build's grids: the TPV28 hills, the
TPV10 60-degree and TPV36 15-degree dipping faults, and a fault surface read
from a text file (grid.f90:625-925, read_2darray 2457-2481); the topography
branch (use_topography: a topography file, the analytical hill, a topography
file with a fault-profile file, grid.f90:1349-1860); the block-list file of an
input file (domain.f90:165-211, input_utils.f90:9-94)."""
import os

import numpy as np
import pytest

import golden_cases as G


@pytest.mark.parametrize('case', G.GRID_CASES)
def test_profile_grid_vs_reference(case, tmp_path):
    z = G.load(case)
    prob = G.grid_problem(case, str(tmp_path))
    for ib, b in enumerate(prob.blocks):
        assert b.mesh == 'curvilinear' and b.profile == str(prob.blocks[0].profile)
        assert (b.topography is not None) == case.startswith('grid_topo')
        X = b.coords()
        ref = z[f'b{ib + 1}_X']
        assert np.abs(X - ref).max() <= 1e-14 * np.abs(ref).max(), ib
        # the fault face, or the free surface of the topography branch, is
        # not planar
        if b.topography:
            assert np.ptp(ref[1][:, 0, :]) > 1e-3
        if b.profile != 'no_path':
            face = ref[0][:, :, -1] if ib == 0 else ref[0][:, :, 0]
            assert np.ptp(face) > 1e-3


def test_profiles_vanish_with_zero_weights(tmp_path):
    """lc = rc = 0 keeps both q faces planar for every formula profile"""
    from wql import geometry
    for prof in geometry.ANALYTICAL_PROFILES:
        X = geometry.curvilinear((9, 21, 13), (-3.0, 0.0, -13.0), (0.0, 20.0, 13.0), 0.0, 0.0, prof)
        assert np.abs(X[0][:, :, 0] + 3.0).max() <= 1e-14 and np.abs(X[0][:, :, -1]).max() <= 1e-14, prof


def test_profile_file_needs_every_row(tmp_path):
    from wql import geometry
    p = tmp_path / 'short.dat'
    p.write_text('0.1 0 0\n0.2 1 0\n')
    with pytest.raises(ValueError):
        geometry.read_fault_profile(str(p), 3, 2)


def test_block_list_file(tmp_path):
    """'&block_list = 'file'' reads the block namelist from that file,
    relative to the input file's directory"""
    from wql.problem import from_namelist
    src = open(os.path.join(G.GOLDEN, 'inputs', 'cart_sw.in')).read()
    start = src.index('&block_list')
    end = src.index('/', src.index('btp(2)%faultsize'))
    (tmp_path / 'sub').mkdir()
    (tmp_path / 'sub' / 'blocks.txt').write_text(src[start:end + 1] + '\n')
    (tmp_path / 'case.in').write_text(src[:start] + "&block_list = 'sub/blocks.txt' /\n" + src[end + 1:])
    a = from_namelist(str(tmp_path / 'case.in'))
    b = G.problem('cart_sw')
    assert [blk.n for blk in a.blocks] == [blk.n for blk in b.blocks]
    assert [blk.bc for blk in a.blocks] == [blk.bc for blk in b.blocks]
    assert [blk.a for blk in a.blocks] == [blk.a for blk in b.blocks]
    (tmp_path / 'bad.in').write_text(src[:start] + "&block_list = 'missing.txt' /\n" + src[end + 1:])
    with pytest.raises(ValueError):
        from_namelist(str(tmp_path / 'bad.in'))


@pytest.mark.parametrize('edit', [("nblocks = 2", "nblocks = 3"),
                                  ("use_mms = F", "use_mms = T")])
def test_unsupported_inputs_fail(edit, tmp_path):
    """input features outside the path raise instead of running as something
    else"""
    from wql.problem import from_namelist
    src = open(os.path.join(G.GOLDEN, 'inputs', 'cart_sw.in')).read()
    assert edit[0] in src
    (tmp_path / 'x.in').write_text(src.replace(edit[0], edit[1], 1))
    with pytest.raises(NotImplementedError):
        from_namelist(str(tmp_path / 'x.in'))


def test_mollified_sources_parse(tmp_path):
    """mollify_source = T reaches the Problem (the sources then spread as
    Gaussians); the oracle, which restates the point form only, refuses it"""
    import oracle as O
    prob = G.problem('mt_smooth')
    assert prob.mollify_source and prob.moment
    assert not G.problem('mt_cart').mollify_source
    with pytest.raises(NotImplementedError):
        O.OracleDomain(prob)


def test_material_files_read(tmp_path):
    """material_source = 'file' fills each block's medium from its three files
    (init_material_from_file); a short file is an error"""
    z = G.load('matfile_sw')
    prob, inp = G.matfile_problem('matfile_sw', str(tmp_path))
    for ib, b in enumerate(prob.blocks):
        assert b.medium is None
        assert np.array_equal(b.lame(), z[f'b{ib + 1}_Mat'])
    (tmp_path / 'mat2_rho.bin').write_bytes((tmp_path / 'mat2_rho.bin').read_bytes()[:-8])
    from wql.problem import from_namelist
    with pytest.raises(ValueError):
        from_namelist(inp)
