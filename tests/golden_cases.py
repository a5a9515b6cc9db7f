"""Helpers shared by the parity tests: the golden cases of tests/golden/."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, 'tests', 'golden')
sys.path.insert(0, os.path.join(ROOT, 'waveqlab3d-gh_amd'))

from wql import problem as P  # noqa: E402

# the upwind and upwind-DRP operators (fd_type / order)
SBP_RHS_CASES = ['up2_cart', 'up4_curv', 'up6_pml', 'up9_curv', 'drp3_curv', 'drp66_cart', 'drp679_curv']
SBP_TRAJ_CASES = ['up4_curv', 'up6_pml', 'drp66_cart']
# the anelastic response (GSLS memory variables); their fixtures keep the
# rates with every term (DFfull) and the memory variables, not DFvol
AN_CASES = ['an_curv', 'an_pml', 'an_drp6_pml']
RHS_CASES = ['cart_sw', 'curv_sw', 'curv_rs', 'pml_locked', 'rw_plastic', 'tpv26_sw', 'mt_cart',
             'mt_curv'] + SBP_RHS_CASES + AN_CASES
TRAJ_CASES = ['cart_sw', 'curv_sw', 'curv_rs', 'curv_rs_long', 'pml_locked', 'plastic_sw', 'locked_pulse',
              'rw_plastic', 'tpv26_sw', 'tpv26_plastic', 'mt_cart', 'mt_curv'] + SBP_TRAJ_CASES + AN_CASES

# Relative-L2 bar for FP64 trajectories.  'curv_rs' stops before rupture:
# slip rates there are ~1e-12 m/s, equal to the absolute step tolerance of
# the reference's Regula Falsi (Interface_Condition.f90:1188), so its
# iteration count -- and the slip rate to within 1e-12 m/s -- is decided by
# last-bit rounding; fields near the fault are ~1e-11.  That window is held to
# 1e-7; the same law during rupture ('curv_rs_long', V ~ 3 m/s) is held to
# the FP64 bar.
TRAJ_TOL = {'curv_rs': 1e-7}


def load(case):
    return np.load(os.path.join(GOLDEN, f'{case}.npz'))


def problem(case):
    return P.from_namelist(os.path.join(GOLDEN, 'inputs', f'{case}.in'))


def pattern(shape_qrs, f1, blk, scale, i0=1):
    """oracle/golden_driver.f90:pat on a (nq, nr, ns) grid -> [ns][nr][nq]."""
    nq, nr, ns = shape_qrs
    k, j, i = np.meshgrid(np.arange(1, ns + 1), np.arange(1, nr + 1), np.arange(i0, i0 + nq),
                          indexing='ij')
    idx = i.astype(np.int64) + 1000 * j + 1000000 * k + 1000000000 * np.int64(f1)
    h = (idx * 7919 + 104729 * blk) % 1000003
    return scale * (2.0 * (h / 1000003.0) - 1.0)


def pattern_state(prob):
    """The state golden_driver's fill_pattern writes, per block:
    F [9][ns][nr][nq], PML Q per side [9][slab], S [4][ns][nr], W [1][ns][nr]."""
    out = []
    for ib, b in enumerate(prob.blocks):
        nq, nr, ns = b.n
        F = np.stack([pattern(b.n, f + 1, ib + 1, 10.0 if f >= 3 else 1.0) for f in range(9)])
        Q = [None] * 6
        for s in range(6):
            if not b.pml[s]:
                continue
            lo, hi = [1, 1, 1], [nq, nr, ns]
            d = s // 2
            if s % 2 == 0:
                hi[d] = b.npml
            else:
                lo[d] = b.n[d] - b.npml + 1
            kk, jj, ii = np.meshgrid(np.arange(lo[2], hi[2] + 1), np.arange(lo[1], hi[1] + 1),
                                     np.arange(lo[0], hi[0] + 1), indexing='ij')
            q = []
            for f in range(9):
                idx = ii.astype(np.int64) + 1000 * jj + 1000000 * kk + 1000000000 * np.int64(f + 1 + 10 * (s + 1))
                h = (idx * 7919 + 104729 * (ib + 3)) % 1000003
                q.append((0.1 * (2.0 * (h / 1000003.0) - 1.0)).ravel())
            Q[s] = np.stack(q)
        kk, jj = np.meshgrid(np.arange(1, ns + 1), np.arange(1, nr + 1), indexing='ij')
        S = []
        for f in range(4):
            idx = 1000 * jj.astype(np.int64) + 1000000 * kk + 1000000000 * np.int64(f + 1)
            h = (idx * 7919 + 104729 * 7) % 1000003
            S.append(0.25 * ((2.0 * (h / 1000003.0) - 1.0) + 1.0))
        idx = 1000 * jj.astype(np.int64) + 1000000 * kk
        h = (idx * 7919 + 104729 * 8) % 1000003
        W = (0.6 + 0.05 * (2.0 * (h / 1000003.0) - 1.0))[None]
        eta = None
        if getattr(prob, 'anelastic', None):
            # memory variables: component c (sxx..syz), mechanism l (golden_driver fill_pattern)
            eta = np.stack([pattern(b.n, 31 + f, ib + 5, 0.05) for f in range(24)])
        out.append(dict(F=F, Q=Q, S=np.stack(S), W=W, eta=eta))
    return out


def pulse(prob, X, args):
    """Gaussian velocity pulse added by golden_driver's add_pulse."""
    x0, y0, z0, sig, amp = args
    r2 = (X[0] - x0) ** 2 + (X[1] - y0) ** 2 + (X[2] - z0) ** 2
    return amp * np.exp(-r2 / (2.0 * sig ** 2))


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


GRID_CASES = ['grid_tpv28', 'grid_tpv10', 'grid_tpv36', 'grid_fractal', 'grid_topo_read', 'grid_topo_analytic',
              'grid_topo_fractal']


def grid_problem(case, tmpdir):
    """the Problem of a fault-profile grid case; the profile-file case gets its
    file (the stored x column, rows 'x y z' with y fastest) written beside a
    copy of its input file in tmpdir"""
    src = os.path.join(GOLDEN, 'inputs', f'{case}.in')
    z = load(case)
    if 'profile_x' not in z and 'topo1' not in z:
        return P.from_namelist(src)
    if 'profile_x' in z:
        x = z['profile_x']
        with open(os.path.join(tmpdir, 'fault.dat'), 'w') as f:
            if x.ndim == 3:  # x, y, z of the topography branch's profile file
                for k in range(x.shape[2]):
                    for j in range(x.shape[1]):
                        f.write(f'{x[0, j, k]:.17e} {x[1, j, k]:.17e} {x[2, j, k]:.17e}\n')
            else:
                nr, ns = x.shape
                for k in range(ns):
                    for j in range(nr):
                        f.write(f'{x[j, k]:.17e} {float(j):.1f} {float(k):.1f}\n')
    for b in (1, 2):
        if f'topo{b}' in z:  # read_topo_from_memory files: rows x y z, z fastest
            t = z[f'topo{b}']
            with open(os.path.join(tmpdir, f'topo{b}.dat'), 'w') as f:
                for i in range(t.shape[1]):
                    for k in range(t.shape[2]):
                        f.write(f'{t[0, i, k]:.17e} {t[1, i, k]:.17e} {t[2, i, k]:.17e}\n')
    dst = os.path.join(tmpdir, f'{case}.in')
    with open(src) as a, open(dst, 'w') as b:
        b.write(a.read())
    return P.from_namelist(dst)


# the slip-weakening problems the reference keys by name (per-node fault
# tables) and init_material's layered media
KEY_RHS_CASES = ['tpv10_sw', 'tpv29_sw', 'tpv31_sw', 'tpv32_sw', 'tpv33_sw', 'tpv34_sw', 'scits_sw', 'tpv36_sw',
                 'tpv37_sw', 'wasatch_sw']
KEY_TRAJ_CASES = ['tpv10_sw', 'tpv29_sw', 'tpv31_sw', 'tpv32_sw', 'tpv33_sw', 'tpv34_sw', 'scits_sw', 'tpv36_sw',
                  'wasatch_sw', 'tpv30_plastic', 'loh1_locked', 'oklahoma_locked']
# the Gaussian-mollified moment-tensor sources (mollify_source = T)
SMOOTH_CASES = ['mt_smooth', 'mt_smooth_lists']
MEDIUM_CASES = ['tpv31_sw', 'tpv32_sw', 'tpv33_sw', 'loh1_locked', 'oklahoma_locked']


# a block medium read from files (material_source = 'file')
MATFILE_CASES = ['matfile_sw']


def matfile_problem(case, tmpdir):
    """(Problem, input path) of a material-file case: the three files of each
    block (raw FP64 lambda, mu, rho, q fastest) written from the fixture's
    medium beside a copy of its input file in tmpdir"""
    z = load(case)
    for b in (1, 2):
        for c, tag in enumerate(('lam', 'mu', 'rho')):
            np.ascontiguousarray(z[f'b{b}_Mat'][c]).astype('<f8').tofile(os.path.join(tmpdir, f'mat{b}_{tag}.bin'))
    dst = os.path.join(tmpdir, f'{case}.in')
    with open(os.path.join(GOLDEN, 'inputs', f'{case}.in')) as a, open(dst, 'w') as f:
        f.write(a.read())
    return P.from_namelist(dst), dst
