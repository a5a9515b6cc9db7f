"""The benchmark workloads as Problems (seeded synthetic inputs).

Units follow the reference (km, km/s, g/cm^3, GPa, MPa, s).  Each builder
takes a `scale` that shrinks the node counts (for tests) while keeping h,
physics and seeds; scale=1 gives the full workload.

  plain    two Cartesian 31x61x61 blocks (one 61^3 body split at a locked
           interface), Gaussian velocity pulse, free surface + absorbing SAT
  rupture  two curvilinear 201x101x101 blocks, planar vertical fault, linear
           slip weakening, Gaussian-random-field topography on the free surface
  rough    two 401x201x201 blocks, self-similar rough fault, rate and state
           (slip law, FP64 Newton solve), 10-node PML on every outer side
           except the free surface
  plastic  two Cartesian 301x151x151 blocks, von Karman heterogeneous medium,
           depth-dependent background stress on the fault and in the
           Drucker-Prager viscoplasticity, TPV26-style slip weakening
  serving  per-scenario two-block problems drawn from {61^3, 121x61x61,
           201x101x101, 241x121x121} with seed-drawn friction
  pml12    the configuration of the reference's PML and anelasticity study
           (simulation/PML-12/input/*.in): two Cartesian 76x101x151 blocks,
           locked interface, 12-node PML, a moment-tensor point source, the
           anelastic response with Q_s = 2 v_s; the operator, the PML and the
           response are switched per deck
"""
import math

import numpy as np

from .problem import LOCKED, RATE_STATE, SLIP_WEAKENING, Block, Fault, Problem

RHO, CS, CP = 2.67, 3.464, 6.0


def _scaled(n, scale):
    return tuple(max(13, int(round((v - 1) * scale)) + 1) for v in n)


def grf2d(shape, h, corr, rms, seed):
    """Gaussian random field on a shape=(n0, n1) grid with spacing h: Gaussian
    covariance of correlation length `corr`, standard deviation `rms`."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape)
    k0 = np.fft.fftfreq(shape[0], d=h) * 2 * np.pi
    k1 = np.fft.fftfreq(shape[1], d=h) * 2 * np.pi
    K2 = k0[:, None] ** 2 + k1[None, :] ** 2
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * np.exp(-K2 * corr ** 2 / 8.0)))
    f -= f.mean()
    s = f.std()
    return f * (rms / s) if s > 0 else f


def self_similar_surface(shape, h, ratio, lam_min, seed):
    """Band-limited self-similar roughness: power spectrum ~ k^-4 (2-D,
    Hurst 1) between 2*pi/L and 2*pi/lam_min, rms amplitude ratio * L."""
    rng = np.random.default_rng(seed)
    L = h * (max(shape) - 1)
    k0 = np.fft.fftfreq(shape[0], d=h) * 2 * np.pi
    k1 = np.fft.fftfreq(shape[1], d=h) * 2 * np.pi
    K = np.sqrt(k0[:, None] ** 2 + k1[None, :] ** 2)
    amp = np.zeros_like(K)
    band = (K >= 2 * np.pi / L) & (K <= 2 * np.pi / lam_min)
    amp[band] = K[band] ** -2.0
    spec = amp * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    f = np.real(np.fft.ifft2(spec))
    f -= f.mean()
    s = f.std()
    return f * (ratio * L / s) if s > 0 else f


def von_karman3d(shape, h, corr, hurst, std, seed):
    """3-D von Karman random field [n0][n1][n2] with correlation lengths corr
    (per axis), Hurst exponent, unit-mean relative perturbation of std."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape)
    ks = [np.fft.fftfreq(n, d=h) * 2 * np.pi for n in shape]
    K2 = (ks[0][:, None, None] * corr[0]) ** 2 + (ks[1][None, :, None] * corr[1]) ** 2 + \
         (ks[2][None, None, :] * corr[2]) ** 2
    P = (1.0 + K2) ** (-(hurst + 1.5) / 2.0)
    f = np.real(np.fft.ifftn(np.fft.fftn(w) * P))
    f -= f.mean()
    s = f.std()
    return f * (std / s) if s > 0 else f


def _blocks(n, h, top=None, fault=None, mesh='curvilinear', pml=False, npml=10, ly_code=2):
    """two blocks either side of x = 0 with n nodes each; y down from the free
    surface y = 0, z centred.  `top` [2*nq-1][ns] y-offsets of the free
    surface over both blocks; `fault` [nr][ns] x-offsets of the fault."""
    nq, nr, ns = n
    Lx, Ly, Lz = (nq - 1) * h, (nr - 1) * h, (ns - 1) * h
    blocks = []
    for ib, (a0, a1) in enumerate(((-Lx, 0.0), (0.0, Lx))):
        bc = (1, 0, ly_code, 1, 1, 1) if ib == 0 else (0, 1, ly_code, 1, 1, 1)
        pm = (False,) * 6
        if pml:
            pm = (True, False, False, True, True, True) if ib == 0 else (False, True, False, True, True, True)
        b = Block(n=n, a=(a0, 0.0, -Lz / 2), b=(a1, Ly, Lz / 2), rho_s_p=(RHO, CS, CP), bc=bc, pml=pm,
                  npml=npml if pml else 0, mesh=mesh, profile='flat')
        blocks.append(b)
    coords = []
    for ib, b in enumerate(blocks):
        X = np.empty((3, ns, nr, nq))
        x = b.a[0] + np.arange(nq) * h
        y = np.arange(nr) * h
        z = b.a[2] + np.arange(ns) * h
        X[0] = x[None, None, :]
        X[1] = y[None, :, None]
        X[2] = z[:, None, None]
        if fault is not None:
            # x' = x + r(y,z) (1 - |x|/Lx): the fault surface r is shared by both blocks
            X[0] += fault.T[:, :, None] * (1.0 - np.abs(X[0]) / Lx)
        if top is not None:
            t = top[ib * (nq - 1): ib * (nq - 1) + nq, :]  # [nq][ns]
            X[1] += t.T[:, None, :] * (1.0 - X[1] / Ly)
        coords.append(X)
    return blocks, coords


def plain(scale=1.0, seed=1):
    """Plain method (FP64 reference case): 61^3 body, Gaussian pulse."""
    n = _scaled((31, 61, 61), scale)
    h = 0.1
    blocks, coords = _blocks(n, h, mesh='cartesian')
    prob = Problem(blocks=blocks, fault=Fault(law=LOCKED), cfl=0.5, name='plain')
    rng = np.random.default_rng(seed)
    Lx, Ly, Lz = (n[0] - 1) * h, (n[1] - 1) * h, (n[2] - 1) * h
    sig = 0.3
    c = (rng.uniform(-Lx + 3 * sig, Lx - 3 * sig), rng.uniform(3 * sig, Ly - 3 * sig),
         rng.uniform(-Lz / 2 + 3 * sig, Lz / 2 - 3 * sig))
    F0 = []
    for X in coords:
        F = np.zeros((9,) + X.shape[1:])
        r2 = (X[0] - c[0]) ** 2 + (X[1] - c[1]) ** 2 + (X[2] - c[2]) ** 2
        F[0:3] = np.exp(-r2 / (2 * sig ** 2))
        F0.append(F)
    return dict(problem=prob, coords=coords, F0=F0, steps=200, precision='fp64')


def sw_fault(mu_s=0.677, mu_d=0.525, Dc=0.4, yc=None, zc=0.0, depth=10.0, width=10.0):
    yc = depth / 2 if yc is None else yc
    return Fault(law=SLIP_WEAKENING, stress_mode=0, T0=(-120.0, 0.0, 70.0),
                 patches=[(yc, zc, 1.5, 1.5, 11.6)], fs=mu_s, fd=mu_d, Dc=Dc,
                 bar_wy=0.8 * depth, bar_wz=0.4 * width, bar_big=1.0e10)


def rupture(scale=1.0, seed=2):
    """Curvilinear rupture: planar fault, slip weakening, GRF topography."""
    n = _scaled((201, 101, 101), scale)
    h = 0.1
    top = grf2d((2 * n[0] - 1, n[2]), h, 2.0, 0.05, seed)
    blocks, coords = _blocks(n, h, top=top)
    Ly, Lz = (n[1] - 1) * h, (n[2] - 1) * h
    prob = Problem(blocks=blocks, fault=sw_fault(depth=Ly, width=Lz), cfl=0.5, name='rupture')
    return dict(problem=prob, coords=coords, F0=None, steps=1500, precision='fp32')


def rough(scale=1.0, seed=3):
    """Rough fault, rate and state, PML on the outer sides."""
    n = _scaled((401, 201, 201), scale)
    h = 0.05
    fault = self_similar_surface((n[1], n[2]), h, 0.01, 4 * h, seed)
    # taper the roughness to zero at the free surface and bottom
    y = np.arange(n[1]) * h
    Ly = (n[1] - 1) * h
    fault *= np.sin(np.pi * y / Ly)[:, None] ** 2
    blocks, coords = _blocks(n, h, fault=fault, pml=True, npml=10)
    Lz = (n[2] - 1) * h
    # TPV101 friction constants with the slip law (rs_law 1, flash weakening
    # off) and the FP64 Newton slip-rate solve
    f = Fault(law=RATE_STATE, stress_mode=1, s0=(-120.0, 0.0, -120.0, 0.0, 75.0, 0.0), nuc=2,
              nucp=(Ly / 2, 0.0, min(3.0, 0.3 * Lz), 25.0, 1.0),
              rs_a=0.008, rs_da=0.008, rs_w=0.4 * Lz, rs_wr=0.1 * Lz, rs_b=0.012, rs_L=0.02,
              rs_V0=1.0e-6, rs_f0=0.6, rs_Vin=1.0e-12, rs_law=1, rs_solver=1)
    prob = Problem(blocks=blocks, fault=f, cfl=0.5, name='rough')
    return dict(problem=prob, coords=coords, F0=None, steps=3000, precision='fp32', init_state=True)


def plastic(scale=1.0, seed=4):
    """Drucker-Prager viscoplasticity in a von Karman medium."""
    n = _scaled((301, 151, 151), scale)
    h = 0.05
    blocks, coords = _blocks(n, h, mesh='cartesian')
    for ib, b in enumerate(blocks):
        nq, nr, ns = n
        dv = von_karman3d((ns, nr, nq), h, (1.0, 5.0, 5.0), 0.2, 0.05, seed + ib)
        rho = RHO * (1 + dv)
        cs = CS * (1 + dv)
        cp = CP * (1 + dv)
        M = np.empty((3, ns, nr, nq))
        M[0] = rho * (cp ** 2 - 2 * cs ** 2)
        M[1] = rho * cs ** 2
        M[2] = rho
        b.material = M
        b.mesh = 'cartesian'
    Ly, Lz = (n[1] - 1) * h, (n[2] - 1) * h
    # TPV26-style fault: the fault resolves the same depth-dependent stress
    # tensor as the plastic background (stress_mode 2), TPV26 friction with
    # cohesion, and forced rupture nucleating at the fault centre
    f = Fault(law=SLIP_WEAKENING, stress_mode=2, fs=0.18, fd=0.12, Dc=0.3,
              bar_wy=0.8 * Ly, bar_wz=0.4 * Lz, bar_big=1.0e10, sw_law=1, coh=(0.4, 0.72, 5.0),
              frc=(0.5 * Ly, 0.0, min(4.0, 0.3 * min(Ly, Lz)), CS, 0.5))
    prob = Problem(blocks=blocks, fault=f, cfl=0.5, plastic=1, pl_c=1.0, pl_nu=0.75, pl_Tv=0.05,
                   s0_kind=1, name='plastic')
    # the time step uses the fastest material
    prob.cfl = 0.5 / 1.15
    return dict(problem=prob, coords=coords, F0=None, steps=2000, precision='fp32')


SERVING_SIZES = [(61, 61, 61), (121, 61, 61), (201, 101, 101), (241, 121, 121)]


def scenario(k, seed=5, scale=1.0):
    """Serving scenario k of the stream: size and friction drawn from the seed."""
    rng = np.random.default_rng([seed, k])
    n = _scaled(SERVING_SIZES[int(rng.integers(len(SERVING_SIZES)))], scale)
    h = 0.1
    blocks, coords = _blocks(n, h, mesh='cartesian')
    Ly, Lz = (n[1] - 1) * h, (n[2] - 1) * h
    mu_s = rng.uniform(0.6, 0.7)
    mu_d = rng.uniform(0.45, 0.55)
    Dc = rng.uniform(0.2, 0.6)
    yc = rng.uniform(0.3 * Ly, 0.7 * Ly)
    zc = rng.uniform(-0.2 * Lz, 0.2 * Lz)
    f = sw_fault(mu_s, mu_d, Dc, yc=yc, zc=zc, depth=Ly, width=Lz)
    # nucleation stress just above the drawn static strength
    f.patches = [(yc, zc, 1.5, 1.5, 120.0 * mu_s - 70.0 + 0.5)]
    prob = Problem(blocks=blocks, fault=f, cfl=0.5, name=f'scenario{k}')
    return dict(problem=prob, coords=coords, F0=None, steps=1000, precision='fp32')


def pml12(scale=1.0, fd_type='traditional', order=6, pml=True, anelastic=True):
    """The decks of the reference's PML study (simulation/PML-12/input/
    <fd>_6_pml-<on|off>_<elastic|anelastic-c2>.in): two Cartesian blocks of
    76x101x151 nodes over 15 x 20 x 30 km each (h = 0.2 km in q and r), a
    locked interface at x = 15 km, the free surface at y = 0 and absorbing
    SAT on the other outer faces, with `pml` a 12-node PML on each of them;
    one order-4 Cos-Bell moment-tensor source (M_xz = -10, t_init 0.2 s) at
    (5, 1.8, 5) km in block 1; with `anelastic` Q_s = 2 v_s, weight_exp 0.6,
    fref 1 Hz; 8 s of simulated time at CFL 0.5."""
    n = _scaled((76, 101, 151), scale)
    h = 0.2
    L = tuple((v - 1) * h for v in n)
    blocks = []
    for ib in range(2):
        bc = (1, 0, 2, 1, 1, 1) if ib == 0 else (0, 1, 2, 1, 1, 1)
        pm = (True, False, False, True, True, True) if ib == 0 else (False, True, False, True, True, True)
        blocks.append(Block(n=n, a=(ib * L[0], 0.0, 0.0), b=((ib + 1) * L[0], L[1], L[2]),
                            rho_s_p=(2.7, 3.464, 6.0), bc=bc, pml=pm if pml else (False,) * 6,
                            npml=12 if pml else 0, mesh='cartesian'))
    src = dict(type=3, duration=0.0, t_init=0.2, m=(0.0, 0.0, 0.0, 0.0, -10.0, 0.0), x=(5.0, 1.8, 5.0), alpha=0)
    prob = Problem(blocks=blocks, fault=Fault(law=LOCKED), cfl=0.5, t_final=8.0, name='pml12',
                   moment=dict(order=4, preference=1, lists={'U': [src], 'V': []}),
                   fd_type=fd_type, order=order, anelastic=(2.0, 0.6, 1.0) if anelastic else None)
    return dict(problem=prob, coords=[b.coords() for b in blocks], F0=None, steps=prob.nt(), precision='fp32')


WORKLOADS = {'plain': plain, 'rupture': rupture, 'rough': rough, 'plastic': plastic, 'pml12': pml12}


def points(prob):
    return sum(b.n[0] * b.n[1] * b.n[2] for b in prob.blocks)
