"""Pins of the friction laws added beside the reference's TPV5/28/101 set.

The slip law with flash weakening ('rate-weakening', Interface_Condition.f90:
438-444), the Gaussian prestress pulse (2387-2420), the constant initial state
(2820-2850), the TPV26/27 strength (1376-1458) and the TPV26 prestress tensor
(2008-2048) are pinned by reference-build trajectories (tests/golden:
rw_plastic, tpv26_sw, tpv26_plastic; test_oracle_golden.py).  What those
cannot pin is checked here:

* KDN (446-449): the reference keys its friction and state law but stops in
  `prestress` for it (2425-2427), so no reference run exists.  With the
  direct-effect parameter a uniform and flash weakening off, the KDN and
  slip-law state rates of one evaluation differ by exactly
  -(V/L)(1 - a) asinh(V/(2 V0) exp(psi/a)) (the steady-state friction Fss
  cancels), which ties KDN to the golden-pinned slip law.  (With V from the
  Newton solve, the hat slip rate that weights KDN equals V to rounding.)
* The fast slip-rate mode (bracketed Newton, no counterpart in the
  reference): it solves the same equation to a residual at FP64 rounding in
  creep and in rupture, and its trajectories stay within 1e-6 (relative L2)
  of the reference's regula-falsi trajectories, whose own stopping rule (an
  absolute 1e-12 step, 1187-1237) leaves them ~3e-7 from the root.
* The TPV26 strength at its limits: before the forced-rupture time with no
  slip a node holds C0 + fs sigma_n; after it the strength is C0 + fd sigma_n,
  C0 = 0.4 + 0.72 (5 - y) above 5 km depth and 0.4 below.
"""
import ctypes as C

import numpy as np
import pytest

import golden_cases as G
import gpu_helpers as H
import oracle as O
from wql import problem as P

FRAME = np.array([1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0])
MAT = np.array([32.04, 32.04, 2.67])  # lambda, mu, rho


def _hat(fault, u, v, slip=0.0, psi=0.6, x=(0.0, 7.5, 7.5), t=0.0):
    fs = O.fault_struct(fault)
    uo, vo, Vel = np.zeros(6), np.zeros(6), C.c_double()
    dp = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    xa, ua, va = np.array(x, float), np.array(u, float), np.array(v, float)
    rc = O.lib().ora_hat(C.byref(fs), dp(ua), dp(va), dp(MAT), dp(MAT), dp(xa), dp(FRAME), slip, psi, t,
                         uo.ctypes.data_as(C.POINTER(C.c_double)), vo.ctypes.data_as(C.POINTER(C.c_double)),
                         C.byref(Vel))
    return rc, uo, vo, Vel.value


def _eta():
    rho, mu = MAT[2], MAT[1]
    Zs = rho * np.sqrt(mu / rho)
    return 1.0 / (2.0 / Zs)


@pytest.mark.parametrize('psi,scale', [(0.3, 0.1), (0.45, 0.1), (0.75, 1e-9), (0.8, 1e-11)])
def test_newton_solves_the_slip_rate_equation(psi, scale):
    """Phi - eta V = sigma_n a asinh(V/(2 V0) exp(psi/a)) to FP64 rounding,
    in rupture (V ~ 1) and in creep (V down to ~1e-12)"""
    f = P.fault_params('TPV101', 'rate-and-state_friction')
    f.rs_solver = 1
    fs = O.fault_struct(f)
    eta = _eta()
    rng = np.random.default_rng(7)
    u, v = np.zeros(6), np.zeros(6)
    u[[1, 2]] = rng.normal(size=2) * scale
    rc, uo, vo, V = _hat(f, u, v, psi=psi, x=(0.0, 7.5, 7.5))
    assert rc == 0 and V > 0.0
    a, sigma = fs.rs_a, 120.0
    tau = np.hypot(uo[4] + 0.0, uo[5] + 75.0)
    fv = sigma * a * np.arcsinh(0.5 * V / fs.rs_V0 * np.exp(psi / a))
    Phi = tau + eta * V
    assert abs(tau - fv) <= 1e-13 * Phi


@pytest.mark.parametrize('case', ['curv_rs_long', 'rw_plastic'])
def test_newton_trajectory_tracks_reference(case):
    z = G.load(case)
    prob = G.problem(case)
    prob.fault.rs_solver = 1
    od = O.OracleDomain(prob)
    args = z['traj_args']
    F0 = H.pulse_init(prob, args)
    for ib in range(2):
        od.blocks[ib]['F'][...] = F0[ib]
    od.init_state()
    od.step(int(args[0]))
    for ib in range(2):
        assert G.rel_l2(od.blocks[ib]['F'], z[f'b{ib + 1}_F']) <= 1e-6
        assert G.rel_l2(od.blocks[ib]['iface']['S'], z[f'b{ib + 1}_S']) <= 1e-6


def _rs_fault(rs_law):
    f = P.fault_params('rate-weakening', 'rate-and-state_friction')
    f.rs_da = 0.0           # a uniform
    f.rs_Vw = 0.0           # no flash weakening
    f.rs_law = rs_law
    # the Newton solve: its V solves the slip-rate equation to rounding, so
    # the hat slip rate (KDN's weight) equals V (the slip law's)
    f.rs_solver = 1
    return f


def test_kdn_state_rate_against_slip_law():
    prob = G.problem('rw_plastic')
    prob.plastic = 0
    dw, ds = {}, {}
    for law in (1, 2):
        prob.fault = _rs_fault(law)
        od, _, _ = H.oracle_rhs(prob, 0.3)
        dw[law] = [od.blocks[ib]['iface']['DW'][0].copy() for ib in range(2)]
        ds[law] = [od.blocks[ib]['iface']['DS'][3].copy() for ib in range(2)]
    f = _rs_fault(2)
    st = G.pattern_state(prob)
    for ib in range(2):
        V = ds[1][ib]
        assert np.array_equal(V, ds[2][ib])  # the same solve: the state law does not enter it
        W = st[ib]['W'][0]
        A = np.arcsinh(V / (2.0 * f.rs_V0) * np.exp(W / f.rs_a))
        want = -(V / f.rs_L) * (1.0 - f.rs_a) * A
        got = dw[2][ib] - dw[1][ib]
        assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()
        assert np.abs(want).max() > 1.0


def test_tpv26_strength_limits():
    """strength C0 + fd sigma_n once weakened, by slip past Dc at any depth or
    by the forced rupture at the nucleation centre; C0 + fs sigma_n holds a
    quiet node before either"""
    f = P.fault_params('TPV26', 'slip-weakening_friction')
    eta = _eta()

    def expect(y, x, slip, t, weakened):
        Omega = 1.0 if y <= 15 else (20.0 - y) / 5.0
        Pf, s22 = 9.8 * y, -2.67 * 9.8 * y
        sxx = Omega * (1.073206 * (s22 + Pf) - Pf) + (1 - Omega) * s22 + Pf
        sxz = Omega * (-0.169029 * (s22 + Pf))
        sigma, tau0 = -sxx, abs(sxz)     # normal (1, 0, 0): T0_n = sxx, T0_l = sxz
        C0 = 0.4 + 0.72 * (5.0 - y) if y <= 5.0 else 0.4
        rc, uo, vo, V = _hat(f, np.zeros(6), np.zeros(6), x=x, slip=slip, t=t)
        assert rc == 0
        fr = 0.12 if weakened else 0.18
        strength = C0 + fr * sigma
        if tau0 > strength:
            assert np.hypot(uo[4], uo[5] + sxz) == pytest.approx(strength, rel=1e-12)
            assert V == pytest.approx((tau0 - strength) / eta, rel=1e-10)
            return True
        assert V == 0.0
        return False

    slipped = 0
    for y in (3.0, 10.0, 17.0):
        assert not expect(y, (0.0, y, 14.0), 0.0, 0.0, False)
        slipped += expect(y, (0.0, y, 14.0), 0.5, 0.0, True)
    assert slipped >= 2
    assert not expect(10.0, (0.0, 10.0, -5.0), 0.0, 0.0, False)
    assert expect(10.0, (0.0, 10.0, -5.0), 0.0, 5.0, True)
