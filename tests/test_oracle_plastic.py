"""Pins the oracle's plasticity pieces that no reference golden covers.

plastic_flow (plastic.f90:221-290, the default model): the closed-form
implicit update is fixed by what it solves.  With trial stress sa, tau the
root of J2 of its deviator and sigma its mean, yield Y = tau + mu sigma; when
Y > 0 the rate lambda solves eta lambda = Y(new stress) with
tau_new = tau - dt lambda G and sigma_new = sigma - dt lambda beta K, the
deviator keeps its direction, and gammap grows by dt lambda.  Below yield
nothing changes and lambda = 0.  The checks recompute tau and sigma of the
output tensor independently, so a dropped term, a wrong sign or a wrong
modulus fails them.

TPV26 background stress (initial_stress_condition.f90:26-76): a pore pressure
rho_w g y and lithostatic vertical stress -rho g y (rho = 2.67, rho_w = 1,
g = 9.8, y depth in km), a depth taper Omega that is 1 above 15 km, 0 below
20 km and linear between.  Checked: the effective vertical stress is
-(rho - rho_w) g y at every depth, the tensor is isotropic below 20 km,
it is continuous across 15 and 20 km and linear in y inside each piece,
and xy, yz shear vanish."""
import numpy as np
import pytest

import golden_cases as G
import oracle as O


def _domain(model, mu=0.3, beta=0.4, eta=0.05, s0_kind=0):
    prob = G.problem('plastic_sw')
    prob.plastic = model
    prob.pl_mu, prob.pl_beta, prob.pl_eta = mu, beta, eta
    prob.s0_kind = s0_kind
    return prob, O.OracleDomain(prob)


def _inv(sa):
    sig = (sa[0] + sa[1] + sa[2]) / 3.0
    d = sa.copy()
    d[:3] -= sig
    J2 = (d[0] ** 2 + d[1] ** 2 + d[2] ** 2) / 2.0 + d[3] ** 2 + d[4] ** 2 + d[5] ** 2
    return np.sqrt(J2), sig, d


@pytest.mark.parametrize('mu,beta,eta', [(0.3, 0.4, 0.05), (0.6, 0.0, 0.2), (0.1, 0.25, 0.0)])
def test_plastic_flow_return_map_invariants(mu, beta, eta):
    prob, od = _domain(2, mu, beta, eta)
    rng = np.random.default_rng(3)
    s0 = np.array(prob.fault.s0, dtype=float)
    for ib in range(2):
        F = od.blocks[ib]['F']
        F[3:] = rng.normal(0.0, 40.0, F[3:].shape)
        od.blocks[ib]['P'][1] = rng.uniform(0.0, 1.0, F.shape[1:])
    trial = [od.blocks[ib]['F'][3:].copy() for ib in range(2)]
    gam0 = [od.blocks[ib]['P'][1].copy() for ib in range(2)]
    od.plastic()
    dt = prob.dt()
    nyield = 0
    for ib in range(2):
        lam_mat = od.blocks[ib]['M']
        Gm = lam_mat[1]
        K = lam_mat[0] + 2.0 / 3.0 * lam_mat[1]
        sa_t = trial[ib] + s0[:, None, None, None]
        sa_n = od.blocks[ib]['F'][3:] + s0[:, None, None, None]
        lam = od.blocks[ib]['P'][0]
        gam = od.blocks[ib]['P'][1]
        tau_t, sig_t, d_t = _inv(sa_t)
        tau_n, sig_n, d_n = _inv(sa_n)
        Y = tau_t + mu * sig_t
        el = Y <= 0
        nyield += int((~el).sum())
        # below yield: unchanged
        assert np.array_equal(sa_n[:, el], sa_t[:, el])
        assert np.all(lam[el] == 0.0)
        assert np.array_equal(gam[el], gam0[ib][el])
        p = ~el
        scale = np.abs(sa_t[:, p]).max()
        assert np.all(lam[p] > 0.0)
        # the deviator is scaled by the signed factor f = (tau - dt lambda G) / tau
        # (the update does not clamp an overshooting return: f < 0 flips it)
        f = (tau_t[p] - dt * lam[p] * Gm[p]) / tau_t[p]
        assert np.allclose(d_n[:, p], f * d_t[:, p], rtol=0, atol=1e-12 * scale)
        assert np.allclose(tau_n[p], np.abs(f) * tau_t[p], rtol=0, atol=1e-12 * scale)
        # the mean stress moves by dt lambda beta K
        assert np.allclose(sig_n[p], sig_t[p] - dt * lam[p] * beta * K[p], rtol=0, atol=1e-12 * scale)
        # implicit consistency: eta lambda = Y of the updated (signed) invariants
        assert np.allclose(eta * lam[p], f * tau_t[p] + mu * sig_n[p], rtol=0, atol=1e-11 * scale)
        assert np.allclose(gam[p], gam0[ib][p] + dt * lam[p], rtol=1e-14, atol=0)
    assert nyield > 100  # the random states exercise both branches


def test_plastic_flow_check_fails_on_a_planted_mistake():
    """the consistency check above rejects an update that drops the
    mu beta K term of the denominator (the mistake is planted here, in a
    copy of the closed form, not in the oracle)"""
    mu, beta, eta, dt, Gm, K = 0.3, 0.4, 0.05, 0.01, 32.0, 53.0
    sa = np.array([-50.0, -40.0, -60.0, 30.0, 5.0, -8.0])
    tau, sig, d = _inv(sa)
    lam_bad = (tau + mu * sig) / (eta + dt * Gm)
    tau_n = tau - dt * lam_bad * Gm
    sig_n = sig - dt * lam_bad * beta * K
    assert abs(eta * lam_bad - (tau_n + mu * sig_n)) > 1e-3


def test_tpv26_background_stress_profile():
    prob, od = _domain(2, s0_kind=1)
    rho, rhow, g = 2.67, 1.0, 9.8
    ys = np.linspace(0.5, 30.0, 119)
    S = np.array([od.initial_stress(np.array([0.0, y, 0.0])) for y in ys])
    # effective vertical stress and vanishing xy, yz shear at every depth
    assert np.allclose(S[:, 1], -(rho - rhow) * g * ys, rtol=1e-14, atol=0)
    assert np.all(S[:, 3] == 0.0) and np.all(S[:, 5] == 0.0)
    # isotropic below 20 km
    deep = ys >= 20.0
    for c in (0, 2):
        assert np.allclose(S[deep, c], S[deep, 1], rtol=1e-14, atol=0)
    assert np.all(S[deep, 4] == 0.0)
    # continuous across 15 and 20 km, linear in y inside each piece
    for y0 in (15.0, 20.0):
        a = od.initial_stress(np.array([0.0, y0 - 1e-9, 0.0]))
        b = od.initial_stress(np.array([0.0, y0 + 1e-9, 0.0]))
        assert np.allclose(a, b, rtol=0, atol=1e-6)
    for lo, hi in ((1.0, 15.0), (20.0, 30.0)):
        y = np.array([lo, 0.5 * (lo + hi), hi])
        s = np.array([od.initial_stress(np.array([0.0, v, 0.0])) for v in y])
        assert np.allclose(s[1], 0.5 * (s[0] + s[2]), rtol=1e-13, atol=1e-12)
    # inside the taper the stress is quadratic in y (Omega linear times a
    # linear profile): the second difference is constant
    y = np.linspace(15.5, 19.5, 5)
    s = np.array([od.initial_stress(np.array([0.0, v, 0.0])) for v in y])
    d2 = s[2:] - 2 * s[1:-1] + s[:-2]
    assert np.allclose(d2, d2[0], rtol=1e-9, atol=1e-9)
    # the shallow tensor is anisotropic (the taper matters)
    assert abs(S[0, 0] - S[0, 1]) > 1e-3 and abs(S[0, 4]) > 1e-3


def test_scec_accumulator_is_the_shear_stress_drop():
    """model 1 ('scec', plastic_flow2): the reference keeps no plastic strain
    for it; the library's FP64 accumulator grows by the drop of the
    deviatoric stress invariant over the shear modulus, (tau_trial -
    tau_returned) / G, recomputed here from the returned stress, and lambda
    is that increment over dt; nodes below yield keep gammap and get
    lambda = 0"""
    prob, od = _domain(1)
    rng = np.random.default_rng(5)
    s0 = np.array(prob.fault.s0, dtype=float)
    for ib in range(2):
        F = od.blocks[ib]['F']
        F[3:] = rng.normal(0.0, 40.0, F[3:].shape)
        od.blocks[ib]['P'][1] = rng.uniform(0.0, 1.0, F.shape[1:])
    trial = [od.blocks[ib]['F'][3:].copy() for ib in range(2)]
    gam0 = [od.blocks[ib]['P'][1].copy() for ib in range(2)]
    od.plastic()
    dt = prob.dt()
    nyield = 0
    for ib in range(2):
        Gm = od.blocks[ib]['M'][1]
        tau_t = _inv(trial[ib] + s0[:, None, None, None])[0]
        tau_n = _inv(od.blocks[ib]['F'][3:] + s0[:, None, None, None])[0]
        lam, gam = od.blocks[ib]['P']
        moved = tau_n < tau_t * (1 - 1e-15)
        nyield += int(moved.sum())
        dg = (tau_t - tau_n) / Gm
        assert np.allclose(gam[moved] - gam0[ib][moved], dg[moved], rtol=1e-9, atol=0)
        assert np.allclose(lam[moved] * dt, dg[moved], rtol=1e-9, atol=0)
        assert np.array_equal(gam[~moved], gam0[ib][~moved]) and np.all(lam[~moved] == 0.0)
    assert nyield > 100
