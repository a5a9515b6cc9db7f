"""HIP path vs the oracle and the reference golden vectors (FP64 mode).

Tolerance: relative L2 <= 1e-12 in FP64 (the reference's own bar for the
FP64 path); absolute for quantities that are zero in exact arithmetic."""
import sys

import numpy as np
import pytest

import golden_cases as G
import gpu_helpers as H
import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _sim(prob, **kw):
    from wql.solver import Simulation
    return Simulation(prob, **kw)


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def _nvariants():
    from wql import native
    return native.lib().wql_tile_variants()


@pytest.mark.parametrize('case', [c for c in G.RHS_CASES if c not in G.SBP_RHS_CASES + G.AN_CASES])
@pytest.mark.parametrize('order', [0, 1])
@pytest.mark.parametrize('tile', list(range(7)))
def test_rhs_vs_golden(case, tile, order):
    """every launch variant, q-tiles fastest (order 0) and s-chunks fastest
    (order 1), against the reference build's rates"""
    _say(f'[gpu] rhs {case} tile={tile} order={order}')
    z = G.load(case)
    prob = G.problem(case)
    t = float(z['rhs_t'])
    with _sim(prob, tile=tile, zchunk=5, order=order) as sim:
        H.load_pattern_state(sim, prob)
        sim.eval_rates(t, what=0)
        for ib in range(2):
            assert G.rel_l2(sim.rates(ib), z[f'b{ib + 1}_DFvol']) <= TOL
            for s in range(6):
                key = f'b{ib + 1}_DQvol{s + 1}'
                if key in z:
                    got = sim.pml_rates(ib, s).reshape(9, -1)
                    ref = z[key].reshape(9, -1)
                    assert G.rel_l2(got, ref) <= TOL, key
        H.load_pattern_state(sim, prob)
        sim.eval_rates(t, what=1, stage=2)
        for ib in range(2):
            assert G.rel_l2(sim.rates(ib), z[f'b{ib + 1}_DFfull']) <= TOL
            DS, DW = sim.fault_rates(ib)
            ref = z[f'b{ib + 1}_DS']
            assert np.abs(DS - ref).max() <= TOL * max(1.0, np.abs(ref).max())
            ref = z[f'b{ib + 1}_DW']
            assert np.abs(DW - ref).max() <= TOL * max(1.0, np.abs(ref).max())


oracle_runs = {}


def _oracle_traj(case):
    """the oracle's blocks after a golden trajectory case"""
    z = G.load(case)
    prob = G.problem(case)
    args = z['traj_args']
    od = O.OracleDomain(prob)
    F0 = H.pulse_init(prob, args)
    for ib in range(2):
        od.blocks[ib]['F'][...] = F0[ib]
    od.init_state()
    od.step(int(args[0]))
    return od.blocks


@pytest.mark.parametrize('tile', [None, 0])
@pytest.mark.parametrize('case', G.TRAJ_CASES)
def test_trajectory_vs_golden(case, tile):
    """tile None: the precision's default launch shape; 0: the fused
    variant that keeps the q-face nodes in compact buffers between stages"""
    _say(f'[gpu] traj {case} tile={tile}')
    z = G.load(case)
    prob = G.problem(case)
    args = z['traj_args']
    with _sim(prob, tile=tile) as sim:
        assert sim.dt == pytest.approx(float(z['dt']), rel=1e-15)
        F0 = H.pulse_init(prob, args)
        for ib in range(2):
            sim.set_fields(ib, F0[ib])
        sim.init_state()
        sim.step(int(args[0]))
        sim.sync()
        assert sim.t == pytest.approx(float(z['t_end']), rel=1e-14)
        for ib in range(2):
            F = sim.fields(ib)
            tol = G.TRAJ_TOL.get(case, TOL)
            assert G.rel_l2(F, z[f'b{ib + 1}_F']) <= tol, f'F block {ib + 1}'
            f = sim.fault(ib)
            for key in ('S', 'W', 'Svel', 'trup'):
                ref = z[f'b{ib + 1}_{key}']
                scale = max(1.0, np.abs(ref).max())
                assert np.abs(f[key] - ref).max() <= max(1e-10, tol) * scale, key
            if f'b{ib + 1}_P' in z and prob.plastic == 2:
                # plastic_flow's lambda and gammap (P%P(:,:,:,1:2)) against the reference
                assert G.rel_l2(sim.plastic(ib), z[f'b{ib + 1}_P']) <= tol, f'P block {ib + 1}'
            elif prob.plastic == 1:
                # the 'scec' model has no accumulator in the reference: the
                # library's FP64 one against the oracle's, on the same run
                if case not in oracle_runs:
                    oracle_runs[case] = _oracle_traj(case)
                P_ref = oracle_runs[case][ib]['P']
                assert G.rel_l2(sim.plastic(ib), P_ref) <= tol, f'P block {ib + 1}'
            for s in range(6):
                key = f'b{ib + 1}_Q{s + 1}'
                if key in z:
                    # wql_get_pml's documented layout [9][slab nodes]: the
                    # reference's Q(i,j,k,f) in column-major order
                    got = sim.pml(ib, s).reshape(9, -1)
                    assert G.rel_l2(got, z[key].reshape(9, -1)) <= TOL, key
            if f'b{ib + 1}_eta' in z:
                # anelastic memory variables after the run
                assert G.rel_l2(sim.anelastic(ib), z[f'b{ib + 1}_eta']) <= tol, f'eta block {ib + 1}'


def test_metrics_and_frames_vs_golden():
    _say('[gpu] metrics curv_sw')
    z = G.load('curv_sw')
    prob = G.problem('curv_sw')
    with _sim(prob) as sim:
        for ib in range(2):
            G_ref = z[f'b{ib + 1}_metric']
            assert G.rel_l2(sim.metrics(ib), G_ref) <= 1e-13
            # face frames: reference dumps include 3 ghost nodes each side
            nq, nr, ns = prob.blocks[ib].n
            nml = z[f'b{ib + 1}_nml']
            off = 0
            for side in range(6):
                a, b = [(nr, ns), (nr, ns), (nq, ns), (nq, ns), (nq, nr), (nq, nr)][side]
                size = (a + 6) * (b + 6) * 3
                parts = [nml[off + k * size: off + (k + 1) * size].reshape(3, b + 6, a + 6)[:, 3:-3, 3:-3]
                         for k in range(3)]
                off += 3 * size
                ref = np.concatenate([p.reshape(3, -1) for p in parts])
                got = sim.frame(ib, side)
                assert np.abs(got - ref).max() <= 1e-13, f'frame block {ib + 1} side {side + 1}'


@pytest.mark.parametrize('n', [(41, 33, 29), (64, 40, 37)])
@pytest.mark.parametrize('mesh', ['cartesian', 'curvilinear'])
def test_rhs_vs_oracle_larger(n, mesh):
    """beyond the golden sizes: every tile variant against the oracle"""
    _say(f'[gpu] rhs-oracle {mesh} {n}')
    prob = G.problem('curv_sw' if mesh == 'curvilinear' else 'cart_sw')
    for b in prob.blocks:
        b.n = n
    od, vol, _ = H.oracle_rhs(prob, 0.3)
    for tile in range(_nvariants()):
        with _sim(prob, tile=tile, zchunk=7) as sim:
            H.load_pattern_state(sim, prob)
            sim.eval_rates(0.3, what=0)
            for ib in range(2):
                assert G.rel_l2(sim.rates(ib), vol[ib]) <= TOL, (tile, ib)
            H.load_pattern_state(sim, prob)
            sim.eval_rates(0.3, what=1, stage=2)
            for ib in range(2):
                assert G.rel_l2(sim.rates(ib), od.blocks[ib]['DF']) <= TOL, (tile, ib)


@pytest.mark.parametrize('mesh', ['cartesian', 'curvilinear'])
def test_rhs_chunk_pairs_vs_oracle(mesh):
    """tile order 3 (s-chunks in pairs, the lower chunk marched downwards):
    the fused kernel at chunk lengths giving odd and even chunk counts
    against the oracle"""
    from wql import native
    _say(f'[gpu] rhs-pairs {mesh}')
    prob = G.problem('curv_sw' if mesh == 'curvilinear' else 'cart_sw')
    for b in prob.blocks:
        b.n = (41, 33, 40)
    od, vol, _ = H.oracle_rhs(prob, 0.3)
    for tile in (1, 5, 0):
        for zc in (3, 4, 5, 7, 28):
            with _sim(prob, tile=tile, zchunk=zc) as sim:
                native.check(native.lib().wql_set_tile_order(sim.ctx, 3, 1))
                H.load_pattern_state(sim, prob)
                sim.eval_rates(0.3, what=0)
                for ib in range(2):
                    assert G.rel_l2(sim.rates(ib), vol[ib]) <= TOL, (tile, zc, ib)
                H.load_pattern_state(sim, prob)
                sim.eval_rates(0.3, what=1, stage=2)
                for ib in range(2):
                    assert G.rel_l2(sim.rates(ib), od.blocks[ib]['DF']) <= TOL, (tile, zc, ib)


@pytest.mark.parametrize('mesh', ['cartesian', 'curvilinear'])
def test_rhs_short_chunks_vs_oracle(mesh):
    """the one- and two-plane s chunks that the library picks for small
    grids, for the fused variants of both precisions' defaults, against the
    oracle"""
    _say(f'[gpu] rhs-short-chunks {mesh}')
    prob = G.problem('curv_sw' if mesh == 'curvilinear' else 'cart_sw')
    for b in prob.blocks:
        b.n = (31, 37, 29)
    od, vol, _ = H.oracle_rhs(prob, 0.3)
    for tile in (1, 0):
        for zc in (1, 2):
            with _sim(prob, tile=tile, zchunk=zc) as sim:
                H.load_pattern_state(sim, prob)
                sim.eval_rates(0.3, what=0)
                for ib in range(2):
                    assert G.rel_l2(sim.rates(ib), vol[ib]) <= TOL, (tile, zc, ib)
                H.load_pattern_state(sim, prob)
                sim.eval_rates(0.3, what=1, stage=2)
                for ib in range(2):
                    assert G.rel_l2(sim.rates(ib), od.blocks[ib]['DF']) <= TOL, (tile, zc, ib)


def test_fp32_constant_metrics_rupture_grid():
    """the rupture grid (x affine in q, z affine in s, topography in y): in
    FP32 the library folds q_x, s_x, q_y, s_y, q_z, s_z into constants (their
    FP64 spread is SBP round-off), so each is constant in the FP32 metrics and
    within 1e-10 of the metric scale of the FP64 metrics; the FP32 rates of
    the fused kernel (variant 0, which skips those loads) and of the
    lead-holding variant (variant 1, which reads every metric) stay within
    FP32 rounding of the oracle's rates"""
    from wql import workloads
    _say('[gpu] fp32 constant metrics')
    W = workloads.WORKLOADS['rupture'](scale=0.3)
    prob, coords = W['problem'], W['coords']
    od = O.OracleDomain(prob, coords=coords)
    H.load_pattern_state(H._NullSim(), prob, od)
    for ib in range(2):
        od.set_rates(ib)
    with _sim(prob, precision='fp64', coords=coords) as sim:
        G64 = [sim.metrics(ib) for ib in range(2)]
    for tile in (0, 1):
        with _sim(prob, precision='fp32', coords=coords, tile=tile) as sim:
            for ib in range(2):
                G32 = sim.metrics(ib)
                scale = np.abs(G64[ib][:9]).max()
                for c in (0, 2, 3, 5, 6, 8):
                    assert np.unique(G32[c]).size == 1, (ib, c)
                    assert np.abs(G32[c] - G64[ib][c]).max() <= 1e-10 * scale + 6e-8 * abs(G64[ib][c]).max(), (ib, c)
            H.load_pattern_state(sim, prob)
            sim.eval_rates(0.0, what=0)
            for ib in range(2):
                assert G.rel_l2(sim.rates(ib), od.blocks[ib]['DF']) <= 1e-5, (tile, ib)


def test_small_grid_default_chunks():
    """the library shortens the default s chunks on a small grid (2x31x61x61)
    and keeps them on the 2x201x101x101 rupture grid"""
    from wql import workloads
    for wl, fp32, fp64 in (('plain', 2, 4), ('rupture', 6, 12)):
        W = workloads.WORKLOADS[wl](scale=1.0)
        for prec, want in (('fp32', fp32), ('fp64', fp64)):
            with _sim(W['problem'], precision=prec, coords=W['coords'], tuned=False) as sim:
                v, zc, _, _ = sim.tile
                assert zc == want, (wl, prec, v, zc)


@pytest.mark.parametrize('case', ['curv_rs', 'pml_locked', 'plastic_sw'])
def test_graph_replays_match_eager(case):
    """a captured step graph (device-side clock) gives the eager launches'
    results bit for bit, over several replays with host work in between"""
    _say(f'[gpu] graph {case}')
    z = G.load(case)
    prob = G.problem(case)
    args = z['traj_args']
    F0 = H.pulse_init(prob, args)
    sims = [_sim(prob), _sim(prob, graph=True)]
    try:
        for sim in sims:
            for ib in range(2):
                sim.set_fields(ib, F0[ib])
            sim.init_state()
        for chunk in (1, 2, 1, 3):
            for sim in sims:
                sim.step(chunk)
                sim.sync()
            assert sims[0].t == sims[1].t
            for ib in range(2):
                assert np.array_equal(sims[0].fields(ib), sims[1].fields(ib)), f'F block {ib + 1}'
                fa, fb = sims[0].fault(ib), sims[1].fault(ib)
                for key in ('S', 'W', 'Svel', 'trup'):
                    assert np.array_equal(fa[key], fb[key]), key
    finally:
        for sim in sims:
            sim.close()


@pytest.mark.parametrize('case', ['curv_sw', 'curv_rs_long', 'pml_locked', 'plastic_sw', 'rw_plastic', 'tpv26_sw',
                                  'tpv26_plastic', 'mt_curv'])
def test_fp32_volume_tracks_fp64(case):
    """FP32 volume fields with the FP64 fault stay close to the FP64 path:
    relative L2 of the fields and of the slip <= 1e-5 at the trajectory's end
    (measured: ~2e-7).  The 12-step creep of curv_rs is left out: its fields
    are ~1e-10 MPa against a 75 MPa prestress, so both paths differ there by
    the FP64 rounding of the friction solve (~1e-14 absolute), not by FP32."""
    _say(f'[gpu] fp32 {case}')
    z = G.load(case)
    prob = G.problem(case)
    args = z['traj_args']
    F0 = H.pulse_init(prob, args)
    out = {}
    for prec in ('fp64', 'fp32'):
        with _sim(prob, precision=prec) as sim:
            for ib in range(2):
                sim.set_fields(ib, F0[ib])
            sim.init_state()
            sim.step(int(args[0]))
            sim.sync()
            out[prec] = [(sim.fields(ib).astype(np.float64), sim.fault(ib)) for ib in range(2)]
    for ib in range(2):
        F64, f64 = out['fp64'][ib]
        F32, f32 = out['fp32'][ib]
        assert G.rel_l2(F32, F64) <= 1e-5, f'F block {ib + 1}'
        slip64 = f64['S'][3]
        if np.abs(slip64).max() > 1e-3:
            assert G.rel_l2(f32['S'][3], slip64) <= 1e-5, 'slip'


@pytest.mark.parametrize('mesh', ['cartesian', 'curvilinear'])
def test_heterogeneous_rhs_vs_oracle(mesh):
    """node-wise lambda, mu, rho (the volume kernels' heterogeneous paths):
    every launch variant against the oracle"""
    _say(f'[gpu] rhs-het {mesh}')
    prob = G.problem('curv_sw' if mesh == 'curvilinear' else 'cart_sw')
    rng = np.random.default_rng(7)
    for ib, b in enumerate(prob.blocks):
        b.n = (41, 33, 29)
        nq, nr, ns = b.n
        base = b.lame()
        k = np.stack(np.meshgrid(np.arange(ns), np.arange(nr), np.arange(nq), indexing='ij'))
        for c in range(3):
            ph = rng.uniform(0, 2 * np.pi, 3)
            wave = np.sin(0.3 * k[0] + ph[0]) * np.cos(0.2 * k[1] + ph[1]) * np.sin(0.25 * k[2] + ph[2])
            base[c] = base[c] * (1.0 + 0.05 * wave)
        b.material = base
    od, vol, _ = H.oracle_rhs(prob, 0.3)
    for tile in range(_nvariants()):
        with _sim(prob, tile=tile, zchunk=6) as sim:
            H.load_pattern_state(sim, prob)
            sim.eval_rates(0.3, what=0)
            for ib in range(2):
                assert G.rel_l2(sim.rates(ib), vol[ib]) <= TOL, (tile, ib)
            H.load_pattern_state(sim, prob)
            sim.eval_rates(0.3, what=1, stage=2)
            for ib in range(2):
                assert G.rel_l2(sim.rates(ib), od.blocks[ib]['DF']) <= TOL, (tile, ib)


@pytest.mark.parametrize('case', ['curv_sw', 'plastic_sw'])
def test_fault_diagnostics(case):
    """peak slip rate (running max of the stage-1 slip rate) and the device
    fault-plane reductions against the same quantities formed on the host"""
    _say(f'[gpu] fault diagnostics {case}')
    z = G.load(case)
    prob = G.problem(case)
    args = z['traj_args']
    F0 = H.pulse_init(prob, args)
    with _sim(prob) as sim:
        for ib in range(2):
            sim.set_fields(ib, F0[ib])
        sim.init_state()
        peak = np.zeros((2,) + sim.fault(0)['trup'].shape)
        for n in range(int(args[0])):
            sim.step(1)
            sim.sync()
            for ib in range(2):
                sv = sim.fault(ib)['Svel']
                peak[ib] = np.maximum(peak[ib], np.sqrt(sv[1] ** 2 + sv[2] ** 2))
        for ib in range(2):
            assert np.array_equal(sim.fault_peak(ib), peak[ib])
            f = sim.fault(ib)
            st = sim.fault_stats(ib)
            rup = f['trup'][0] < 1e9
            assert st['max_slip'] == pytest.approx(max(0.0, f['S'][3].max()), rel=1e-15)
            sr = np.sqrt(f['Svel'][1] ** 2 + f['Svel'][2] ** 2)
            assert st['max_slip_rate'] == pytest.approx(sr.max(), rel=1e-15)
            assert st['max_peak_slip_rate'] == pytest.approx(peak[ib].max(), rel=1e-15)
            assert st['ruptured_nodes'] == rup.sum()
            if rup.any():
                assert st['mean_slip_ruptured'] == pytest.approx(f['S'][3][rup].mean(), rel=1e-12)
                assert st['first_rupture_time'] == f['trup'][0][rup].min()


@pytest.mark.parametrize('tile', [None, 0])
@pytest.mark.parametrize('s0_kind', [0, 1])
def test_plastic_flow_model_vs_oracle(s0_kind, tile):
    """the default plasticity model (plastic_flow, plastic.f90:221-290) with
    the constant (0) and the TPV26 depth (1) background stress: fields after
    six LSRK steps against the oracle, whose return map and stress profile
    are pinned by tests/test_oracle_plastic.py"""
    _say(f'[gpu] plastic model 2 s0_kind={s0_kind} tile={tile}')
    prob = G.problem('plastic_sw')
    prob.plastic = 2
    prob.pl_mu, prob.pl_beta, prob.pl_eta = 0.05, 0.3, 0.01
    prob.s0_kind = s0_kind
    args = G.load('plastic_sw')['traj_args']
    F0 = H.pulse_init(prob, args)
    od = O.OracleDomain(prob)
    for ib in range(2):
        od.blocks[ib]['F'][...] = F0[ib]
    od.step(6)
    with _sim(prob, tile=tile) as sim:
        for ib in range(2):
            sim.set_fields(ib, F0[ib])
        sim.step(6)
        sim.sync()
        for ib in range(2):
            assert G.rel_l2(sim.fields(ib), od.blocks[ib]['F']) <= TOL, ib
    # the background stress yields: the plastic rate is nonzero somewhere
    assert max(np.abs(od.blocks[ib]['P'][0]).max() for ib in range(2)) > 0.0


@pytest.mark.parametrize('tile', [None, 0])
def test_host_update_between_steps(tile):
    """fields set from the host between steps reach every kernel, including
    the q-face buffers that variant 24 keeps between stages: two steps, the
    fields halved from the host, two more steps, against the oracle doing the
    same"""
    _say(f'[gpu] host update between steps tile={tile}')
    prob = G.problem('cart_sw')
    args = G.load('cart_sw')['traj_args']
    F0 = H.pulse_init(prob, args)
    od = O.OracleDomain(prob)
    for ib in range(2):
        od.blocks[ib]['F'][...] = F0[ib]
    with _sim(prob, tile=tile) as sim:
        for ib in range(2):
            sim.set_fields(ib, F0[ib])
        sim.step(2)
        od.step(2)
        for ib in range(2):
            F1 = 0.5 * od.blocks[ib]['F']
            od.blocks[ib]['F'][...] = F1
            sim.set_fields(ib, F1)
        sim.step(2)
        sim.sync()
        od.step(2)
        for ib in range(2):
            assert G.rel_l2(sim.fields(ib), od.blocks[ib]['F']) <= TOL, ib


@pytest.mark.parametrize('shapes', [((37, 21, 29), (13, 21, 29)), ((65, 22, 31), (38, 22, 31)),
                                    ((13, 13, 13), (33, 13, 13))])
def test_rhs_block_shapes_vs_oracle(shapes):
    """blocks of different lengths along q, and q/r lengths whose last tile
    sits at its natural position (q 38, r 22: the tile starts exactly six
    nodes from the face) or is right-aligned (remainders 1 and 5), plus
    single-tile blocks: the fused (default and lead-holding) and role-split
    kernels against the oracle"""
    _say(f'[gpu] rhs shapes {shapes}')
    prob = G.problem('curv_sw')
    for b, n in zip(prob.blocks, shapes):
        b.n = n
    od, vol, _ = H.oracle_rhs(prob, 0.3)
    for tile in (None, 0, 6):
        kw = {} if tile is None else dict(tile=tile, zchunk=5)
        with _sim(prob, **kw) as sim:
            H.load_pattern_state(sim, prob)
            sim.eval_rates(0.3, what=0)
            for ib in range(2):
                assert G.rel_l2(sim.rates(ib), vol[ib]) <= TOL, (tile, ib)
            H.load_pattern_state(sim, prob)
            sim.eval_rates(0.3, what=1, stage=2)
            for ib in range(2):
                assert G.rel_l2(sim.rates(ib), od.blocks[ib]['DF']) <= TOL, (tile, ib)


@pytest.mark.parametrize('prec', ['fp64', 'fp32'])
def test_mixed_pair_vs_oracle(prec):
    """a Cartesian block beside a curvilinear one, with node-wise material
    on the curvilinear block only: the pair runs the curvilinear,
    heterogeneous kernels, so the library keeps uniform metrics in G and
    constant lambda, mu, rho in M for the other block.  Rates (every launch
    variant in FP64) and a 6-step trajectory against the oracle: relative L2
    <= 1e-12 in FP64, <= 1e-5 in FP32."""
    _say(f'[gpu] mixed pair {prec}')
    prob = G.problem('curv_sw')
    prob.blocks[0].mesh = 'cartesian'
    rng = np.random.default_rng(11)
    M = prob.blocks[1].lame()
    M *= 1.0 + 0.05 * rng.standard_normal(M.shape[1:])[None]
    prob.blocks[1].material = M
    tol = TOL if prec == 'fp64' else 1e-5
    od, vol, _ = H.oracle_rhs(prob, 0.3)
    for tile in (range(7) if prec == 'fp64' else (None,)):
        with _sim(prob, precision=prec, tile=tile) as sim:
            H.load_pattern_state(sim, prob)
            sim.eval_rates(0.3, what=0)
            for ib in range(2):
                assert G.rel_l2(sim.rates(ib), vol[ib]) <= tol, (tile, ib)
            H.load_pattern_state(sim, prob)
            sim.eval_rates(0.3, what=1, stage=2)
            for ib in range(2):
                assert G.rel_l2(sim.rates(ib), od.blocks[ib]['DF']) <= tol, (tile, ib)
    od = O.OracleDomain(prob)
    F0 = H.pulse_init(prob, [0, -1.0, 7.5, 0.0, 0.4, 1.0])
    with _sim(prob, precision=prec) as sim:
        for ib in range(2):
            sim.set_fields(ib, F0[ib])
            od.blocks[ib]['F'][...] = F0[ib]
        sim.step(6)
        sim.sync()
        od.step(6)
        for ib in range(2):
            assert G.rel_l2(sim.fields(ib), od.blocks[ib]['F']) <= tol, ib
            f = sim.fault(ib)
            ref = od.blocks[ib]['iface']['S']
            assert np.abs(f['S'] - ref).max() <= tol * max(1.0, np.abs(ref).max()), ib


@pytest.mark.parametrize('case', ['curv_rs', 'curv_rs_long', 'rw_plastic'])
def test_newton_mode_vs_oracle(case):
    """the fast slip-rate solve (rs_solver 1, bracketed Newton) against the
    oracle's same solve at the FP64 bar, and against the reference build's
    regula-falsi trajectory within 1e-6 during rupture (the reference's own
    stopping rule leaves it ~3e-7 from the root)"""
    _say(f'[gpu] newton {case}')
    z = G.load(case)
    prob = G.problem(case)
    prob.fault.rs_solver = 1
    args = z['traj_args']
    od = O.OracleDomain(prob)
    F0 = H.pulse_init(prob, args)
    for ib in range(2):
        od.blocks[ib]['F'][...] = F0[ib]
    od.init_state()
    od.step(int(args[0]))
    with _sim(prob) as sim:
        for ib in range(2):
            sim.set_fields(ib, F0[ib])
        sim.init_state()
        sim.step(int(args[0]))
        sim.sync()
        for ib in range(2):
            F = sim.fields(ib)
            assert G.rel_l2(F, od.blocks[ib]['F']) <= TOL, f'F block {ib + 1}'
            f = sim.fault(ib)
            for key in ('S', 'W', 'Svel'):
                ref = od.blocks[ib]['iface'][key]
                assert np.abs(f[key] - ref).max() <= TOL * max(1.0, np.abs(ref).max()), key
            if case != 'curv_rs':
                assert G.rel_l2(F, z[f'b{ib + 1}_F']) <= 1e-6


def test_stage_by_stage_equals_step():
    """wql_stage(1..5) reproduces wql_step bit for bit; stages out of order
    and a step started while one is part done are refused"""
    from wql import native
    _say('[gpu] stage by stage')
    z = G.load('curv_rs_long')
    prob = G.problem('curv_rs_long')
    F0 = H.pulse_init(prob, z['traj_args'])
    L = native.lib()
    with _sim(prob) as a, _sim(prob) as b:
        for sim in (a, b):
            for ib in range(2):
                sim.set_fields(ib, F0[ib])
            sim.init_state()
        a.step(3)
        a.sync()
        assert L.wql_stage(b.ctx, 2) == native.WQL_E_STATE
        for _ in range(3):
            for s in range(1, 6):
                b.stage(s)
                if s == 2:
                    assert L.wql_step(b.ctx, 1) == native.WQL_E_STATE
        b.sync()
        assert a.t == b.t
        for ib in range(2):
            assert np.array_equal(a.fields(ib), b.fields(ib)), ib
            fa, fb = a.fault(ib), b.fault(ib)
            for key in ('S', 'W', 'Svel', 'trup'):
                assert np.array_equal(fa[key], fb[key]), key
