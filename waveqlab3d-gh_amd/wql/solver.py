"""Python host mirror of the reference's domain/time-stepping API over libwql.

``Simulation`` plays the role of the reference's `domain_type` plus
`time_step_RK` (domain.f90, time_step.f90): it builds the two blocks from a
``Problem``, keeps the whole state on the device and advances it.
"""
import ctypes as C

import numpy as np

from . import geometry, native
from .native import check, dp


def _p(a):
    return a.ctypes.data_as(dp)


def fault_struct(f):
    s = native.FaultParams()
    s.law, s.stress_mode = f.law, f.stress_mode
    s.T0[:] = f.T0
    s.s0[:] = f.s0
    s.npatch = len(f.patches)
    for k, pt in enumerate(f.patches):
        s.patch[k][:] = pt
    s.nuc = f.nuc
    s.nucp[:] = f.nucp
    for k in ('fs', 'fd', 'Dc', 'bar_wy', 'bar_wz', 'bar_big', 'rs_a', 'rs_da', 'rs_w', 'rs_wr',
              'rs_b', 'rs_L', 'rs_V0', 'rs_f0', 'rs_Vin', 'rs_law', 'rs_fw', 'rs_Vw', 'rs_dVw', 'rs_init',
              'rs_W0', 'sw_law', 'rs_solver'):
        setattr(s, k, getattr(f, k))
    s.coh[:] = f.coh
    s.frc[:] = f.frc
    s.key = f.key
    return s


def problem_desc(prob, precision='fp64', dt=None):
    d = native.ProblemDesc()
    d.precision = native.WQL_FP32 if precision == 'fp32' else native.WQL_FP64
    d.dt = prob.dt() if dt is None else dt
    d.t0 = 0.0
    d.plastic = prob.plastic
    for k in ('pl_c', 'pl_nu', 'pl_Tv', 'pl_mu', 'pl_beta', 'pl_eta', 's0_kind'):
        setattr(d, k, getattr(prob, k))
    d.fault = fault_struct(prob.fault)
    d.fd_type = native.FD_TYPES[prob.fd_type]
    d.fd_order = prob.order
    if prob.anelastic:
        d.anelastic = 1
        d.an_c, d.an_weight_exp, d.an_fref = (float(v) for v in prob.anelastic)
    return d


def block_descs(prob, coords=None):
    arr = (native.BlockDesc * 2)()
    keep = []
    for ib, b in enumerate(prob.blocks):
        X = np.ascontiguousarray(coords[ib] if coords is not None else b.coords(), dtype=np.float64)
        keep.append(X)
        d = arr[ib]
        d.nq, d.nr, d.ns = b.n
        d.bc[:] = b.bc
        d.pml[:] = [int(v) for v in b.pml]
        d.npml = b.npml
        d.cartesian = 1 if b.mesh == 'cartesian' else 0
        d.X = _p(X)
        Mb = b.material if b.material is not None else b.medium_material(X)
        if Mb is not None:
            M = np.ascontiguousarray(Mb, dtype=np.float64)
            keep.append(M)
            d.M = _p(M)
        d.rho_s_p[:] = b.rho_s_p
    return arr, keep


def estimate_bytes(prob, precision='fp64'):
    """Device bytes a Simulation of ``prob`` would hold (admission control)."""
    pd = problem_desc(prob, precision)
    arr = (native.BlockDesc * 2)()
    for ib, b in enumerate(prob.blocks):
        d = arr[ib]
        d.nq, d.nr, d.ns = b.n
        d.pml[:] = [int(v) for v in b.pml]
        d.npml = b.npml
        d.cartesian = 1 if b.mesh == 'cartesian' else 0
    return native.lib().wql_estimate_bytes(C.byref(pd), arr)


class Simulation:
    def __init__(self, prob, precision='fp64', tile=None, zchunk=None, coords=None, graph=False, dt=None,
                 order=None, tuned=True):
        L = native.lib()
        self.prob = prob
        self.precision = precision
        self.ctx = C.c_void_p()
        pd = problem_desc(prob, precision, dt)
        arr, keep = block_descs(prob, coords)
        check(L.wql_create(C.byref(pd), arr, C.byref(self.ctx)))
        self.dt = pd.dt
        if order is not None:
            order = (order, 2)
        if tuned and tile is None and zchunk is None and order is None:
            from . import autotune
            best = autotune.lookup(prob, precision)
            if best is not None:
                tile, zchunk = best['variant'], best['zchunk']
                order = (best.get('order', 0), best.get('group', 2))
        if tile is not None or zchunk is not None:
            v0, z0 = C.c_int(), C.c_int()
            check(L.wql_get_tile(self.ctx, C.byref(v0), C.byref(z0)))
            check(L.wql_set_tile(self.ctx, v0.value if tile is None else tile, z0.value if zchunk is None else zchunk))
        if order is not None:
            check(L.wql_set_tile_order(self.ctx, *order))
        if prob.moment:
            for ib in range(2):
                X = coords[ib] if coords is not None else prob.blocks[ib].coords()
                self.set_sources(ib, X)
        if graph:
            check(L.wql_use_graph(self.ctx, 1))

    def set_sources(self, ib, X):
        """the problem's moment-tensor sources of block ib, located on the
        block's node coordinates X[3][ns][nr][nq] (init_moment_tensor)"""
        from .problem import block_sources
        src, held, order, alpha = block_sources(self.prob, ib)
        n = len(src)
        arr = (native.Source * max(n, 1))()
        for k, s_ in enumerate(src):
            arr[k].type = s_['type']
            arr[k].duration = s_['duration']
            arr[k].t_init = s_['t_init']
            arr[k].m[:] = list(s_['m'])
            arr[k].x[:] = list(s_['x'])
        X = np.ascontiguousarray(X, dtype=np.float64)
        if getattr(self.prob, 'mollify_source', False):
            # every source of the block's list, held or not (the Gaussian
            # reaches every node of the block)
            check(native.lib().wql_set_sources_smooth(self.ctx, ib, n, arr, _p(X)))
            return []
        nq, nr, ns = self.prob.blocks[ib].n
        xyz = np.ascontiguousarray([s_['x'] for s_ in src] or [[0.0, 0.0, 0.0]], dtype=np.float64)
        check(native.lib().wql_locate_sources(_p(X), nq, nr, ns, n, _p(xyz), arr))
        for k in range(n):
            if not held[k]:
                arr[k].node[:] = [-2, -2, -2]
                arr[k].node_x[:] = [-1.0, -1.0, -1.0]
        check(native.lib().wql_set_sources(self.ctx, ib, n, arr, order, alpha))
        return [tuple(arr[k].node) for k in range(n)]

    @property
    def tile(self):
        """(variant, zchunk, order, group) of the main volume kernel's launch shape"""
        v, z, o, g = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(native.lib().wql_get_tile(self.ctx, C.byref(v), C.byref(z)))
        check(native.lib().wql_get_tile_order(self.ctx, C.byref(o), C.byref(g)))
        return v.value, z.value, o.value, g.value

    def close(self):
        if self.ctx:
            native.lib().wql_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state -----------------------------------------------------------
    def _shape(self, ib, nf):
        nq, nr, ns = self.prob.blocks[ib].n
        return (nf, ns, nr, nq)

    def set_fields(self, ib, F):
        F = np.ascontiguousarray(F, dtype=np.float64)
        assert F.shape == self._shape(ib, 9)
        check(native.lib().wql_set_fields(self.ctx, ib, _p(F)))

    def fields(self, ib):
        F = np.empty(self._shape(ib, 9))
        check(native.lib().wql_get_fields(self.ctx, ib, _p(F)))
        return F

    def rates(self, ib):
        F = np.empty(self._shape(ib, 9))
        check(native.lib().wql_get_rates(self.ctx, ib, _p(F)))
        return F

    def metrics(self, ib):
        G = np.empty(self._shape(ib, 10))
        check(native.lib().wql_get_metrics(self.ctx, ib, _p(G)))
        return G

    def frame(self, ib, side):
        nq, nr, ns = self.prob.blocks[ib].n
        fp = [nr * ns, nr * ns, nq * ns, nq * ns, nq * nr, nq * nr][side]
        a = np.empty((9, fp))
        check(native.lib().wql_get_frame(self.ctx, ib, side, _p(a)))
        return a

    def slab_shape(self, ib, side):
        b = self.prob.blocks[ib]
        n = list(b.n)
        n[side // 2] = b.npml
        return (9, n[0] * n[1] * n[2])

    def set_pml(self, ib, side, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        check(native.lib().wql_set_pml(self.ctx, ib, side, _p(Q)))

    def pml(self, ib, side):
        Q = np.empty(self.slab_shape(ib, side))
        check(native.lib().wql_get_pml(self.ctx, ib, side, _p(Q)))
        return Q

    def pml_rates(self, ib, side):
        Q = np.empty(self.slab_shape(ib, side))
        check(native.lib().wql_get_pml_rates(self.ctx, ib, side, _p(Q)))
        return Q

    def set_fault(self, ib, S=None, W=None):
        L = native.lib()
        Sp = _p(np.ascontiguousarray(S, dtype=np.float64)) if S is not None else None
        Wp = _p(np.ascontiguousarray(W, dtype=np.float64)) if W is not None else None
        keep = (S, W)
        check(L.wql_set_fault(self.ctx, ib, Sp, Wp))

    def fault(self, ib):
        nq, nr, ns = self.prob.blocks[ib].n
        out = dict(S=np.empty((4, ns, nr)), W=np.empty((1, ns, nr)), trup=np.empty((1, ns, nr)),
                   Svel=np.empty((3, ns, nr)), Uhat=np.empty((6, ns, nr)), Vhat=np.empty((6, ns, nr)))
        check(native.lib().wql_get_fault(self.ctx, ib, *(_p(out[k]) for k in
                                                         ('S', 'W', 'trup', 'Svel', 'Uhat', 'Vhat'))))
        return out

    def fault_faces(self):
        """fields on the fault: block 0's +q face and block 1's -q face,
        each [9][ns][nr]"""
        nr, ns = self.prob.blocks[0].n[1:]
        U, V = np.empty((9, ns, nr)), np.empty((9, ns, nr))
        check(native.lib().wql_get_fault_faces(self.ctx, _p(U), _p(V)))
        return U, V

    def fault_peak(self, ib):
        """peak tangential slip rate per fault node [1][ns][nr]"""
        nq, nr, ns = self.prob.blocks[ib].n
        out = np.empty((1, ns, nr))
        check(native.lib().wql_get_fault_peak(self.ctx, ib, _p(out)))
        return out

    def fault_stats(self, ib=0):
        """device-side fault-plane statistics of one block's copy"""
        out = np.empty(6)
        check(native.lib().wql_fault_stats(self.ctx, ib, _p(out)))
        keys = ('max_slip', 'max_slip_rate', 'max_peak_slip_rate', 'ruptured_nodes', 'mean_slip_ruptured',
                'first_rupture_time')
        return dict(zip(keys, out.tolist()))

    def plastic(self, ib):
        """plastic state [2][ns][nr][nq]: strain rate lambda of the last
        update, accumulated plastic strain gammap (the reference's P%P)"""
        P = np.empty(self._shape(ib, 2))
        check(native.lib().wql_get_plastic(self.ctx, ib, _p(P)))
        return P

    def set_anelastic(self, ib, eta):
        """anelastic memory variables [24][ns][nr][nq] (component sxx..syz
        times mechanism 1..4)"""
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        assert eta.shape == self._shape(ib, 24)
        check(native.lib().wql_set_anelastic(self.ctx, ib, _p(eta)))

    def anelastic(self, ib, what=0):
        """memory variables (what 0) or their rates (what 1), [24][ns][nr][nq]"""
        a = np.empty(self._shape(ib, 24))
        check(native.lib().wql_get_anelastic(self.ctx, ib, what, _p(a)))
        return a

    def fault_rates(self, ib):
        nq, nr, ns = self.prob.blocks[ib].n
        DS, DW = np.empty((4, ns, nr)), np.empty((1, ns, nr))
        check(native.lib().wql_get_fault_rates(self.ctx, ib, _p(DS), _p(DW)))
        return DS, DW

    # -- stepping ----------------------------------------------------------
    def init_state(self):
        check(native.lib().wql_init_state(self.ctx))

    def eval_rates(self, t, what=1, stage=2):
        check(native.lib().wql_eval_rates(self.ctx, t, what, stage))

    def step(self, n=1):
        check(native.lib().wql_step(self.ctx, n))

    def stage(self, s):
        """RK stage s (1..5) of the current step"""
        check(native.lib().wql_stage(self.ctx, s))

    def sync(self):
        check(native.lib().wql_sync(self.ctx))

    @property
    def t(self):
        return native.lib().wql_time(self.ctx)

    def device_bytes(self):
        return native.lib().wql_device_bytes(self.ctx)

    def profile(self, on=True):
        check(native.lib().wql_profile(self.ctx, 1 if on else 0))

    def volume_kernel_ms(self):
        return native.lib().wql_volume_kernel_ms(self.ctx)


# wql_grid_desc.profile of the reference's profile_type names (others: planar)
PROFILE_CODES = {'analytical_test_problem': 1, 'analytical_tpv28': 2, 'analytical_tpv10': 3, 'analytical_tpv36': 4}


def device_grid(block):
    """node coordinates X[3][ns][nr][nq] of a Block generated on the device
    by wql_grid_block (the same grid as block.coords())"""
    nq, nr, ns = block.n
    if block.mesh != 'cartesian' and block.topography:
        # the topography branch: its faces from the surface data on the host,
        # the volume interpolation on the device
        f = geometry.topography_faces(block.n, block.a, block.b, lc=block.lc, rc=block.rc, profile=block.profile,
                                      profile_path=block.profile_path, faultsize=block.faultsize,
                                      topography=block.topography, topography_path=block.topography_path,
                                      topo=block.topo)
        faces = np.ascontiguousarray(np.concatenate([a.T.ravel() for c in 'XYZ' for a in f[c]]))
        X = np.empty((3, ns, nr, nq))
        check(native.lib().wql_grid_faces(nq, nr, ns, _p(faces), _p(X)))
        return X
    g = native.GridDesc()
    g.nq, g.nr, g.ns = block.n
    g.a[:] = block.a
    g.b[:] = block.b
    g.mesh = 0 if block.mesh == 'cartesian' else 1
    g.profile = PROFILE_CODES.get(block.profile, 0)
    g.lc, g.rc = block.lc, block.rc
    keep = []
    off, side = None, 0
    if block.fault_offset is not None:
        off, side = block.fault_offset, 1 if block.bc[1] == 0 else 0
    if block.mesh != 'cartesian' and block.profile == 'read_from_memomry_fractal':
        # the file's x column times the face weight, as an offset field
        if block.lc != 0.0 and block.rc != 0.0:
            raise ValueError('a fault profile file on both q faces of one block is not supported')
        if block.lc != 0.0 or block.rc != 0.0:
            X1 = geometry.read_fault_profile(block.profile_path, block.n[1], block.n[2])
            off, side = (block.rc * X1, 0) if block.rc != 0.0 else (block.lc * X1, 1)
    if off is not None:
        fx = np.ascontiguousarray(np.asarray(off, dtype=np.float64).T)  # [ns][nr]
        keep.append(fx)
        g.fault_x = _p(fx)
        g.fault_side = side
    if block.top_offset is not None:
        ty = np.ascontiguousarray(np.asarray(block.top_offset, dtype=np.float64).T)  # [ns][nq]
        keep.append(ty)
        g.top_y = _p(ty)
    nq, nr, ns = block.n
    X = np.empty((3, ns, nr, nq))
    check(native.lib().wql_grid_block(C.byref(g), _p(X)))
    return X
