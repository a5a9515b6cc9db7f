"""ctypes binding of the C oracle (liboracle.so).  TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, as the checker of the HIP path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
dp = C.POINTER(C.c_double)


class OraFault(C.Structure):
    _fields_ = [('law', C.c_int), ('stress_mode', C.c_int), ('T0', C.c_double * 3), ('s0', C.c_double * 6),
                ('npatch', C.c_int), ('patch', (C.c_double * 5) * 4), ('nuc', C.c_int), ('nucp', C.c_double * 5),
                ('fs', C.c_double), ('fd', C.c_double), ('Dc', C.c_double),
                ('bar_wy', C.c_double), ('bar_wz', C.c_double), ('bar_big', C.c_double),
                ('rs_a', C.c_double), ('rs_da', C.c_double), ('rs_w', C.c_double), ('rs_wr', C.c_double),
                ('rs_b', C.c_double), ('rs_L', C.c_double), ('rs_V0', C.c_double), ('rs_f0', C.c_double),
                ('rs_Vin', C.c_double),
                ('rs_law', C.c_int), ('rs_fw', C.c_double), ('rs_Vw', C.c_double), ('rs_dVw', C.c_double),
                ('rs_init', C.c_int), ('rs_W0', C.c_double), ('sw_law', C.c_int), ('coh', C.c_double * 3),
                ('frc', C.c_double * 5), ('rs_solver', C.c_int)]


class OraSource(C.Structure):
    _fields_ = [('type', C.c_int), ('duration', C.c_double), ('t_init', C.c_double), ('m', C.c_double * 6),
                ('x', C.c_double * 3), ('node', C.c_int * 3), ('node_x', C.c_double * 3)]


class OraBlock(C.Structure):
    _fields_ = [('nq', C.c_int), ('nr', C.c_int), ('ns', C.c_int), ('cartesian', C.c_int),
                ('bc', C.c_int * 6), ('pml', C.c_int * 6), ('npml', C.c_int),
                ('X', dp), ('metric', dp), ('M', dp), ('nml', dp * 6), ('F', dp), ('DF', dp),
                ('Q', dp * 6), ('DQ', dp * 6), ('P', dp),
                ('nsrc', C.c_int), ('src_order', C.c_int), ('src_alpha', C.c_int), ('src', C.POINTER(OraSource)),
                ('sbp', C.c_int), ('anelastic', C.c_int), ('an_tau', C.c_double * 4), ('an_w', C.c_double * 4),
                ('eta', dp), ('deta', dp), ('qinv', dp)]


class SbpHalf(C.Structure):
    """wql_sbp_half (waveqlab3d-gh_amd/csrc/wql_sbp_ops.h)"""
    _fields_ = [('nl', C.c_int), ('nr', C.c_int), ('lo', C.c_int), ('hi', C.c_int), ('band', C.c_double * 16),
                ('exact', C.c_double * 16), ('left', (C.c_double * 16) * 8), ('right', (C.c_double * 16) * 8)]


class SbpOp(C.Structure):
    """wql_sbp_op: one upwind / upwind-DRP operator, tables per direction"""
    _fields_ = [('family', C.c_char_p), ('order', C.c_int), ('band', C.c_int), ('tau0', C.c_double),
                ('v', SbpHalf * 3), ('s', SbpHalf * 3)]


def sbp_ops():
    """every operator table the library and the oracle share"""
    L = lib()
    out, k = [], 0
    while True:
        p = L.ora_sbp_table(k)
        if not p:
            return out
        out.append(SbpOp.from_address(p))
        k += 1


class OraIface(C.Structure):
    _fields_ = [('S', dp), ('DS', dp), ('W', dp), ('DW', dp), ('trup', dp), ('Svel', dp),
                ('Uhat', dp), ('Vhat', dp)]


class OraDomain(C.Structure):
    _fields_ = [('b', OraBlock * 2), ('I', OraIface * 2), ('fault', OraFault), ('plastic', C.c_int),
                ('pl_c', C.c_double), ('pl_nu', C.c_double), ('pl_Tv', C.c_double),
                ('pl_mu', C.c_double), ('pl_beta', C.c_double), ('pl_eta', C.c_double),
                ('s0_kind', C.c_int), ('t', C.c_double), ('dt', C.c_double)]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, 'liboracle.so')
        if not os.path.exists(path):
            raise RuntimeError('oracle/liboracle.so not built (make -C oracle)')
        _lib = C.CDLL(path)
        _lib.ora_metrics.argtypes = [C.POINTER(OraBlock)]
        _lib.ora_set_rates.argtypes = [C.POINTER(OraBlock)]
        _lib.ora_boundary.argtypes = [C.POINTER(OraBlock)]
        _lib.ora_interface.argtypes = [C.POINTER(OraDomain), C.c_int, C.c_double, C.c_int]
        _lib.ora_step.argtypes = [C.POINTER(OraDomain), C.c_int]
        _lib.ora_init_state.argtypes = [C.POINTER(OraDomain)]
        _lib.ora_sources.argtypes = [C.POINTER(OraBlock), C.c_double]
        _lib.ora_locate_sources.argtypes = [C.POINTER(OraBlock), C.c_int, dp, C.POINTER(OraSource)]
        _lib.ora_d1.argtypes = [dp, C.c_int, C.c_int, C.c_double]
        _lib.ora_d1.restype = C.c_double
        _lib.ora_sbp_op.argtypes = [C.c_int, C.c_int]
        _lib.ora_anelastic_init.argtypes = [C.POINTER(OraBlock), C.c_double, C.c_double, C.c_double]
        _lib.ora_sbp_table.argtypes = [C.c_int]
        _lib.ora_sbp_table.restype = C.c_void_p
        _lib.ora_rk_coeffs.argtypes = [dp, dp, dp]
        _lib.ora_plastic.argtypes = [C.POINTER(OraDomain)]
        _lib.ora_initial_stress.argtypes = [C.POINTER(OraDomain), dp, dp]
        _lib.ora_slab_points.argtypes = [C.POINTER(OraBlock), C.c_int]
        _lib.ora_slab_points.restype = C.c_size_t
        _lib.ora_hat.argtypes = [C.POINTER(OraFault)] + [dp] * 6 + [C.c_double] * 3 + [dp, dp, dp]
    return _lib


def _p(a):
    return a.ctypes.data_as(dp)


def fault_struct(f):
    if getattr(f, 'key', 0):
        raise NotImplementedError('the oracle does not restate the keyed fault problems (their parity is '
                                  'against the reference build)')
    s = OraFault()
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
    return s


class OracleDomain:
    """Owns numpy arrays for both blocks and both interface copies."""

    def __init__(self, prob, coords=None):
        L = lib()
        self.prob = prob
        self.d = OraDomain()
        self.keep = []
        self.blocks = []
        for ib, bs in enumerate(prob.blocks):
            nq, nr, ns = bs.n
            NP = nq * nr * ns
            X = bs.coords() if coords is None else coords[ib]
            arr = dict(X=np.ascontiguousarray(X, dtype=np.float64), metric=np.zeros((10, ns, nr, nq)),
                       M=bs.lame(), F=np.zeros((9, ns, nr, nq)), DF=np.zeros((9, ns, nr, nq)),
                       P=np.zeros((2, ns, nr, nq)))
            fp = [nr * ns, nr * ns, nq * ns, nq * ns, nq * nr, nq * nr]
            arr['nml'] = [np.zeros((9, fp[s])) for s in range(6)]
            b = self.d.b[ib]
            b.nq, b.nr, b.ns = nq, nr, ns
            b.cartesian = 1 if bs.mesh == 'cartesian' else 0
            b.bc[:] = bs.bc
            b.pml[:] = [int(v) for v in bs.pml]
            b.npml = bs.npml
            b.sbp = L.ora_sbp_op({'traditional': 0, 'upwind': 1, 'upwind_drp': 2}[prob.fd_type], prob.order)
            if b.sbp < -1:
                raise ValueError(f'no {prob.fd_type} operator of order {prob.order}')
            for k in ('X', 'metric', 'M', 'F', 'DF', 'P'):
                setattr(b, k, _p(arr[k]))
            for s in range(6):
                b.nml[s] = _p(arr['nml'][s])
            arr['Q'], arr['DQ'] = [None] * 6, [None] * 6
            for s in range(6):
                if bs.pml[s]:
                    sp = L.ora_slab_points(C.byref(b), s)
                    arr['Q'][s] = np.zeros((9, sp))
                    arr['DQ'][s] = np.zeros((9, sp))
                    b.Q[s] = _p(arr['Q'][s])
                    b.DQ[s] = _p(arr['DQ'][s])
            if L.ora_metrics(C.byref(b)) != 0:
                raise RuntimeError('non-unit normal in face frame')
            if getattr(prob, 'anelastic', None):
                # memory variables and Q^-1; the unrelaxed moduli replace M
                for k, nf in (('eta', 24), ('deta', 24), ('qinv', 2)):
                    arr[k] = np.zeros((nf, ns, nr, nq))
                    setattr(b, k, _p(arr[k]))
                if L.ora_anelastic_init(C.byref(b), *[float(v) for v in prob.anelastic]) != 0:
                    raise ValueError(f'anelastic weight_exp {prob.anelastic[1]} has no weight set')
            self.blocks.append(arr)
            FPn = nr * ns
            ia = dict(S=np.zeros((4, ns, nr)), DS=np.zeros((4, ns, nr)), W=np.zeros((1, ns, nr)),
                      DW=np.zeros((1, ns, nr)), trup=np.full((1, ns, nr), 1.0e9), Svel=np.zeros((3, ns, nr)),
                      Uhat=np.zeros((6, ns, nr)), Vhat=np.zeros((6, ns, nr)))
            I = self.d.I[ib]
            for k, v in ia.items():
                setattr(I, k, _p(v))
            arr['iface'] = ia
        self.d.fault = fault_struct(prob.fault)
        if prob.plastic and prob.s0_kind > 1:
            raise NotImplementedError('the oracle restates the constant and TPV26/27 plastic background stresses')
        self.d.plastic = prob.plastic
        for k in ('pl_c', 'pl_nu', 'pl_Tv', 'pl_mu', 'pl_beta', 'pl_eta', 's0_kind'):
            setattr(self.d, k, getattr(prob, k))
        self.d.t = 0.0
        self.d.dt = prob.dt()
        if getattr(prob, 'mollify_source', False):
            raise NotImplementedError('the oracle restates the point-source form only (the mollified sources '
                                      'are pinned against reference-build goldens)')
        if getattr(prob, 'moment', None):
            from wql.problem import block_sources
            for ib in range(2):
                src, held, order, alpha = block_sources(prob, ib)
                self.set_sources(ib, [dict(s_, unlocated=not h) for s_, h in zip(src, held)], order, alpha)

    def init_state(self):
        lib().ora_init_state(C.byref(self.d))

    def set_sources(self, ib, sources, order=2, alpha=0):
        """moment-tensor sources of block ib: dicts with type, duration,
        t_init, m (6), x (3); located on the block's grid by the oracle"""
        n = len(sources)
        arr = (OraSource * max(n, 1))()
        xyz = np.ascontiguousarray([s_['x'] for s_ in sources] or [[0.0, 0.0, 0.0]], dtype=np.float64)
        lib().ora_locate_sources(C.byref(self.d.b[ib]), n, _p(xyz), arr)
        for k, s_ in enumerate(sources):
            arr[k].type = s_['type']
            arr[k].duration = s_['duration']
            arr[k].t_init = s_['t_init']
            arr[k].m[:] = list(s_['m'])
            if s_.get('unlocated'):
                arr[k].node[:] = [-2, -2, -2]
                arr[k].node_x[:] = [-1.0, -1.0, -1.0]
        self.keep.append(arr)
        b = self.d.b[ib]
        b.nsrc, b.src_order, b.src_alpha = n, order, alpha
        b.src = C.cast(arr, C.POINTER(OraSource))
        return [(tuple(arr[k].node), tuple(arr[k].node_x)) for k in range(n)]

    def sources(self, ib, t):
        lib().ora_sources(C.byref(self.d.b[ib]), t)

    def set_rates(self, ib):
        lib().ora_set_rates(C.byref(self.d.b[ib]))

    def boundary(self, ib):
        lib().ora_boundary(C.byref(self.d.b[ib]))

    def interface(self, ib1, t, stage):
        rc = lib().ora_interface(C.byref(self.d), ib1, t, stage)
        if rc:
            raise RuntimeError(f'oracle interface error {rc}')

    def plastic(self):
        lib().ora_plastic(C.byref(self.d))

    def initial_stress(self, x):
        out = np.zeros(6)
        lib().ora_initial_stress(C.byref(self.d), np.ascontiguousarray(x, dtype=np.float64).ctypes.data_as(
            C.POINTER(C.c_double)), out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def step(self, n):
        rc = lib().ora_step(C.byref(self.d), n)
        if rc:
            raise RuntimeError(f'oracle step error {rc}')

    @property
    def t(self):
        return self.d.t
