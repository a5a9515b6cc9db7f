"""Problem description: the reference's input-file API restated as data.

A ``Problem`` holds what the reference's &problem_list / &block_list
namelists carry for the hot path (domain.f90:105-136, datatypes.f90:255-266)
plus the fault law parameters that the reference keys by problem name
(Interface_Condition.f90:1268-2435).  ``from_namelist`` reads a reference
input file; ``fault_params`` turns a problem key into explicit parameters.
"""
import math
import os
import re
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import geometry

LOCKED, SLIP_WEAKENING, RATE_STATE = 0, 1, 2


@dataclass
class Block:
    n: Tuple[int, int, int]
    a: Tuple[float, float, float]
    b: Tuple[float, float, float]
    rho_s_p: Tuple[float, float, float] = (2.67, 3.464, 6.0)   # rho, cs, cp
    bc: Tuple[int, ...] = (1, 0, 2, 1, 1, 1)                   # Lx Rx Ly Ry Lz Rz
    pml: Tuple[bool, ...] = (False,) * 6
    npml: int = 0
    lc: float = 0.0
    rc: float = 0.0
    mesh: str = 'cartesian'                                    # 'cartesian' | 'curvilinear'
    profile: str = 'analytical_test_problem'
    profile_path: Optional[str] = None                         # 'read_from_memomry_fractal' profile file
    # use_topography: the topography branch of Curve_Grid_3D3 with this
    # topography_type ('analytical_topography' or 'read_topo_from_memory'),
    # its file, the &problem_list topo scale and btp%faultsize
    topography: Optional[str] = None
    topography_path: Optional[str] = None
    topo: float = 1.0
    faultsize: Optional[Tuple[int, int]] = None
    fault_offset: Optional[np.ndarray] = None                  # [nr][ns] x-offsets of the fault face
    top_offset: Optional[np.ndarray] = None                    # [nq][ns] y-offsets of the r=1 face
    material: Optional[np.ndarray] = None                      # [3][ns][nr][nq] lambda, mu, rho (heterogeneous)
    medium: Optional[str] = None                               # init_material's keyed medium (problem name)

    def coords(self):
        if self.mesh == 'cartesian':
            return geometry.cartesian(self.n, self.a, self.b)
        if self.topography:
            return geometry.curvilinear_topography(
                self.n, self.a, self.b, lc=self.lc, rc=self.rc, profile=self.profile,
                profile_path=self.profile_path, faultsize=self.faultsize, topography=self.topography,
                topography_path=self.topography_path, topo=self.topo)
        fl = fr = None
        if self.fault_offset is not None:
            if self.bc[1] == 0:
                fr = self.fault_offset
            else:
                fl = self.fault_offset
        return geometry.curvilinear(self.n, self.a, self.b, self.lc, self.rc,
                                    self.profile, fault_left=fl, fault_right=fr, top=self.top_offset,
                                    profile_path=self.profile_path)

    def medium_material(self, X=None):
        """lambda, mu, rho of the keyed medium at the nodes X (default: this
        block's grid), or None without one"""
        if self.medium is None:
            return None
        return problem_material(self.medium, self, X)

    def lame(self):
        """M = (lambda, mu, rho) on every node (material.f90:181-188)."""
        if self.material is not None:
            return np.ascontiguousarray(self.material, dtype=np.float64)
        if self.medium is not None:
            return self.medium_material()
        rho, cs, cp = self.rho_s_p
        nq, nr, ns = self.n
        M = np.empty((3, ns, nr, nq))
        M[0] = rho * (cp ** 2 - 2.0 * cs ** 2)
        M[1] = rho * cs ** 2
        M[2] = rho
        return M

    def spacing(self):
        return tuple((self.b[d] - self.a[d]) / float(self.n[d] - 1) for d in range(3))


@dataclass
class Fault:
    law: int = LOCKED
    stress_mode: int = 0
    T0: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    s0: Tuple[float, ...] = (0.0,) * 6
    patches: List[Tuple[float, float, float, float, float]] = field(default_factory=list)
    nuc: int = 0
    nucp: Tuple[float, ...] = (0.0,) * 5
    fs: float = 0.0
    fd: float = 0.0
    Dc: float = 1.0
    bar_wy: float = 1e30
    bar_wz: float = 1e30
    bar_big: float = 0.0
    rs_a: float = 0.0
    rs_da: float = 0.0
    rs_w: float = 0.0
    rs_wr: float = 1.0
    rs_b: float = 0.0
    rs_L: float = 1.0
    rs_V0: float = 1.0
    rs_f0: float = 0.0
    rs_Vin: float = 0.0
    rs_law: int = 0            # state evolution: 0 aging, 1 slip law + flash weakening, 2 KDN
    rs_fw: float = 0.0         # flash-weakening residual friction
    rs_Vw: float = 0.0         # flash-weakening velocity (<= 0: off) ...
    rs_dVw: float = 0.0        # ... + rs_dVw (1 - B(y) B(z)) outside the velocity-weakening box
    rs_init: int = 0           # initial state: 0 steady state from the prestress, 1 constant rs_W0
    rs_W0: float = 0.0
    sw_law: int = 0            # slip-weakening strength: 0 linear, 1 TPV26/27 cohesion + forced rupture
    coh: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    frc: Tuple[float, ...] = (0.0, 0.0, 1.0, 1.0, 1.0)
    rs_solver: int = 0         # 0 regula falsi (the reference's), 1 bracketed Newton
    # a slip-weakening problem the reference keys by name (native.FAULT_KEYS:
    # TPV10, TPV29/30, TPV31-34, SCITS2016, TPV36/37, wasatch100) whose
    # prestress, strength, cohesion and forced-rupture time the library
    # evaluates at every fault node; 0: the fields above describe the fault
    key: int = 0


def fault_params(problem: str, coupling: str) -> Fault:
    """Explicit parameters for the reference's problem keys."""
    from .native import FAULT_KEYS
    if coupling == 'locked':
        return Fault(law=LOCKED)
    if coupling == 'slip-weakening_friction' and problem in FAULT_KEYS:
        # prestress 1802-2435, Tau_strength 1268-1797, Frictional_Cohesion
        # 2491-2580, Time_of_Forced_Rupture 2583-2660, evaluated per node
        return Fault(law=SLIP_WEAKENING, key=FAULT_KEYS[problem])
    if coupling == 'slip-weakening_friction':
        if problem == 'TPV5':   # Interface_Condition.f90:1283-1299, 1822-1830
            return Fault(law=SLIP_WEAKENING, stress_mode=0, T0=(-120.0, 0.0, 70.0),
                         patches=[(7.5, -7.5, 1.5, 1.5, 8.0), (7.5, 0.0, 1.5, 1.5, 11.6),
                                  (7.5, 7.5, 1.5, 1.5, -8.0)],
                         fs=0.677, fd=0.525, Dc=0.4, bar_wy=15.0, bar_wz=15.0, bar_big=1.0e10)
        if problem == 'TPV28':  # 1459-1474, 2049-2079, initial_stress_condition.f90:78-84
            return Fault(law=SLIP_WEAKENING, stress_mode=1, s0=(-60.0, 0.0, 60.0, 0.0, 29.38, 0.0),
                         nuc=1, nucp=(7.5, 0.0, 1.4, 0.6, 11.6),
                         fs=0.677, fd=0.373, Dc=0.4, bar_wy=15.0, bar_wz=18.0, bar_big=1.0e10)
        if problem in ('TPV26', 'TPV27'):
            # 1376-1458 (strength with cohesion and forced rupture), 2008-2048
            # (depth-dependent prestress, initial_stress_condition.f90:26-76)
            return Fault(law=SLIP_WEAKENING, stress_mode=2, fs=0.18, fd=0.12, Dc=0.3,
                         bar_wy=20.0, bar_wz=20.0, bar_big=1.0e10, sw_law=1,
                         coh=(0.4, 0.72, 5.0), frc=(10.0, -5.0, 4.0, 3.464, 0.5))
    if coupling == 'rate-and-state_friction' and problem == 'rate-weakening':
        # 410-416 (a box), 438-444 (slip law with flash weakening),
        # 2387-2420 (prestress + pulse), 2820-2850 (constant state),
        # initial_stress_condition.f90:184-192 (tensor, with its pi constant)
        pi = 3.14159265359
        psi = 50.0 / 180.0 * pi
        s1, s5 = -126.0, 42.0
        s3 = (1.0 - 2.0 * s5 / (s1 * math.tan(2.0 * psi))) * s1
        return Fault(law=RATE_STATE, stress_mode=1, s0=(s1, 0.5 * (s1 + s3), s3, 0.0, s5, 0.0),
                     nuc=3, nucp=(8.0, 0.0, 3.0 * 1.5 * 0.3, -0.2429, 0.0),
                     rs_a=0.016, rs_da=0.016, rs_w=20.0, rs_wr=2.5, rs_b=0.02, rs_L=2.0 * 0.2572,
                     rs_V0=1.0e-6, rs_f0=0.7, rs_Vin=0.0, rs_law=1, rs_fw=0.13, rs_Vw=0.17, rs_dVw=5.0,
                     rs_init=1, rs_W0=0.4367)
    if coupling == 'rate-and-state_friction' and problem == 'KDN':
        # the reference keys friction (402-408) and the state law (446-449)
        # for 'KDN' but has no prestress or initial state for it: its run
        # stops in prestress (2425-2427).  The law is reachable here through
        # an explicit Fault(rs_law=2, ...).
        raise ValueError("problem 'KDN' has no prestress in the reference (Interface_Condition.f90:2425-2427); "
                         "give an explicit Fault(law=RATE_STATE, rs_law=2, ...) instead")
    if coupling == 'rate-and-state_friction' and problem in ('TPV101', 'TPV102'):
        # 386-400, 433-436, 2324-2385, 2756-2780, initial_stress_condition.f90:114-120
        return Fault(law=RATE_STATE, stress_mode=1, s0=(-120.0, 0.0, -120.0, 0.0, 75.0, 0.0),
                     nuc=2, nucp=(7.5, 0.0, 3.0, 25.0, 1.0),
                     rs_a=0.008, rs_da=0.008, rs_w=15.0, rs_wr=3.0, rs_b=0.012, rs_L=0.02,
                     rs_V0=1.0e-6, rs_f0=0.6, rs_Vin=1.0e-12)
    raise ValueError(f'unsupported problem/coupling: {problem}/{coupling}')


@dataclass
class Problem:
    blocks: List[Block]
    fault: Fault
    cfl: float = 0.5
    t_final: float = 0.0
    plastic: int = 0           # 0 elastic, 1 Drucker-Prager return map ('scec'), 2 default model
    pl_c: float = 1.36
    pl_nu: float = 0.1934
    pl_Tv: float = 0.03
    pl_mu: float = 0.0
    pl_beta: float = 0.0
    pl_eta: float = 0.0
    s0_kind: int = 0           # plastic background stress: 0 = fault.s0 everywhere, 1 = TPV26/27, 2 = TPV29/30 depth profile
    name: str = 'default'
    w_stride: int = 1          # fault output every w_stride steps (time_step.f90:137)
    w_fault: bool = False      # write the fault output files (fault_output.f90)
    # moment-tensor point sources (&moment_list, moment_tensor.f90:82-270):
    # None, or dict(order, preference, lists) with lists {'union': [...]} or
    # {'U': [...], 'V': [...]} of source dicts (type, duration, t_init, m, x,
    # alpha); see block_sources
    moment: Optional[dict] = None
    # mollify_source: the sources spread as Gaussians over every node
    # (set_moment_tensor_smooth) instead of the discrete delta
    mollify_source: bool = False
    # spatial operator (&problem_list fd_type, order; domain.f90:106, 135-136):
    # 'traditional' (sixth-order central), 'upwind' (order 2-9) or
    # 'upwind_drp' (order 3-7, 66, 679); the order is unused for traditional
    fd_type: str = 'traditional'
    order: int = 5
    # anelastic response (response = 'anelastic', &anelastic_list; material.f90:
    # 11-145): None, or (c, weight_exp, fref) with Q_s = c v_s, Q_p = Q_s / 2
    anelastic: Optional[tuple] = None

    def dt(self):
        """Global time step (domain.f90:222-235, block.f90:437-452)."""
        dts = []
        for b in self.blocks:
            h = min(b.spacing())
            dts.append(self.cfl * h / math.sqrt(b.rho_s_p[1] ** 2 + b.rho_s_p[2] ** 2))
        return min(dts)

    def nt(self):
        return int(math.floor(self.t_final / self.dt()))


SOURCE_TYPES = {'gaussian': 0, 'LOH_discontinuity': 1, 'Brune': 2, 'Cos-Bell': 3, 'cosine': 4,
                'LOH_moment_time_history': 5, 'sine_moment_time_history': 6}


def _tensor_lines(lines, begin, end):
    """source lines between two markers, blank and '!' lines skipped
    (moment_tensor.f90:24-30), or None without the begin marker"""
    if begin not in lines:
        return None
    out = []
    for ln in lines[lines.index(begin) + 1:]:
        if ln.strip() == end:
            return out
        t = ln.strip()
        if not t or t.startswith('!'):
            continue
        out.append(t)
    raise ValueError(f'moment tensor list: {end} missing')


def _source(line):
    """one source line: type duration t_init mXX mYY mZZ mXY mXZ mYZ x y z alpha
    (read list-directed, moment_tensor.f90:218-223)"""
    tok = re.findall(r"'[^']*'|\"[^\"]*\"|[^\s,]+", line)
    if len(tok) < 13:
        raise ValueError(f'moment tensor line needs 13 values: {line!r}')
    kind = tok[0].strip("'\"")
    if kind not in SOURCE_TYPES:
        raise ValueError(f'invalid source type {kind!r} (moment_tensor.f90:970)')
    v = [float(t.replace('d', 'e').replace('D', 'e')) for t in tok[1:12]]
    return dict(type=SOURCE_TYPES[kind], duration=v[0], t_init=v[1], m=tuple(v[2:8]), x=tuple(v[8:11]),
                alpha=int(float(tok[12].replace('d', 'e').replace('D', 'e'))))


def read_moment(text):
    """&moment_list and the source lists of an input file (block.f90:63-85,
    moment_tensor.f90:82-270), or None when use_moment_tensor is off"""
    ml = read_namelist(text, 'moment_list')
    if not ml.get('use_moment_tensor', False):
        return None
    lines = [ln.rstrip('\n').rstrip('\r') for ln in text.splitlines()]
    lines = [ln.strip() if ln.strip().startswith('!---') else ln for ln in lines]
    pref = int(ml.get('tensor_block_preference', 1))
    out = dict(order=int(ml.get('order', 2)), preference=2 if pref == 2 else 1)
    union = _tensor_lines(lines, '!---begin:tensor_list---', '!---end:tensor_list---')
    if union is not None:
        out['lists'] = {'union': [_source(ln) for ln in union]}
    else:
        out['lists'] = {}
        for tag in ('U', 'V'):
            body = _tensor_lines(lines, f'!---begin:tensor_list{tag}---', f'!---end:tensor_list{tag}---')
            out['lists'][tag] = [_source(ln) for ln in body] if body is not None else None
    return out


def block_sources(prob, ib):
    """(sources, held, order, alpha) of block ib (0 or 1): the block's list
    (the unified one, or 'U' for block 1 and 'V' for block 2), whether the
    block may hold each source (with the unified list only the preferred
    block does, moment_tensor.f90:346-358), the operator order, and the
    first source's alpha, which the reference uses for all of a block's
    sources (singular_source's alpha(1))"""
    m = prob.moment
    if not m:
        return [], [], 2, 0
    if 'union' in m['lists']:
        src = m['lists']['union']
        held = [ib + 1 == m['preference']] * len(src)
    else:
        src = m['lists'].get('U' if ib == 0 else 'V')
        if src is None:
            raise ValueError('error reading the moment tensor list of block %d' % (ib + 1))
        held = [True] * len(src)
    return src, held, m['order'], (src[0]['alpha'] if src else 0)


def _nml_value(tok):
    tok = tok.strip()
    if tok in ('T', '.true.', '.TRUE.'):
        return True
    if tok in ('F', '.false.', '.FALSE.'):
        return False
    if tok[:1] in "'\"":
        return tok.strip("'\"")
    t = tok.replace('d', 'e').replace('D', 'e')
    try:
        return int(t)
    except ValueError:
        return float(t)


def read_namelist(text, group):
    """Minimal reader for the reference's &group ... / namelists."""
    text = re.sub(r'!.*', '', text)
    m = re.search(r'&' + group + r'\b(.*?)/\s*$', text, re.S | re.M)
    if not m:
        return {}
    body = m.group(1)
    out = {}
    for key, val in re.findall(r"([A-Za-z_][\w%()]*)\s*=\s*((?:'[^']*'|[^=])*?)(?=,?\s*[A-Za-z_][\w%()]*\s*=|\s*$)", body, re.S):
        vals = [v for v in re.findall(r"'[^']*'|[^,\s]+", val) if v.strip(',')]
        parsed = [_nml_value(v) for v in vals]
        out[key.lower()] = parsed[0] if len(parsed) == 1 else parsed
    return out


MEDIA = ('LOH1', 'LOH1_Harmonic', 'TPV31', 'TPV32', 'TPV33', 'OKLAHOMA')


def problem_material(problem, block, X=None):
    """lambda, mu, rho [3][ns][nr][nq] of init_material's medium for this
    problem name at the block's nodes X (default: block.coords(); material.
    f90:179-690: 'LOH1', 'LOH1_Harmonic', 'TPV31', 'TPV32', 'TPV33',
    'OKLAHOMA'), or None for the homogeneous rho_s_p medium of every other
    problem"""
    from . import native
    nq, nr, ns = block.n
    X = np.ascontiguousarray(block.coords() if X is None else X, dtype=np.float64)
    M = np.empty((3, ns, nr, nq))
    dp = native.dp
    rc = native.lib().wql_problem_material(problem.encode(), nq, nr, ns, X.ctypes.data_as(dp), M.ctypes.data_as(dp))
    if rc < 0:
        native.check(rc)
    return M if rc == 1 else None


def _data_path(p, base):
    """a file named in an input file: as given when it exists (the reference
    opens it from the run directory), else relative to the input file's
    directory; None for the reference's placeholders"""
    if not p or p.strip().lower() in ('no-path', 'no_path', 'none'):
        return None
    p = p.strip()
    if os.path.isabs(p) or os.path.exists(p):
        return p
    return os.path.join(base, p)


def read_material_files(paths, n):
    """lambda, mu, rho [3][ns][nr][nq] from the three files of
    init_material_from_file (material.f90:706-747, read_file_distributed over
    the block's global array): raw native FP64, q fastest"""
    if len(paths) != 3 or any(p is None for p in paths):
        raise ValueError("material_source = 'file' needs three material_path files (lambda, mu, rho)")
    nq, nr, ns = n
    M = np.empty((3, ns, nr, nq))
    for k, p in enumerate(paths):
        try:
            a = np.fromfile(p, dtype=np.float64, count=nq * nr * ns)
        except OSError:
            raise ValueError(f'error opening material file {p}') from None
        if a.size != nq * nr * ns:
            raise ValueError(f'material file {p}: {a.size} values, the block has {nq * nr * ns}')
        M[k] = a.reshape(ns, nr, nq)
    return M


def block_list_file(text, base):
    """the file named by '&block_list = 'file'' in an input file, relative to
    the input file's directory unless absolute (locate_block_list_file,
    input_utils.f90:9-94), or None"""
    for line in text.splitlines():
        line = line.split('!', 1)[0].strip()
        m = re.match(r'&block_list\s*=\s*([\'"])(.*?)\1', line, re.I)
        if m:
            return m.group(2) if os.path.isabs(m.group(2)) else os.path.join(base, m.group(2))
        if re.match(r'&block_list\b', line, re.I):
            return None
    return None


def from_namelist(path):
    """Build a Problem from a reference input file (&problem_list, &block_list;
    the block list may sit in its own file, domain.f90:165-211)."""
    text = open(path).read()
    base = os.path.dirname(os.path.abspath(path))
    pl = read_namelist(text, 'problem_list')
    # inputs outside this library's path fail here rather than run as
    # something else
    if int(pl.get('nblocks', 2)) != 2:
        raise NotImplementedError('the path runs two blocks joined by one interface')
    if read_namelist(text, 'mms_list').get('use_mms', False):
        raise NotImplementedError('use_mms: the manufactured-solution forcing is not built')
    blf = block_list_file(text, base)
    if blf is not None:
        try:
            btext = open(blf).read()
        except OSError:
            raise ValueError(f'error opening block list file {blf}') from None
        bl = read_namelist(btext, 'block_list')
    else:
        bl = read_namelist(text, 'block_list')
    problem = pl.get('problem', 'TPV5')
    coupling = pl.get('coupling', 'locked')
    mesh = pl.get('type_of_mesh', 'cartesian')
    blocks = []
    for i in (1, 2):
        g = lambda k, d=None: bl.get(f'btp({i})%{k}', d)
        lq, rq = g('lqrs'), g('rqrs')
        pml_l = g('pml_lqrs', [False] * 3)
        pml_r = g('pml_rqrs', [False] * 3)
        blocks.append(Block(
            n=tuple(g('nqrs')), a=tuple(float(v) for v in g('aqrs')), b=tuple(float(v) for v in g('bqrs')),
            rho_s_p=tuple(float(v) for v in g('rho_s_p', (2.67, 3.464, 6.0))),
            bc=(lq[0], rq[0], lq[1], rq[1], lq[2], rq[2]),
            pml=(pml_l[0], pml_r[0], pml_l[1], pml_r[1], pml_l[2], pml_r[2]),
            npml=int(g('npml', 0) or 0), lc=float(g('lc', 0.0) or 0.0), rc=float(g('rc', 0.0) or 0.0),
            mesh=mesh, profile=g('profile_type', 'analytical_test_problem'),
            profile_path=_data_path(g('profile_path', None), base)))
        if pl.get('use_topography', False) and mesh == 'curvilinear':
            # the topography branch of Curve_Grid_3D3 (domain.f90:280-288:
            # faultsize sizes a profile file only with a topography file)
            b = blocks[-1]
            b.topography = g('topography_type', 'none')
            b.topography_path = _data_path(g('topography_path', None), base)
            b.topo = float(pl.get('topo', 1.0))
            fs = g('faultsize', None)
            if fs and b.profile == 'read_from_memomry_fractal' and b.topography == 'read_topo_from_memory':
                b.faultsize = (int(fs[0]), int(fs[1]))
    for i, b in enumerate(blocks, 1):
        if str(pl.get('material_source', 'hardcode')) == 'file':
            # init_material_from_file (material.f90:706-747): lambda, mu, rho
            # from the block's three material_path files
            paths = bl.get(f'btp({i})%material_path', None) or []
            b.material = read_material_files([_data_path(p, base) for p in paths], b.n)
            continue
        # init_material's keyed medium, evaluated on the block's grid when a
        # simulation is set up (Block.medium_material)
        b.medium = problem if problem in MEDIA else None
    response = pl.get('response', 'elastic')
    model = pl.get('plastic_model', 'default')
    plastic = 0
    if response == 'plastic':
        plastic = 1 if model == 'scec' else 2
    prob = Problem(blocks=blocks, fault=fault_params(problem, coupling),
                   cfl=float(pl.get('cfl', 0.5)), t_final=float(pl.get('t_final', 0.0)),
                   plastic=plastic, name=pl.get('name', 'default'), w_stride=int(pl.get('w_stride', 1)),
                   w_fault=bool(pl.get('w_fault', False)), moment=read_moment(text),
                   mollify_source=bool(pl.get('mollify_source', False)),
                   fd_type=pl.get('fd_type', 'traditional'), order=int(pl.get('order', 5)))
    if response == 'anelastic':
        al = read_namelist(text, 'anelastic_list')
        prob.anelastic = (float(al.get('c', 1.0)), float(al.get('weight_exp', 0.0)), float(al.get('fref', 1.0)))
    if plastic == 2:
        # per-block mu_beta_eta (plastic_material.f90:38-44); one set for both blocks here
        mbe = [tuple(float(v) for v in bl.get(f'btp({i})%mu_beta_eta', (0.0, 0.0, 0.0))) for i in (1, 2)]
        if mbe[0] != mbe[1]:
            raise ValueError('the two blocks give different mu_beta_eta; the library takes one set')
        prob.pl_mu, prob.pl_beta, prob.pl_eta = mbe[0]
    if plastic:
        f = fault_params(problem, 'slip-weakening_friction') if problem == 'TPV28' else prob.fault
        prob.s0_kind = 1 if problem in ('TPV26', 'TPV27') else 2 if problem in ('TPV29', 'TPV30') else 0
        if problem == 'TPV30':
            prob.pl_c, prob.pl_nu, prob.pl_Tv = 1.18, 0.1680, 0.05
        if prob.fault.stress_mode == 0 and problem == 'TPV28':
            prob.fault.s0 = f.s0
    return prob
