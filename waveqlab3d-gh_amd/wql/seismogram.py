"""Station seismograms in the reference's format (seismogram.f90).

The input file's &output_list switches station output on
(output_seismograms = T) and says where the stations are and where the
files go (seismogram.f90:54-104):

    station_list            'infile' (default): the lines between
                            '!---begin:station_list---' and
                            '!---end:station_list---' of the input file, or
                            per block '!---begin:station_listU---' (block 1) /
                            '..._listV---' (block 2) when that list is absent;
                            'extfile' / 'exfile': the same markers in
                            station_list_file
    station_file_directory  'seismogram' (default): files go to
                            <dir>/block<b>/ ('block<b>/' for '' or '.')
    station_xyz_index       F: <name>_<i>_<j>_<k>_block<b>.dat with the
                            station's 1-based node; T: <name>_<x>_<y>_<z>.dat
                            with the requested coordinates as f20.3

Each station line holds x y z; every block takes the node nearest to it
within half the local grid diagonal (Find_Coordinates, seismogram.f90:531-650,
here wql_find_stations), so a station on the fault has a file in both
blocks' directories, and one off a block none there.  Each output step
(every w_stride steps, at the start of the step, time_step.f90:137) appends
one '(4f15.10)' line: t, vx, vy, vz at the station's node
(seismogram.f90:460-478).
"""
import ctypes as C
import os
import re

import numpy as np

from . import native
from .native import check
from .problem import read_namelist

DEFAULTS = dict(output_seismograms=False, station_xyz_index=False, station_list='infile', station_list_file='',
                station_file_directory='seismogram')


def output_options(path):
    """&output_list of the input file with the reference's defaults"""
    opts = dict(DEFAULTS)
    with open(path) as f:
        nml = read_namelist(f.read(), 'output_list')
    for k in DEFAULTS:
        if k in nml:
            opts[k] = nml[k]
    return opts


def _between(lines, begin, end):
    try:
        i = lines.index(begin)
    except ValueError:
        return None
    out = []
    for ln in lines[i + 1:]:
        if ln == end:
            return out
        out.append(ln)
    raise ValueError(f'station list: {end} missing')


def station_coordinates(path, opts, block):
    """[n][3] requested coordinates of the stations block `block` (1 or 2)
    looks for (seismogram.f90:128-326)"""
    src = path
    if opts['station_list'] in ('extfile', 'exfile'):
        if not opts['station_list_file']:
            raise ValueError('station_list_file must be given with station_list = extfile')
        src = opts['station_list_file']
    with open(src) as f:
        lines = [ln.rstrip('\n').rstrip('\r') for ln in f]
    body = _between(lines, '!---begin:station_list---', '!---end:station_list---')
    if body is None:
        if src != path:
            raise ValueError('!---begin:station_list--- not found in the external station file')
        tag = 'U' if block == 1 else 'V'
        body = _between(lines, f'!---begin:station_list{tag}---', f'!---end:station_list{tag}---')
        if body is None:
            raise ValueError('error reading seismogram station list')
    xyz = [[float(v.replace('d', 'e').replace('D', 'e')) for v in re.split(r'[\s,]+', ln.strip())[:3]]
           for ln in body]
    return np.array(xyz, dtype=np.float64).reshape(-1, 3)


def find_stations(X, xyz):
    """0-based (i, j, k) node of each station in the block grid
    X[3][ns][nr][nq], -1 where the station is off the block"""
    X = np.ascontiguousarray(X, dtype=np.float64)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    _, ns, nr, nq = X.shape
    ijk = np.empty((len(xyz), 3), dtype=np.int32)
    check(native.lib().wql_find_stations(X.ctypes.data_as(native.dp), nq, nr, ns, len(xyz),
                                         xyz.ctypes.data_as(native.dp), ijk.ctypes.data_as(C.POINTER(C.c_int))))
    return ijk


def fmt_f(v, w, d):
    """Fortran Fw.d: right-justified fixed point, asterisks when too wide"""
    s = f'{v:{w}.{d}f}'
    return '*' * w if len(s) > w else s


def file_name(name, opts, block, ijk, xyz):
    """path of a station's file (seismogram.f90:113-125, 377-392)"""
    d = str(opts['station_file_directory']).strip()
    bdir = f'{d}/block{block}' if d and d != '.' else f'block{block}'
    if opts['station_xyz_index']:
        xs, ys, zs = (fmt_f(v, 20, 3).strip() for v in xyz)
        fn = f'{name.strip()}_{xs}_{ys}_{zs}.dat'
    else:
        i, j, k = (int(v) + 1 for v in ijk)
        fn = f'{name.strip()}_{i}_{j}_{k}_block{block}.dat'
    return os.path.join(bdir, fn)


class Seismograms:
    """Station files of one run: init_seismogram for both blocks on
    construction (stations located on `coords`, the blocks' node
    coordinates), write_seismogram per `write`, destroy_seismogram on close."""

    def __init__(self, sim, inp, coords, name=None, opts=None):
        from .problem import read_namelist as _rn
        self.sim = sim
        self.opts = output_options(inp) if opts is None else opts
        if name is None:
            with open(inp) as f:
                name = _rn(f.read(), 'problem_list').get('name', 'default')
        self.files = [[], []]
        self.nodes = [np.zeros((0, 3), np.int32), np.zeros((0, 3), np.int32)]
        if not self.opts['output_seismograms']:
            return
        for ib in range(2):
            xyz = station_coordinates(inp, self.opts, ib + 1)
            ijk = find_stations(coords[ib], xyz)
            found = (ijk >= 0).all(axis=1)
            nodes = np.ascontiguousarray(ijk[found], dtype=np.int32)
            for s in np.nonzero(found)[0]:
                path = file_name(name, self.opts, ib + 1, ijk[s], xyz[s])
                os.makedirs(os.path.dirname(path) or '.', exist_ok=True)
                self.files[ib].append(open(path, 'w'))
            self.nodes[ib] = nodes
            check(native.lib().wql_set_stations(sim.ctx, ib, len(nodes), nodes.ctypes.data_as(C.POINTER(C.c_int))))

    def write(self):
        """one line per station at the simulation's current time and fields"""
        for ib in range(2):
            n = len(self.files[ib])
            if n == 0:
                continue
            v = np.empty((n, 3))
            check(native.lib().wql_get_stations(self.sim.ctx, ib, v.ctypes.data_as(native.dp)))
            t = self.sim.t
            for f, row in zip(self.files[ib], v):
                f.write(''.join(fmt_f(x, 15, 10) for x in (t, row[0], row[1], row[2])) + '\n')

    def close(self):
        for fl in self.files:
            for f in fl:
                f.close()
        self.files = [[], []]

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read(path):
    """[nrecords][4] t vx vy vz of a station file"""
    return np.loadtxt(path, ndmin=2)
