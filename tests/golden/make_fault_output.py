"""Write the fault-output and station-seismogram fixtures of
tests/golden/fault_output/ and tests/golden/seismogram/ from the reference
build.

Runs only in the development container: it needs the reference Fortran built
by `make -C oracle ref` (oracle/_ref/waveqlab3d, never committed) and MPICH's
mpiexec.  Each case is the golden case of the same name (make_golden.py) run
through the reference's own main program with fault output switched on
(`w_fault = T`), two MPI ranks (one block per rank), for `t_final` with an
output every `w_stride` steps.  The eight files the reference writes,
<name>_interface.{Uface,Vface,S,Uhat,Vhat,Svel,trup,state} (fault_output.f90:13-135),
are copied unchanged next to the input file that produced them.

The seismogram cases switch fault output off and station output on
(&output_list output_seismograms = T): the station list sits in the input
file between the '!---begin:station_list---' / '!---end:station_list---'
lines, or in an external file (station_list = 'extfile'), and the reference
writes one '(4f15.10)' file per station and block (seismogram.f90:12-489);
every file it writes is copied, with its path below the run directory
(seismogram.f90:113-125, 377-392), into tests/golden/seismogram/<case>/.
"""
import os
import shutil
import subprocess
import tempfile
import sys

from make_golden import CASES, HERE, MPIEXEC, ROOT, namelist

REF_MAIN = os.path.join(ROOT, 'oracle', '_ref', 'waveqlab3d')
OUT = os.path.join(HERE, 'fault_output')
EXTS = ('Uface', 'Vface', 'S', 'Uhat', 'Vhat', 'Svel', 'trup', 'state')

# case: (golden case, t_final, w_stride)
RUNS = {'fo_cart_sw': ('cart_sw', 0.3, 2), 'fo_curv_sw': ('curv_sw', 0.3, 2)}

SG_OUT = os.path.join(HERE, 'seismogram')
# stations: on a node, between nodes, on the fault (both blocks), on the free
# surface, inside block 2, and off both blocks
STATIONS = ['-1.5 7.5 0.0', '-0.8 6.9 0.37', '0.0 7.5 0.0', '-2.0 6.0 -0.25', '1.2 8.1 -0.6', '5.0 7.5 0.0']
# case: (golden case, t_final, w_stride, &output_list body, station list in the input file?, extra files)
SG_RUNS = {
    'sg_curv_sw': ('curv_sw', 0.3, 1, " output_seismograms = T,\n output_fault_topo = F /", True, {}),
    'sg_cart_sw': ('cart_sw', 0.3, 3,
                   " output_seismograms = T,\n station_xyz_index = T,\n station_list = 'extfile',\n"
                   " station_list_file = 'stations.txt',\n station_file_directory = 'seis_out',\n"
                   " output_fault_topo = F /", False,
                   {'stations.txt': '# station list\n!---begin:station_list---\n' + '\n'.join(STATIONS) +
                    '\n!---end:station_list---\n'}),
}


def run_sg(name, case, t_final, w_stride, output_list, inline, extra, work):
    os.makedirs(work, exist_ok=True)
    kw = dict(CASES[case][0])
    kw['t_final'] = t_final
    text = namelist(name, **kw).replace(' w_stride = 1,', f' w_stride = {w_stride},')
    old = " output_seismograms = F,\n output_fault_topo = F /"
    assert old in text and ' w_fault = F /' in text
    text = text.replace(old, output_list)
    if inline:
        text += '\n!---begin:station_list---\n' + '\n'.join(STATIONS) + '\n!---end:station_list---\n'
    inp = os.path.join(work, f'{name}.in')
    with open(inp, 'w') as f:
        f.write(text)
    for fn, body in extra.items():
        with open(os.path.join(work, fn), 'w') as f:
            f.write(body)
    for d in ('data', 'seismogram/block1/data', 'seismogram/block2/data', 'seis_out/block1/data',
              'seis_out/block2/data'):
        os.makedirs(os.path.join(work, d), exist_ok=True)
    r = subprocess.run([MPIEXEC, '-n', '2', REF_MAIN, f'{name}.in'], cwd=work, capture_output=True, text=True,
                       timeout=900)
    if r.returncode != 0:
        print(r.stdout[-3000:], r.stderr[-3000:])
        raise SystemExit(f'{name} failed')
    dst = os.path.join(SG_OUT, name)
    shutil.rmtree(dst, ignore_errors=True)
    os.makedirs(dst)
    shutil.copy(inp, os.path.join(dst, f'{name}.in'))
    for fn in extra:
        shutil.copy(os.path.join(work, fn), os.path.join(dst, fn))
    for top in ('seismogram', 'seis_out'):
        for root, _, files in os.walk(os.path.join(work, top)):
            for fn in files:
                if fn.endswith('.dat'):
                    rel = os.path.relpath(os.path.join(root, fn), work)
                    os.makedirs(os.path.dirname(os.path.join(dst, rel)), exist_ok=True)
                    shutil.copy(os.path.join(root, fn), os.path.join(dst, rel))
                    print(f'{name}: {rel}')


def run(name, case, t_final, w_stride, work):
    os.makedirs(work, exist_ok=True)
    kw = dict(CASES[case][0])
    kw['t_final'] = t_final
    text = namelist(name, **kw).replace(' w_stride = 1,', f' w_stride = {w_stride},')
    text = text.replace(' w_fault = F /', ' w_fault = T /')
    assert f'w_stride = {w_stride}' in text and 'w_fault = T' in text
    inp = os.path.join(work, f'{name}.in')
    with open(inp, 'w') as f:
        f.write(text)
    for d in ('data', 'seismogram/block1/data', 'seismogram/block2/data'):
        os.makedirs(os.path.join(work, d), exist_ok=True)
    r = subprocess.run([MPIEXEC, '-n', '2', REF_MAIN, f'{name}.in'], cwd=work, capture_output=True, text=True,
                       timeout=900)
    if r.returncode != 0:
        print(r.stdout[-3000:], r.stderr[-3000:])
        raise SystemExit(f'{name} failed')
    os.makedirs(OUT, exist_ok=True)
    shutil.copy(inp, os.path.join(OUT, f'{name}.in'))
    for e in EXTS:
        src = os.path.join(work, 'data', f'{name}_interface.{e}')
        shutil.copy(src, os.path.join(OUT, f'{name}_interface.{e}'))
        print(f'{name}_interface.{e}: {os.path.getsize(src)} bytes')


if __name__ == '__main__':
    if not os.path.exists(REF_MAIN):
        sys.exit('build the reference first: make -C oracle ref')
    work = tempfile.mkdtemp(prefix='wql_golden_')
    for nm in sys.argv[1:] or list(RUNS) + list(SG_RUNS):
        if nm in SG_RUNS:
            run_sg(nm, *SG_RUNS[nm], os.path.join(work, nm))
        else:
            case, t_final, w_stride = RUNS[nm]
            run(nm, case, t_final, w_stride, os.path.join(work, nm))
