"""Write the fault-output fixtures of tests/golden/fault_output/ from the
reference build.

Runs only in the development container: it needs the reference Fortran built
by `make -C oracle ref` (oracle/_ref/waveqlab3d, never committed) and MPICH's
mpiexec.  Each case is the golden case of the same name (make_golden.py) run
through the reference's own main program with fault output switched on
(`w_fault = T`), two MPI ranks (one block per rank), for `t_final` with an
output every `w_stride` steps.  The eight files the reference writes,
<name>_interface.{Uface,Vface,S,Uhat,Vhat,Svel,trup,state} (fault_output.f90:13-135),
are copied unchanged next to the input file that produced them.
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
    for nm in sys.argv[1:] or list(RUNS):
        case, t_final, w_stride = RUNS[nm]
        run(nm, case, t_final, w_stride, os.path.join(work, nm))
