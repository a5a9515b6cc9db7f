"""The reference's fault output files, written from the device state.

WaveQLab3D writes eight files per run (fault_output.f90:13-135), appending one
record per output step: raw little-endian FP32 (ps = 4, mpi3dbasic.f90:13),
each record the Fortran array (nr, ns, nfields) of the fault plane, i.e. C
order [nfields][ns][nr]:

    <name>_interface.Uface  9  block 1's fields on the fault  F(pq1, :, :, 1:9)
    <name>_interface.Vface  9  block 2's fields on the fault  F(mq2, :, :, 1:9)
    <name>_interface.S      4  slip (n, m, l) and cumulative |slip|
    <name>_interface.Uhat   6  slip rate v^-u^ (1-3), P^ traction + prestress (4-6)
    <name>_interface.Vhat   6  slip rate v^-u^ (1-3), Q^ traction + prestress (4-6)
    <name>_interface.Svel   3  slip rate (Uhat 1-3)
    <name>_interface.trup   1  rupture time
    <name>_interface.state  1  state variable W

The records of step n are written inside its first RK stage (time_step.f90:137,
when mod(n-1, w_stride) == 0), after the interface solve and before the field
update: Uface, Vface, S and state hold the state at the start of step n
(domain.f90:741-779), Uhat, Vhat, Svel and trup the values the first stage of
step n forms (CouplingForcing.f90:213-228, Interface_Condition.f90:480-505).
The library keeps those stage-1 values on the device, so `write_step` takes
the first group before the step and the second after it.  The interface
state is block 1's copy (block index 0 here), as in the reference's 2-rank
run where block 1's rank writes it.
"""
import numpy as np

EXTS = ('Uface', 'Vface', 'S', 'Uhat', 'Vhat', 'Svel', 'trup', 'state')
NFIELDS = dict(Uface=9, Vface=9, S=4, Uhat=6, Vhat=6, Svel=3, trup=1, state=1)


class FaultOutput:
    """Appends records to the eight files of one run (`name` is the
    reference's problem name, e.g. 'data/tpv5')."""

    def __init__(self, name):
        self.files = {e: open(f'{name}_interface.{e}', 'wb') for e in EXTS}

    def close(self):
        for f in self.files.values():
            f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _write(self, ext, a):
        a = np.asarray(a)
        assert a.shape[0] == NFIELDS[ext], (ext, a.shape)
        self.files[ext].write(np.ascontiguousarray(a, dtype='<f4').tobytes())

    def write_step(self, sim, nsteps=1):
        """one output step followed by nsteps - 1 plain steps: the
        start-of-step records, the output step, its stage-1 records"""
        U, V = sim.fault_faces()
        f = sim.fault(0)
        self._write('Uface', U)
        self._write('Vface', V)
        self._write('S', f['S'])
        self._write('state', f['W'])
        sim.step(1)
        g = sim.fault(0)
        self._write('Uhat', g['Uhat'])
        self._write('Vhat', g['Vhat'])
        self._write('Svel', g['Uhat'][:3])
        self._write('trup', g['trup'])
        if nsteps > 1:
            sim.step(nsteps - 1)


def run(sim, nt, w_stride, name, stations=None, fault=True):
    """the reference's time loop (main.f90:83-96) with its output every
    `w_stride` steps (write_output, time_step.f90:137): nt steps, records at
    steps 1, 1 + w_stride, ...; fault output when `fault` (w_fault), and the
    station lines of `stations` (a seismogram.Seismograms) at the start of
    each output step"""
    out = FaultOutput(name) if fault else None
    try:
        n = 1
        while n <= nt:
            k = min(w_stride, nt - n + 1)
            if stations is not None:
                stations.write()
            if out is not None:
                out.write_step(sim, k)
            else:
                sim.step(k)
            n += k
    finally:
        if out is not None:
            out.close()
    sim.sync()


def read(name, ext, nr, ns):
    """records of one file as [nrecords][nfields][ns][nr] float32"""
    a = np.fromfile(f'{name}_interface.{ext}', '<f4')
    per = NFIELDS[ext] * nr * ns
    if a.size % per:
        raise ValueError(f'{name}_interface.{ext}: {a.size} values is not a whole number of records of {per}')
    return a.reshape(-1, NFIELDS[ext], ns, nr)
