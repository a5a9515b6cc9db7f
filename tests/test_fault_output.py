"""The reference's fault output files (fault_output.f90:13-135): layout, record
count and the GPU path's files against the ones the reference build wrote
for the same input (tests/golden/fault_output/, made by
tests/golden/make_fault_output.py)."""
import os
import sys

import numpy as np
import pytest

import golden_cases as G
from wql import fault_output as FO
from wql.problem import from_namelist

FO_DIR = os.path.join(G.GOLDEN, 'fault_output')
CASES = ['fo_cart_sw', 'fo_curv_sw']


def _ref(case):
    prob = from_namelist(os.path.join(FO_DIR, f'{case}.in'))
    return prob, os.path.join(FO_DIR, case)


def _nrec(prob):
    nt = prob.nt()
    return (nt - 1) // prob.w_stride + 1


@pytest.mark.parametrize('case', CASES)
def test_reference_files_have_the_documented_layout(case):
    """one record per output step of [nfields][ns][nr] FP32, Svel = Uhat 1-3"""
    prob, name = _ref(case)
    assert prob.w_fault and prob.w_stride == 2
    nr, ns = prob.blocks[0].n[1:]
    recs = {e: FO.read(name, e, nr, ns) for e in FO.EXTS}
    for e, a in recs.items():
        assert a.shape == (_nrec(prob), FO.NFIELDS[e], ns, nr), e
        assert np.isfinite(a).all(), e
    assert np.array_equal(recs['Svel'], recs['Uhat'][:, :3])
    # the first record is the initial state: no slip yet, fields at rest
    assert np.abs(recs['S'][0]).max() == 0.0
    assert np.abs(recs['Uface'][0][:3]).max() == 0.0


class _FakeSim:
    """stands in for Simulation: distinct values per call and step"""
    def __init__(self, nr, ns):
        self.nr, self.ns, self.n = nr, ns, 0

    def _a(self, nf, tag):
        return np.arange(nf * self.ns * self.nr, dtype=float).reshape(nf, self.ns, self.nr) + 1000 * self.n + tag

    def fault_faces(self):
        return self._a(9, 0.25), self._a(9, 0.5)

    def fault(self, ib):
        return dict(S=self._a(4, 0.125), W=self._a(1, 0.375), trup=self._a(1, 0.625), Uhat=self._a(6, 0.75),
                    Vhat=self._a(6, 0.875))

    def step(self, k):
        self.n += k

    def sync(self):
        pass


def test_writer_record_order(tmp_path):
    """start-of-step records before the step, stage-1 records after it, one
    output every w_stride steps, FP32 [nfields][ns][nr]"""
    nr, ns, nt, ws = 5, 4, 7, 3
    sim = _FakeSim(nr, ns)
    name = str(tmp_path / 'x')
    FO.run(sim, nt, ws, name)
    steps = [0, 3, 6]  # steps done before output steps 1, 4, 7
    ref = _FakeSim(nr, ns)
    for e in FO.EXTS:
        a = FO.read(name, e, nr, ns)
        assert a.shape == (3, FO.NFIELDS[e], ns, nr)
        for r, n0 in enumerate(steps):
            ref.n = n0 if e in ('Uface', 'Vface', 'S', 'state') else n0 + 1
            f = ref.fault(0)
            U, V = ref.fault_faces()
            want = dict(Uface=U, Vface=V, S=f['S'], state=f['W'], Uhat=f['Uhat'], Vhat=f['Vhat'],
                        Svel=f['Uhat'][:3], trup=f['trup'])[e]
            assert np.array_equal(a[r], want.astype(np.float32)), (e, r)


@pytest.mark.gpu
@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
@pytest.mark.parametrize('case', CASES)
def test_fault_output_matches_reference_build(case, precision, tmp_path):
    """the eight files of the HIP path against the reference build's, record
    by record.  FP64: the values agree to the FP64 path's 1e-12, so the FP32
    records are equal to within two float32 ulp (plus 1e-9 of the file's
    largest value for values that cancel to near zero).  FP32 volume: within
    1e-5 of the file's largest value."""
    from wql.solver import Simulation
    print(f'[gpu] fault output {case} {precision}', file=sys.stderr, flush=True)
    prob, ref_name = _ref(case)
    nr, ns = prob.blocks[0].n[1:]
    name = str(tmp_path / case)
    with Simulation(prob, precision=precision) as sim:
        if prob.fault.law == 2:
            sim.init_state()
        FO.run(sim, prob.nt(), prob.w_stride, name)
    for e in FO.EXTS:
        assert os.path.getsize(f'{name}_interface.{e}') == os.path.getsize(f'{ref_name}_interface.{e}'), e
        got = FO.read(name, e, nr, ns).astype(np.float64)
        ref = FO.read(ref_name, e, nr, ns).astype(np.float64)
        scale = max(np.abs(ref).max(), 1e-30)
        if precision == 'fp64':
            ok = np.abs(got - ref) <= 2.4e-7 * np.abs(ref) + 1e-9 * scale
        else:
            ok = np.abs(got - ref) <= 1e-5 * scale
        assert ok.all(), (e, float(np.abs(got - ref).max()), scale)
