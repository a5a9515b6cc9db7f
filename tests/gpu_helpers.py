"""Shared set-up for the GPU parity tests: run the same case on the HIP
library and on the oracle."""
import numpy as np

import golden_cases as G
import oracle as O
from wql.solver import Simulation


def load_pattern_state(sim, prob, od=None):
    st = G.pattern_state(prob)
    for ib in range(2):
        sim.set_fields(ib, st[ib]['F'])
        for s in range(6):
            if st[ib]['Q'][s] is not None:
                sim.set_pml(ib, s, st[ib]['Q'][s])
        sim.set_fault(ib, st[ib]['S'], st[ib]['W'])
        if st[ib]['eta'] is not None:
            sim.set_anelastic(ib, st[ib]['eta'])
        if od is not None:
            if st[ib]['eta'] is not None:
                od.blocks[ib]['eta'][...] = st[ib]['eta']
            od.blocks[ib]['F'][...] = st[ib]['F']
            for s in range(6):
                if st[ib]['Q'][s] is not None:
                    od.blocks[ib]['Q'][s][...] = st[ib]['Q'][s]
            od.blocks[ib]['iface']['S'][...] = st[ib]['S']
            od.blocks[ib]['iface']['W'][...] = st[ib]['W']
    return st


def oracle_rhs(prob, t):
    od = O.OracleDomain(prob)
    load_pattern_state(_NullSim(), prob, od)
    for ib in range(2):
        od.set_rates(ib)
    # set_rates (domain.f90:908-951): the moment-tensor sources follow the
    # volume rates of both blocks
    for ib in range(2):
        od.sources(ib, t)
    vol = [od.blocks[ib]['DF'].copy() for ib in range(2)]
    dqv = [[None if od.blocks[ib]['DQ'][s] is None else od.blocks[ib]['DQ'][s].copy() for s in range(6)]
           for ib in range(2)]
    for ib in (1, 2):
        od.boundary(ib - 1)
        od.interface(ib, t, 2)
    return od, vol, dqv


class _NullSim:
    def set_anelastic(self, *a): pass
    def set_fields(self, *a): pass
    def set_pml(self, *a): pass
    def set_fault(self, *a, **k): pass


def pulse_init(prob, args):
    """initial F per block for a traj case: pulse (if any) on zero fields"""
    out = []
    for ib, b in enumerate(prob.blocks):
        X = b.coords()
        F = np.zeros((9,) + X.shape[1:])
        if len(args) > 1:
            F[0:3] += G.pulse(prob, X, args[1:])
        out.append(F)
    return out
