"""The README driver loop (README.md:83-115) run on the oracle.

Mirrors ``ref_run_steps`` in oracle/ref_harness.f90 step for step so a run of
the C restatement can be compared bit for bit with the reference's own loop.
"""
from __future__ import annotations

from oracle.oracle import LGV, NHT, QQ, VV


def oracle_run_steps(O, st, nl, integ, kind, p, nsteps, step0, dt, damping=1.0):
    for t in range(nsteps):
        if integ == VV:
            O.integrate(VV, 0, st, dt)
        elif integ == NHT:
            O.integrate(NHT, 0, st, dt)
            O.integrate(QQ, 0, st, dt)
        elif integ == LGV:
            O.integrate(LGV, 0, st, dt, step=step0 + t, damping=damping)
        if O.needs_rebuild(st, nl):
            O.nlist_build(st, nl)
        O.call("or_initialize_forces", st)
        O.pair_loop(st, nl, kind, p)
        if integ == VV:
            O.integrate(VV, 1, st, dt)
        elif integ == NHT:
            O.integrate(NHT, 1, st, dt)
            O.integrate(QQ, 1, st, dt)
        elif integ == LGV:
            O.integrate(LGV, 1, st, dt, damping=damping)
        O.call("or_calculate_thermo", st)
    return st
