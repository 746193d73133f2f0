"""Writes tests/golden/reporters/*.txt: frames of the SUCROSE reference's
reporters (reporters.f90) for the seeded states of tests/_states.report_state.

The reference is compiled from its own Fortran sources by ``make -C oracle
ref``; each frame is written by its own Fortran runtime through
oracle/ref_harness.f90 (ref_report).  tests/test_reporters.py compares the
library's formatter with these files where the reference build is absent.

    python tests/golden/make_reporters.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import Reference  # noqa: E402
from tests._states import report_state  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reporters")
KINDS = {1: "state_data", 2: "trajectory", 3: "config", 4: "checkpoint"}


def main():
    os.makedirs(OUT, exist_ok=True)
    ref = Reference(private=True)
    for seed, n, nb in ((1, 20, 0), (2, 37, 12), (3, 5, 4)):
        st, rt = report_state(np.random.default_rng(seed), n, nbonds=nb)
        ref.load(st)
        for kind, name in KINDS.items():
            ref.report(kind, os.path.join(OUT, f"{name}_{seed}.txt"), 1000 + seed, rt)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
