"""The Fortran binding (sucrose_amd/fortran/sucrose_gpu.f90) used from the
reference's own modules: the C1 Lennard-Jones NVE run driven once by the
reference on the host and once through the binding on the GPU
(tests/fortran/lj_dropin.f90)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "lj_dropin")

needs_exe = pytest.mark.skipif(not os.path.exists(EXE),
                               reason="drop-in driver not built (make -C tests/fortran, after make -C oracle ref)")


@needs_exe
def test_binding_stops_cleanly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "sx_create failed" in r.stdout


@pytest.mark.gpu
@needs_exe
def test_lj_nve_dropin_matches_reference_loop():
    print("lj_dropin: 864 LJ atoms, 1000 NVE steps, reference loop vs GPU binding", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "DROPIN OK" in r.stdout
