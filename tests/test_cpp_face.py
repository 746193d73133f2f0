"""The C++ face (include/sucrose.hpp), shaped like the reference's C++
classes, compiled against the library (tests/cpp/cpp_face.cpp): it builds,
it fails cleanly without a device, and on the device the pair-by-pair
calculate_forces(i, j) path reproduces the whole-loop calculate_all()."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "cpp_face")


def test_cpp_face_builds():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(EXE)


def test_cpp_face_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    if not os.path.exists(EXE):
        pytest.skip("not built")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "sucrose error" in r.stdout


@pytest.mark.gpu
def test_cpp_face_pairs_match_loop():
    print("cpp_face: 108 LJ atoms, 20 VV steps through the C++ face; pair-by-pair vs whole loop", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "CPP FACE OK" in r.stdout
