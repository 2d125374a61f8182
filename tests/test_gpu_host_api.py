"""The C++ host mirror of the reference interfaces (constitutiveModel,
constitutiveEq::New, sparseSolver<symmTensor>, fvMatrix<symmTensor>::solve)
driven from C++ the way rheoFoam calls them, checked against the oracle
inside the driver (tests/cpp/test_host_api.cpp, built by build())."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_host_api_against_oracle():
    exe = os.path.join(ROOT, "tests", "cpp", "test_host_api")
    assert os.path.exists(exe), "build() did not produce the C++ host-API driver"
    print("[case] C++ host API driver", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
