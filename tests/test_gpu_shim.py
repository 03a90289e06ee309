"""The C++ adapter (include/dmftgpu_shim.hpp) with the reference's call surface,
driven from C++ with a stand-in for Eigen::MatrixXcd, checked against the oracle."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "test_shim")


@pytest.mark.gpu
def test_cpp_shim_drop_in():
    assert os.path.exists(BIN), "built by __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "SHIM OK" in r.stdout
