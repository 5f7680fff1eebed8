"""Runs the C++ program that uses the drop-in headers (include/frovedis_hip/)
with the reference's own signatures on std::vector storage."""
import subprocess

import pytest

from test_abi import _build_shim

pytestmark = pytest.mark.gpu


def test_cpp_dropin_program(tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    exe = tmp_path / "shim_test"
    _build_shim(exe)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "shim ok" in r.stdout
