"""The reference's C++ call surface (include/litho_cpp.hpp) on the GPU: runs the
C++ test program tests/cpp/test_litho_cpp.cpp (built by build()), which
follows the reference's tests/test_ctm_optimizer.cpp test_synthetic and the
full-chip consistency checks."""
import subprocess

import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu


def test_cpp_call_surface():
    exe = PKG / "test_litho_cpp"
    if not exe.exists():
        pytest.fail(f"{exe} is missing: run __graft_entry__.build() first")
    print("case C++ call surface", flush=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "all checks passed" in p.stdout
