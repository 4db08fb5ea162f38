"""The C++ shim with the reference's call shapes (spectre_amd/cpp/spectre_amd.hpp):
it compiles against the C ABI (CPU), and its test program passes on a HIP
device (GPU): exact derivatives of polynomials, oracle-identical
TimeDerivative results, lift scaling and error reporting."""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
EXE = ROOT / "tests" / "cpp" / "_build" / "test_shim"


def test_shim_compiles():
    subprocess.run(["make", "-s", "-C", str(ROOT / "tests" / "cpp")], check=True)
    assert EXE.exists()


@pytest.mark.gpu
def test_shim_program_passes_on_device():
    assert EXE.exists(), "build the shim test first (__graft_entry__.build())"
    out = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "OK (0 failures)" in out.stdout
