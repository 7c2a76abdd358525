"""Runs the C++ drop-in API test programs (tests/cpp/*.cpp) and the C programs of the libviennacl-style
API (tests/cpp/*.c), built by `make -C viennacl-dev_amd tests`, on the device.  Each program prints one FAIL line per failed check and exits nonzero."""
import pathlib
import subprocess

import pytest

BIN = pathlib.Path(__file__).resolve().parent / "cpp" / "bin"
SOURCES = sorted(list((pathlib.Path(__file__).resolve().parent / "cpp").glob("*.cpp")) +
                 list((pathlib.Path(__file__).resolve().parent / "cpp").glob("*.c")))


@pytest.mark.gpu
@pytest.mark.parametrize("src", SOURCES, ids=[s.stem for s in SOURCES])
def test_cpp_program(dev, src):
    exe = BIN / src.stem
    assert exe.exists(), f"{exe} is not built (make -C viennacl-dev_amd tests)"
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=110)
    print(p.stdout[-4000:])
    assert p.returncode == 0, p.stdout[-4000:] + p.stderr[-2000:]
    assert "ALL PASSED" in p.stdout


def test_cpp_programs_are_built():
    missing = [s.stem for s in SOURCES if not (BIN / s.stem).exists()]
    assert not missing, f"C++ API tests not built: {missing} (make -C viennacl-dev_amd tests)"
