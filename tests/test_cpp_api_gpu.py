"""Runs the C++ drop-in API suite (tests/cpp/api_suite.cpp, compiled
against include/tensor.hpp; its inputs and expected values come from
tests/golden/reference_vectors.txt, the reference suite's constants)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "api_suite")


@pytest.mark.gpu
def test_cpp_api_suite(tmp_path):
    assert os.path.exists(EXE), "tests/cpp/api_suite is not built (python -m gputils_amd.build)"
    env = dict(os.environ, GPUTILS_GOLDEN_DIR=os.path.join(ROOT, "tests", "golden"))
    p = subprocess.run([EXE], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=600)
    print(p.stdout)
    print(p.stderr)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]


def test_cpp_api_compiles_with_plain_gxx(tmp_path):
    """The drop-in header needs only a C++17 compiler and the C ABI."""
    src = os.path.join(ROOT, "tests", "cpp", "api_suite.cpp")
    out = tmp_path / "t.o"
    p = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        f"-I{os.path.join(ROOT, 'include')}", src], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-4000:]
