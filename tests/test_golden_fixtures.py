"""The committed fixtures under tests/golden/ are what make_golden.py writes
from tests/golden_values.py (so the C++ suite, which reads the text file,
checks the same reference constants as the Python tests)."""
import os

import numpy as np

from tests.golden import make_golden

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_reference_vectors_file_matches_golden_values():
    want = make_golden.golden_vectors()
    got = {}
    with open(os.path.join(HERE, "reference_vectors.txt")) as f:
        for line in f:
            name, count, *vals = line.split()
            assert int(count) == len(vals)
            got[name] = [float(v) for v in vals]
    assert got == want
    assert got["DOT_AB"] == [604.0] and len(got["T234A"]) == 24


def test_python_written_bt_fixture():
    B = make_golden.read_bt(os.path.join(HERE, "b_d.bt"))
    i, j, k = np.meshgrid(np.arange(3), np.arange(3), np.arange(2), indexing="ij")
    assert np.array_equal(B, 1.0 + k + 2 * j + 6 * i)
