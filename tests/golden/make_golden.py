"""Writes the small fixtures under tests/golden/.

b_d.bt: the 3x3x2 FP64 tensor B[i][j][k] = 1 + k + 2j + 6i that GPUtils'
Python test writes before the C++ suite reads it back (python/test/test.py
and test/testTensor.cu X:193-202), in the .bt format: three little-endian
uint64 (rows, cols, mats) followed by the column-major data with the matrix
index slowest (include/tensor.cuh:762-825).

reference_vectors.txt: every list and number of tests/golden_values.py (the
reference suite's inputs and expected values), one per line as
``NAME COUNT v0 v1 ...`` with values printed to round-trip exactly, for the
C++ drop-in suite (tests/cpp/api_suite.cpp), which has no Python."""
import os

import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def write_bt(x: np.ndarray, path: str) -> None:
    x = np.asarray(x)
    if x.ndim > 3:
        raise ValueError("at most 3 dimensions")
    shape = x.shape + (1,) * (3 - x.ndim)
    rows, cols, mats = shape
    data = np.asarray(x).reshape(rows, cols, mats)
    with open(path, "wb") as f:
        f.write(np.array([rows, cols, mats], dtype="<u8").tobytes())
        f.write(np.ascontiguousarray(np.transpose(data, (2, 1, 0))).tobytes())


def read_bt(path: str, dtype=np.float64) -> np.ndarray:
    with open(path, "rb") as f:
        rows, cols, mats = np.frombuffer(f.read(24), dtype="<u8")
        data = np.frombuffer(f.read(), dtype=dtype)
    return np.transpose(data.reshape(int(mats), int(cols), int(rows)), (2, 1, 0))


def golden_vectors() -> dict:
    """Name -> list of floats for every numeric constant of golden_values."""
    import golden_values as gv
    out = {}
    for name in sorted(vars(gv)):
        if not name.isupper():
            continue
        v = getattr(gv, name)
        if isinstance(v, (int, float)):
            out[name] = [float(v)]
        elif isinstance(v, (list, tuple)) and all(isinstance(e, (int, float)) for e in v):
            out[name] = [float(e) for e in v]
        elif isinstance(v, dict):
            for key, e in sorted(v.items()):
                out[f"{name}_{key.upper()}"] = [float(e)]
    return out


def write_vectors(path: str) -> None:
    with open(path, "w") as f:
        for name, vals in golden_vectors().items():
            f.write(name + " " + str(len(vals)) + " " + " ".join(repr(x) for x in vals) + "\n")


def main():
    B = np.array([[[1, 2], [3, 4], [5, 6]],
                  [[7, 8], [9, 10], [11, 12]],
                  [[13, 14], [15, 16], [17, 18]]], dtype=np.float64)
    write_bt(B, os.path.join(HERE, "b_d.bt"))
    write_vectors(os.path.join(HERE, "reference_vectors.txt"))


if __name__ == "__main__":
    main()
