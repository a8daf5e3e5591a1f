"""Converts the reference's pickled test fixtures to .npy without unpickling.

The reference's golden results (tests/data/oculomotor_{euler,rk,ie,im}.pkl of
avramidis/sodecl, read by tests/sodecl_python_tests.py:49,92,135,178) are
pickled numpy arrays.  Nothing here executes pickle: the opcode stream is
walked with pickletools.genops, and the array is rebuilt from the three
pieces numpy's reduce protocol stores as plain data -- the shape tuple, the
dtype string and the Fortran-order flag -- plus the single raw-bytes payload.

Usage:  python tests/golden/decode_reference_fixtures.py [REFERENCE_ROOT]
Writes tests/golden/oculomotor_<solver>.npy (shape (orbits*neq, nsnap)).
"""
import os
import pickletools
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def decode(path):
    data = open(path, "rb").read()
    ops = list(pickletools.genops(data))
    payload = [arg for op, arg, _ in ops if op.name in ("BINBYTES", "SHORT_BINBYTES", "BINBYTES8")]
    payload = max(payload, key=len)           # the array data; the other is b"b"
    # shape: the first TUPLE2/TUPLE3 after the first MARK that follows REDUCE
    ints, shape, fortran, dtype = [], None, None, None
    for i, (op, arg, _) in enumerate(ops):
        if op.name == "MARK" and shape is None and i > 0:
            j = i + 1
            version = ops[j][1]; j += 1
            dims = []
            while ops[j][0].name.startswith("BININT"):
                dims.append(ops[j][1]); j += 1
            if ops[j][0].name in ("TUPLE1", "TUPLE2", "TUPLE3") and len(dims) >= 1 and version == 1:
                shape = tuple(dims)
        if op.name == "BINUNICODE" and arg in ("f8", "<f8") and dtype is None:
            dtype = "f8"
        if op.name == "BINUNICODE" and arg in ("<", ">", "|"):
            endian = arg
    # Fortran flag: the boolean right before the payload
    for i, (op, arg, _) in enumerate(ops):
        if op.name in ("BINBYTES", "SHORT_BINBYTES", "BINBYTES8") and arg is payload:
            fortran = ops[i - 1][0].name == "NEWTRUE"
    assert shape is not None and dtype == "f8" and endian == "<", (shape, dtype)
    arr = np.frombuffer(payload, dtype="<f8")
    assert arr.size == int(np.prod(shape))
    return arr.reshape(shape, order="F" if fortran else "C").copy(), fortran


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for name in ("euler", "rk", "ie", "im"):
        arr, fortran = decode(os.path.join(ref, "tests", "data", f"oculomotor_{name}.pkl"))
        out = os.path.join(HERE, f"oculomotor_{name}.npy")
        np.save(out, arr)
        print(f"{out}: shape={arr.shape} fortran={fortran} min={arr.min():.6g} max={arr.max():.6g}")


if __name__ == "__main__":
    main()
