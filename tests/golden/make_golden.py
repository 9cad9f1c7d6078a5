"""Writes the golden fixtures under tests/golden/ (run in a container that has
the reference tree; the fixtures are committed, the reference is not).

* vie_reference_tensors.json -- the six independent NumPy Galerkin tensors of
  the reference's tests/vie-reference.csv (orders 36/52), as data.
* ref_operator_small.json -- products and GMRES solutions of the reference's
  own compiled FFT operator (oracle/_ref) on the lattices of
  tests/vie_fft_test.c:45-46 and a 4x4x4 block, for fixed inputs.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
REF = os.environ.get("REFERENCE_ROOT", "/root/reference")


def cplx(a):
    a = np.asarray(a, dtype=np.complex128).reshape(-1)
    return [[float(v.real), float(v.imag)] for v in a]


def tensors():
    rows = []
    with open(os.path.join(REF, "tests", "vie-reference.csv")) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            p = line.strip().split(",")
            num = [float(v) for v in p[1:]]
            rows.append({"name": p[0], "k0": num[0], "test_center": num[1:4], "test_size": num[4:7],
                         "source_center": num[7:10], "source_size": num[10:13],
                         "tensor": [[num[13 + 2 * i], num[14 + 2 * i]] for i in range(9)]})
    with open(os.path.join(HERE, "vie_reference_tensors.json"), "w") as f:
        json.dump({"source": "reference tests/vie-reference.csv (generate_vie_reference.py, orders 36/52)",
                   "tolerance": "3e-10 + 3e-8 |N| (tests/vie_kernel_test.c:226)", "cases": rows}, f, indent=1)


def operators():
    from oracle import oracle as O
    R = O.ReferenceLibrary()
    R.L.refshim_backend_set(b"builtin")
    block = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(-1).tolist()
    cases = [
        ("holes_5x3x2", (5, 3, 2), [0, 0, 0, 4, 2, 1, 1, 0, 0, 0, 1, 0, 0, 0, 1, 1, 1, 1], 0.25, 0.8, 3 - 0.4j),
        ("far_9x5x3", (9, 5, 3), [0, 0, 0, 8, 4, 2, 4, 1, 0], 0.2, 4.0, 3.0),
        ("one_cell", (1, 1, 1), [0, 0, 0], 0.25, 0.2, 3.0),
        ("block_4", (4, 4, 4), block, 0.1, 2.0, 5 - 1j),
    ]
    out = []
    for name, shape, xyz, h, k, er in cases:
        p = R.plan(shape, xyz, h, k, er)
        n = 3 * (len(xyz) // 3)
        x = np.sin(0.7 * np.arange(n)) + 1j * np.cos(0.3 * np.arange(n))
        y = p.apply(x)
        pz = p.precondition(x)
        ok, sol, rep = p.gmres(y, rtol=1e-11, maxit=150)
        out.append({"name": name, "shape": list(shape), "xyz": list(xyz), "pitch": h, "k0": k,
                    "epsilon": [er.real, er.imag] if isinstance(er, complex) else [float(er), 0.0],
                    "x": cplx(x), "y": cplx(y), "precondition": cplx(pz), "solve": cplx(sol),
                    "iterations": rep.iterations, "restarts": rep.restarts,
                    "relative_residual": rep.relative_residual})
    with open(os.path.join(HERE, "ref_operator_small.json"), "w") as f:
        json.dump({"source": "reference vie_fft.c/vie_iterative.c compiled by oracle/Makefile (builtin FFT)",
                   "x": "sin(0.7 i) + j cos(0.3 i)", "cases": out}, f)


if __name__ == "__main__":
    tensors()
    operators()
    print("golden fixtures written")
