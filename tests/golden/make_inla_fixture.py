"""Writes tests/golden/inla_126_csc.npz: the reference's INLA test matrix
data/A_126_126_ns42_nt3.dat (format of main.C:50-70: ``n n nnz``, ia[n+1], ja[nnz],
a[nnz]) as compressed numpy arrays, so the GPU suite can feed the reference's own
spatio-temporal precision matrix to the HIP path where the reference is not present.
The arrays are data only (a test vector the reference holds); the expected outputs are
in goldens.json["inla_126"] (dense numpy FP64, make_golden.py).

Run from the repository root: ``python tests/golden/make_inla_fixture.py``.
"""
from __future__ import annotations

import hashlib
import os

import numpy as np

SRC = "/root/reference/data/A_126_126_ns42_nt3.dat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "inla_126_csc.npz")


def main():
    raw = open(SRC, "rb").read()
    tok = raw.split()
    n, nnz = int(tok[0]), int(tok[2])
    p = 3
    ia = np.array([int(t) for t in tok[p:p + n + 1]], dtype=np.uint64)
    p += n + 1
    ja = np.array([int(t) for t in tok[p:p + nnz]], dtype=np.uint64)
    p += nnz
    a = np.array([float(t) for t in tok[p:p + nnz]], dtype=np.float64)
    np.savez_compressed(OUT, ia=ia, ja=ja, a=a, n=np.array(n), ns=np.array(42), nt=np.array(3), nd=np.array(0),
                        sha256=np.array(hashlib.sha256(raw).hexdigest()))
    print("wrote", OUT, "n", n, "nnz", nnz)


if __name__ == "__main__":
    main()
