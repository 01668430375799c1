"""Test inputs shared by the CPU and GPU suites: formulas, plus the reference's INLA test
matrix data/A_126_126_ns42_nt3.dat stored as data (golden/inla_126_csc.npz)."""
from __future__ import annotations

import json
import os

import numpy as np

from oracle import oracle

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def goldens() -> dict:
    with open(os.path.join(GOLDEN_DIR, "goldens.json")) as f:
        return json.load(f)


def inla_126():
    """The reference's INLA matrix (ns 42, nt 3, nd 0) as lower CSC: (ia, ja, a, ns, nt, nd)."""
    z = np.load(os.path.join(GOLDEN_DIR, "inla_126_csc.npz"))
    return z["ia"], z["ja"], z["a"], int(z["ns"]), int(z["nt"]), int(z["nd"])


def inla_126_sha256() -> str:
    return str(np.load(os.path.join(GOLDEN_DIR, "inla_126_csc.npz"))["sha256"])


def rhs_formula(n: int, nrhs: int = 1) -> np.ndarray:
    """Deterministic right-hand side used by the golden vectors: b[k, j] = cos(k + 0.5 j) + 0.25."""
    k = np.arange(n)[:, None]
    j = np.arange(nrhs)[None, :]
    return np.cos(k + 0.5 * j) + 0.25


def spde_like(nx: int = 6, ny: int = 7, nt: int = 3, nd: int = 0, kappa2: float = 0.3, rho: float = 0.9):
    """INLA-style spatio-temporal precision (the class of data/A_126_126_ns42_nt3.dat):
    Q_s = K K with K = kappa^2 I + (grid Laplacian of nx x ny), temporal AR(1) precision
    Q_t (tridiagonal), Q = Q_t (x) Q_s: block tridiagonal with sparse blocks, SPD, not
    diagonally dominant.  With nd > 0, nd dense fixed-effect rows are appended (small
    covariates, corner 1 + sum of their squares), which keeps it SPD.
    Returns (A dense, ns, nt, nd)."""
    ns = nx * ny
    Lg = np.zeros((ns, ns))
    for x in range(nx):
        for y in range(ny):
            i = x * ny + y
            for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                xx, yy = x + dx, y + dy
                if 0 <= xx < nx and 0 <= yy < ny:
                    Lg[i, i] += 1.0
                    Lg[i, xx * ny + yy] -= 1.0
    K = kappa2 * np.eye(ns) + Lg
    Qs = K @ K
    Qt = np.zeros((nt, nt))
    for t in range(nt):
        Qt[t, t] = 1.0 + (rho * rho if 0 < t < nt - 1 else 0.0)
        if t + 1 < nt:
            Qt[t, t + 1] = Qt[t + 1, t] = -rho
    Qt += 0.01 * np.eye(nt)
    Q = np.kron(Qt, Qs)
    n = ns * nt + nd
    A = np.zeros((n, n))
    A[:ns * nt, :ns * nt] = Q
    if nd > 0:
        rng = np.random.default_rng(1234)
        Xc = 0.05 * rng.standard_normal((nd, ns * nt))
        A[ns * nt:, :ns * nt] = Xc
        A[:ns * nt, ns * nt:] = Xc.T
        # Schur complement of the arrow stays SPD: corner >> X Q^-1 X^T
        A[ns * nt:, ns * nt:] = np.eye(nd) * (1.0 + 10.0 * np.abs(Xc).sum())
    return A, ns, nt, nd


def closed_form(ns: int, nt: int, nd: int):
    ia, ja, a = oracle.closed_form_csc(ns, nt, nd)
    return ia, ja, a


def complex_symmetric(ns: int, nt: int, nd: int, seed: int = 0, imag_scale: float = 0.4):
    """A complex-symmetric (A = A^T) block-arrow matrix: the generated real benchmark matrix
    plus i times a symmetric N(0, imag_scale^2) matrix on the same pattern (and an imaginary
    diagonal), so that every pivot of L L^T is complex and nonzero.  Returns dense A."""
    A = oracle.mf_to_dense(oracle.generate(seed, ns, nt, nd), ns, nt, nd).astype(np.complex128)
    rng = np.random.default_rng(seed + 17)
    n = A.shape[0]
    M = np.zeros((n, n))
    for c in range(n):
        bc = c // ns
        for r in range(c, n):
            br = r // ns
            if c >= ns * nt or r >= ns * nt or br == bc or br == bc + 1:
                M[r, c] = rng.standard_normal()
    M = np.tril(M) + np.tril(M, -1).T
    return A + 1j * imag_scale * M


def complex_csc_lower(A: np.ndarray, ns: int, nt: int, nd: int):
    """Lower-triangle CSC (uint64 indices, complex values) of the block-arrow pattern of A."""
    n = A.shape[0]
    ia, ja, a = [0], [], []
    for c in range(n):
        bc = c // ns
        for r in range(c, n):
            br = r // ns
            if c >= ns * nt or r >= ns * nt or br == bc or br == bc + 1:
                if A[r, c] != 0:
                    ja.append(r)
                    a.append(A[r, c])
        ia.append(len(ja))
    return np.array(ia, dtype=np.uint64), np.array(ja, dtype=np.uint64), np.array(a, dtype=np.complex128)
