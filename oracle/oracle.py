"""CPU oracle for the RGF hot path -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, by ``__graft_entry__.smoke()`` (as the checker) and by
``bench.py``'s cpu_baseline leg; the product package ``rgf_amd`` never imports
it.  It wraps ``librgf_oracle.so`` (a restatement of rjanalik/RGF's block
algorithm on host MKL over the reference's MF layout, see rgf_oracle.c) and
adds dense numpy helpers used as independent pins:

* ``mf_to_dense`` / ``dense_to_mf``: the MF layout of RGF.H:85-138 and
  CWC_utility.cu:998-1029 (getPos) restated in numpy;
* ``read_dat``: the CSC file format of main.C:50-70;
* ``closed_form_csc``: the formula behind the reference's data/A_9_9_ns3_nt3.dat
  and data/A_11_11_ns3_nt3_nd2.dat (A[c,c] = 10(c+1), A[r,c] = -0.1 (r-c) on the
  block-arrow pattern);
* ``reference_selinv_dense``: RGF.H:422-506's formula evaluated densely, to pin
  the compat selected inversion independently of the LAPACK restatement.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "librgf_oracle.so")
_lib = None
_lock = threading.Lock()

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_L = C.c_long


def build() -> None:
    """Compile librgf_oracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(_LIB_PATH):
            build()
        lib = C.CDLL(_LIB_PATH)
        lib.rgf_oracle_mf_size.argtypes = [_L, _L, _L]
        lib.rgf_oracle_mf_size.restype = _L
        lib.rgf_oracle_diag_pos.argtypes = [_L, _L, _L, _lp]
        lib.rgf_oracle_getpos.argtypes = [_L, _L, _L, _L, _L]
        lib.rgf_oracle_getpos.restype = _L
        lib.rgf_oracle_csc_to_mf.argtypes = [_up, _up, _dp, _L, _L, _L, _dp]
        lib.rgf_oracle_csc_to_mf.restype = _L
        lib.rgf_oracle_factor.argtypes = [_dp, _L, _L, _L]
        lib.rgf_oracle_logdet.argtypes = [_dp, _L, _L, _L]
        lib.rgf_oracle_logdet.restype = C.c_double
        lib.rgf_oracle_solve.argtypes = [_dp, _L, _L, _L, _dp, _L]
        lib.rgf_oracle_selinv.argtypes = [_dp, _L, _L, _L, _dp, C.c_int]
        lib.rgf_oracle_residual_norm.argtypes = [_up, _up, _dp, _L, _dp, _dp]
        lib.rgf_oracle_residual_norm.restype = C.c_double
        lib.rgf_oracle_residual_norm_normalized.argtypes = [_up, _up, _dp, _L, _dp, _dp]
        lib.rgf_oracle_residual_norm_normalized.restype = C.c_double
        lib.rgf_oracle_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        lib.rgf_oracle_gen_normal.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64]
        lib.rgf_oracle_gen_normal.restype = C.c_double
        lib.rgf_oracle_gen_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        lib.rgf_oracle_gen_uniform.restype = C.c_double
        lib.rgf_oracle_generate.argtypes = [C.c_uint64, C.c_uint32, _L, _L, _L, C.c_double,
                                            C.c_double, C.c_double, C.c_double, _dp]
        _lib = lib
        return lib


# ---------------------------------------------------------------- layout
def mf_size(ns: int, nt: int, nd: int) -> int:
    """Length of the MF array, RGF.H:95."""
    return 2 * (nt - 1) * ns * ns + ns * ns + (ns * nt + nd) * nd


def diag_pos(ns: int, nt: int, nd: int) -> np.ndarray:
    out = np.zeros(ns * nt + nd, dtype=np.int64)
    load().rgf_oracle_diag_pos(ns, nt, nd, out)
    return out


def getpos(r: int, c: int, ns: int, nt: int, nd: int) -> int:
    return int(load().rgf_oracle_getpos(r, c, ns, nt, nd))


def _snode(ns, nt, nd, c):
    return c * ns * (2 * ns + nd)


def _lda(ns, nt, nd, c):
    if c < nt - 1:
        return 2 * ns + nd
    if c < nt:
        return ns + nd
    return nd


def mf_to_dense(mf: np.ndarray, ns: int, nt: int, nd: int, symmetric: bool = True,
                lower_only_diag: bool = True) -> np.ndarray:
    """Dense n x n matrix from the logical blocks of an MF array (numpy restatement
    of the layout in RGF.H:114-138 / CWC_utility.cu:998-1029).  Diagonal blocks and
    the corner contribute their lower triangle; with ``symmetric`` the result is
    mirrored, otherwise the lower-triangular factor is returned."""
    n = ns * nt + nd
    A = np.zeros((n, n))
    for c in range(nt):
        lda = _lda(ns, nt, nd, c)
        s = _snode(ns, nt, nd, c)
        rows = lda
        P = mf[s:s + lda * ns].reshape(ns, lda).T   # column-major panel, rows x ns
        D = P[:ns, :]
        A[c * ns:(c + 1) * ns, c * ns:(c + 1) * ns] = np.tril(D) if lower_only_diag else D
        if c < nt - 1:
            A[(c + 1) * ns:(c + 2) * ns, c * ns:(c + 1) * ns] = P[ns:2 * ns, :]
            arow = 2 * ns
        else:
            arow = ns
        if nd > 0:
            A[nt * ns:, c * ns:(c + 1) * ns] = P[arow:arow + nd, :]
        assert rows == arow + nd
    if nd > 0:
        s = (nt - 1) * ns * (2 * ns + nd) + ns * (ns + nd)
        Cn = mf[s:s + nd * nd].reshape(nd, nd).T
        A[nt * ns:, nt * ns:] = np.tril(Cn)
    if symmetric:
        A = np.tril(A) + np.tril(A, -1).T
    return A


def dense_to_mf(A: np.ndarray, ns: int, nt: int, nd: int) -> np.ndarray:
    """Inverse of ``mf_to_dense``: lower logical blocks of A into a fresh MF array."""
    mf = np.zeros(mf_size(ns, nt, nd))
    for c in range(nt):
        lda = _lda(ns, nt, nd, c)
        s = _snode(ns, nt, nd, c)
        P = np.zeros((lda, ns))
        P[:ns, :] = np.tril(A[c * ns:(c + 1) * ns, c * ns:(c + 1) * ns])
        if c < nt - 1:
            P[ns:2 * ns, :] = A[(c + 1) * ns:(c + 2) * ns, c * ns:(c + 1) * ns]
            arow = 2 * ns
        else:
            arow = ns
        if nd > 0:
            P[arow:arow + nd, :] = A[nt * ns:, c * ns:(c + 1) * ns]
        mf[s:s + lda * ns] = P.T.reshape(-1)
    if nd > 0:
        s = (nt - 1) * ns * (2 * ns + nd) + ns * (ns + nd)
        mf[s:s + nd * nd] = np.tril(A[nt * ns:, nt * ns:]).T.reshape(-1)
    return mf


def dense_to_csc_lower(A: np.ndarray, ns: int, nt: int, nd: int, keep_zeros: bool = False):
    """Lower-triangle CSC (uint64 indices) of the block-arrow pattern of A."""
    n = A.shape[0]
    ia = [0]
    ja = []
    a = []
    for c in range(n):
        bc = c // ns
        for r in range(c, n):
            br = r // ns
            inpat = c >= ns * nt or r >= ns * nt or br == bc or br == bc + 1
            if not inpat:
                continue
            v = A[r, c]
            if v != 0.0 or keep_zeros:
                ja.append(r)
                a.append(v)
        ia.append(len(ja))
    return (np.array(ia, dtype=np.uint64), np.array(ja, dtype=np.uint64),
            np.array(a, dtype=np.float64))


def csc_to_dense(ia, ja, a, n: int) -> np.ndarray:
    A = np.zeros((n, n))
    for c in range(n):
        for k in range(int(ia[c]), int(ia[c + 1])):
            A[int(ja[k]), c] = a[k]
    return np.tril(A) + np.tril(A, -1).T


def read_dat(path: str):
    """main.C:50-70: ``n n nnz`` then ia[n+1], ja[nnz], a[nnz]."""
    tok = open(path).read().split()
    n, nnz = int(tok[0]), int(tok[2])
    p = 3
    ia = np.array(tok[p:p + n + 1], dtype=np.uint64); p += n + 1
    ja = np.array(tok[p:p + nnz], dtype=np.uint64); p += nnz
    a = np.array(tok[p:p + nnz], dtype=np.float64)
    return n, ia, ja, a


def write_dat(path: str, n: int, ia, ja, a) -> None:
    with open(path, "w") as f:
        f.write(f"{n}\n{n}\n{len(a)}\n")
        f.write("\n".join(str(int(v)) for v in ia) + "\n")
        f.write("\n".join(str(int(v)) for v in ja) + "\n")
        f.write("\n".join(repr(float(v)) for v in a) + "\n")


def closed_form_csc(ns: int, nt: int, nd: int):
    """Formula behind data/A_9_9_ns3_nt3.dat and data/A_11_11_ns3_nt3_nd2.dat:
    A[c,c] = 10 (c+1), A[r,c] = -0.1 (r-c) for every lower entry of the pattern."""
    n = ns * nt + nd
    A = np.zeros((n, n))
    for c in range(n):
        bc = c // ns
        for r in range(c, n):
            br = r // ns
            if c >= ns * nt or r >= ns * nt or br == bc or br == bc + 1:
                A[r, c] = 10.0 * (c + 1) if r == c else -(r - c) / 10.0
    return dense_to_csc_lower(A, ns, nt, nd)


# ---------------------------------------------------------------- oracle calls
def csc_to_mf(ia, ja, a, ns, nt, nd) -> np.ndarray:
    mf = np.zeros(mf_size(ns, nt, nd))
    rc = load().rgf_oracle_csc_to_mf(np.ascontiguousarray(ia, dtype=np.uint64),
                                     np.ascontiguousarray(ja, dtype=np.uint64),
                                     np.ascontiguousarray(a, dtype=np.float64), ns, nt, nd, mf)
    if rc != 0:
        raise ValueError(f"CSC entry {-rc - 1} is outside the block-arrow pattern")
    return mf


def factor(mf_a: np.ndarray, ns, nt, nd) -> np.ndarray:
    """FirstStageFactor on a copy of A in MF layout; returns the factor MF."""
    mf = np.array(mf_a, dtype=np.float64, copy=True)
    info = load().rgf_oracle_factor(mf, ns, nt, nd)
    if info != 0:
        raise np.linalg.LinAlgError(f"not SPD at column {info}")
    return mf


def logdet(mf_l: np.ndarray, ns, nt, nd) -> float:
    return float(load().rgf_oracle_logdet(np.ascontiguousarray(mf_l), ns, nt, nd))


def solve(mf_l: np.ndarray, ns, nt, nd, b: np.ndarray) -> np.ndarray:
    """b: n x nrhs (or n); returns x with the same shape."""
    n = ns * nt + nd
    b2 = b.reshape(n, -1)
    nrhs = b2.shape[1]
    rhs = np.ascontiguousarray(b2.T).reshape(-1).copy()   # column-major, ld = n
    load().rgf_oracle_solve(np.ascontiguousarray(mf_l), ns, nt, nd, rhs, nrhs)
    return rhs.reshape(nrhs, n).T.reshape(b.shape)


def selinv(mf_l: np.ndarray, ns, nt, nd, exact: bool = True) -> np.ndarray:
    d = np.zeros(ns * nt + nd)
    rc = load().rgf_oracle_selinv(np.ascontiguousarray(mf_l), ns, nt, nd, d, 1 if exact else 0)
    if rc != 0:
        raise MemoryError("oracle selinv allocation failed")
    return d


def residual_norm(ia, ja, a, n, x, b, normalized=False) -> float:
    f = load().rgf_oracle_residual_norm_normalized if normalized else load().rgf_oracle_residual_norm
    return float(f(np.ascontiguousarray(ia, dtype=np.uint64), np.ascontiguousarray(ja, dtype=np.uint64),
                   np.ascontiguousarray(a, dtype=np.float64), n,
                   np.ascontiguousarray(x, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64)))


def philox4x32_10(ctr, key) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint32)
    load().rgf_oracle_philox4x32_10(np.asarray(ctr, dtype=np.uint32), np.asarray(key, dtype=np.uint32), out)
    return out


def gen_normal(seed, matrix_id, kind, block, e) -> float:
    return float(load().rgf_oracle_gen_normal(seed, matrix_id, kind, block, e))


def gen_uniform(seed, matrix_id, idx) -> float:
    return float(load().rgf_oracle_gen_uniform(seed, matrix_id, idx))


def generate(seed, ns, nt, nd, matrix_id=0, sigma_c=0.1, sigma_a=0.05, theta_d=1.0, theta_c=1.0):
    """Host generator for the benchmark matrices; returns A in MF layout."""
    mf = np.zeros(mf_size(ns, nt, nd))
    rc = load().rgf_oracle_generate(seed, matrix_id, ns, nt, nd, sigma_c, sigma_a, theta_d, theta_c, mf)
    if rc != 0:
        raise MemoryError("oracle generator allocation failed")
    return mf


# ---------------------------------------------------------------- dense pins
def reference_selinv_dense(A: np.ndarray, ns: int, nt: int, nd: int) -> np.ndarray:
    """RGF.H:422-506's recursion, evaluated with dense numpy blocks of
    L = cholesky(A):  S_dd = X_dd^T X_dd,  S_ii = X_i^T [I + B_i^T S_{i+1} B_i +
    D_i^T S_dd D_i] X_i  (no cross terms).  Returns the diagonal."""
    Lc = np.linalg.cholesky(A)
    n = A.shape[0]
    out = np.zeros(n)
    Sdd = None
    if nd > 0:
        Xdd = np.linalg.inv(Lc[nt * ns:, nt * ns:])
        Sdd = Xdd.T @ Xdd
        out[nt * ns:] = np.diag(Sdd)
    Snext = None
    for i in range(nt - 1, -1, -1):
        sl = slice(i * ns, (i + 1) * ns)
        X = np.linalg.inv(Lc[sl, sl])
        T = np.eye(ns)
        if i < nt - 1:
            B = Lc[(i + 1) * ns:(i + 2) * ns, sl]
            T = T + B.T @ Snext @ B
        if nd > 0:
            D = Lc[nt * ns:, sl]
            T = T + D.T @ Sdd @ D
        Snext = X.T @ T @ X
        out[sl] = np.diag(Snext)
    return out
