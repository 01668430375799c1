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


# ---------------------------------------------------------------- complex (RGF<CPX>)
# The reference's complex instantiation (RGF.H:756-780) runs the same sweep with block
# updates that transpose without conjugating (trsm 'R','L','T','N' at RGF.H:260, 287, 317;
# gemm 'N','T' at RGF.H:276-308; selinv gemm 'T','N' at RGF.H:442-501): complex-symmetric
# algebra, A = A^T = L L^T.  Restated here with dense numpy blocks (test infrastructure).
def zchol_sym(S: np.ndarray) -> np.ndarray:
    """Complex-symmetric Cholesky L L^T = S (lower part of S read), principal square root
    of each pivot (the diagonal-block potrf of RGF.H:254, 284, 312, 335 in that algebra)."""
    S = np.tril(np.array(S, dtype=np.complex128, copy=True))
    n = S.shape[0]
    L = np.zeros_like(S)
    for j in range(n):
        d = S[j, j]
        if d == 0 or not np.isfinite(d):
            raise np.linalg.LinAlgError(f"zero pivot at column {j + 1}")
        s = np.sqrt(d)
        L[j, j] = s
        L[j + 1:, j] = S[j + 1:, j] / s
        S[j + 1:, j + 1:] -= np.tril(np.outer(L[j + 1:, j], L[j + 1:, j]))
    return L


def zfactor_blocks(A: np.ndarray, ns: int, nt: int, nd: int) -> dict:
    """FirstStageFactor (RGF.H:233-340) in complex-symmetric algebra on dense blocks of A:
    S_i = A_ii - B_{i-1} B_{i-1}^T, L_ii = chol(S_i), B_i = A_{i+1,i} L_ii^-T,
    D_i = (A_di - D_{i-1} B_{i-1}^T) L_ii^-T, L_dd = chol(A_dd - sum D_i D_i^T)."""
    A = np.asarray(A, dtype=np.complex128)
    blk = lambda r, c: A[r * ns:(r + 1) * ns, c * ns:(c + 1) * ns]
    Lii, Bs, Ds = [], [], []
    Sdd = A[nt * ns:, nt * ns:].copy() if nd > 0 else None
    for i in range(nt):
        S = blk(i, i).copy()
        Ad = A[nt * ns:, i * ns:(i + 1) * ns].copy() if nd > 0 else None
        if i > 0:
            S -= Bs[i - 1] @ Bs[i - 1].T
            if nd > 0:
                Ad -= Ds[i - 1] @ Bs[i - 1].T
        L = zchol_sym(S)
        X = np.linalg.inv(L)
        Lii.append(L)
        if i < nt - 1:
            Bs.append(blk(i + 1, i) @ X.T)
        if nd > 0:
            D = Ad @ X.T
            Ds.append(D)
            Sdd -= D @ D.T
    Ldd = zchol_sym(Sdd) if nd > 0 else None
    return {"ns": ns, "nt": nt, "nd": nd, "L": Lii, "B": Bs, "D": Ds, "Ldd": Ldd}


def zblocks_to_dense(F: dict) -> np.ndarray:
    """The dense lower-triangular factor L of zfactor_blocks (for L L^T = A)."""
    ns, nt, nd = F["ns"], F["nt"], F["nd"]
    n = ns * nt + nd
    L = np.zeros((n, n), dtype=np.complex128)
    for i in range(nt):
        L[i * ns:(i + 1) * ns, i * ns:(i + 1) * ns] = F["L"][i]
        if i < nt - 1:
            L[(i + 1) * ns:(i + 2) * ns, i * ns:(i + 1) * ns] = F["B"][i]
        if nd > 0:
            L[nt * ns:, i * ns:(i + 1) * ns] = F["D"][i]
    if nd > 0:
        L[nt * ns:, nt * ns:] = F["Ldd"]
    return L


def zlogdet(F: dict) -> complex:
    """logDet (RGF.H:616-627, Blas.H:1084-1092): 2 sum log L_jj, principal logarithms."""
    s = sum(np.log(np.diag(L)).sum() for L in F["L"])
    if F["nd"] > 0:
        s += np.log(np.diag(F["Ldd"])).sum()
    return complex(2.0 * s)


def zsolve(F: dict, b: np.ndarray) -> np.ndarray:
    """SecondStageSolve (RGF.H:344-418): forward y_i = L_ii^-1 (b_i - B_{i-1} y_{i-1}),
    y_d = L_dd^-1 (b_d - sum D_i y_i); backward x_d = L_dd^-T y_d,
    x_i = L_ii^-T (y_i - B_i^T x_{i+1} - D_i^T x_d)."""
    ns, nt, nd = F["ns"], F["nt"], F["nd"]
    b = np.asarray(b, dtype=np.complex128)
    shape = b.shape
    b = b.reshape(ns * nt + nd, -1)
    y = np.zeros_like(b)
    x = np.zeros_like(b)
    for i in range(nt):
        r = b[i * ns:(i + 1) * ns].copy()
        if i > 0:
            r -= F["B"][i - 1] @ y[(i - 1) * ns:i * ns]
        y[i * ns:(i + 1) * ns] = np.linalg.solve(F["L"][i], r)
    if nd > 0:
        r = b[nt * ns:].copy()
        for i in range(nt):
            r -= F["D"][i] @ y[i * ns:(i + 1) * ns]
        y[nt * ns:] = np.linalg.solve(F["Ldd"], r)
        x[nt * ns:] = np.linalg.solve(F["Ldd"].T, y[nt * ns:])
    for i in range(nt - 1, -1, -1):
        r = y[i * ns:(i + 1) * ns].copy()
        if i < nt - 1:
            r -= F["B"][i].T @ x[(i + 1) * ns:(i + 2) * ns]
        if nd > 0:
            r -= F["D"][i].T @ x[nt * ns:]
        x[i * ns:(i + 1) * ns] = np.linalg.solve(F["L"][i].T, r)
    return x.reshape(shape)


def zselinv(F: dict, exact: bool = True) -> np.ndarray:
    """diag(A^-1) by the Takahashi recursion (ThirdStageRGF, RGF.H:422-506) with X_i = L_ii^-1:
    S_dd = X_dd^T X_dd; S_ii = X_i^T [I + B_i^T S_{i+1} B_i + D_i^T S_dd D_i] X_i, plus with
    `exact` the arrow cross terms and S_{d,i} = -(S_{d,i+1} G_i + S_dd H_i) that the
    reference's formula leaves out (exact for every nd)."""
    ns, nt, nd = F["ns"], F["nt"], F["nd"]
    out = np.zeros(ns * nt + nd, dtype=np.complex128)
    Sdd = Sd_next = S_next = None
    if nd > 0:
        Xdd = np.linalg.inv(F["Ldd"])
        Sdd = Xdd.T @ Xdd
        out[nt * ns:] = np.diag(Sdd)
    for i in range(nt - 1, -1, -1):
        X = np.linalg.inv(F["L"][i])
        S = X.T @ X
        Sd = None
        if nd > 0:
            H = F["D"][i] @ X
            S = S + H.T @ Sdd @ H
            Sd = -(Sdd @ H)
        if i < nt - 1:
            G = F["B"][i] @ X
            S = S + G.T @ S_next @ G
            if nd > 0 and exact:
                S = S + G.T @ Sd_next.T @ H + H.T @ Sd_next @ G
                Sd = Sd - Sd_next @ G
        out[i * ns:(i + 1) * ns] = np.diag(S)
        S_next, Sd_next = S, Sd
    return out
