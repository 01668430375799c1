"""Parity of the HIP path (through the C ABI) with the CPU oracle and the goldens.

Bar: FP64 logdet to rel 1e-10 (tests use 1e-11 at small sizes), solves to 1e-9,
diag(A^-1) to 1e-9 (see each assert)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import rgf_amd
from rgf_amd import _lib
from oracle import oracle
import fixtures

pytestmark = pytest.mark.gpu
G = fixtures.goldens()
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def dptr(a):
    import ctypes as C
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ------------------------------------------------------------------ kernels
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(16, 16, 4), (64, 64, 16), (130, 70, 33), (5, 200, 129), (300, 270, 40)])
@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4])
def test_gemm_kernel(gpu, ta, tb, m, n, k, tile):
    """FP64 product against numpy, for the automatic tile shape (0) and each forced
    one (1 = 64x64, 2 = 128x64, 3 = 128x128, 4 = 32x32; bits 12-14 of the debug flags)."""
    say(f"gemm tile={tile} ta={ta} tb={tb} m={m} n={n} k={k}")
    rng = np.random.default_rng(m * 7 + n + k)
    lda, ldb, ldc = (k if ta else m) + 3, (n if tb else k) + 1, m + 2
    A = np.asfortranarray(rng.standard_normal((lda, m if ta else k)))
    B = np.asfortranarray(rng.standard_normal((ldb, k if tb else n)))
    C = np.asfortranarray(rng.standard_normal((ldc, n)))
    opA = (A[:k, :m].T if ta else A[:m, :k])
    opB = (B[:n, :k].T if tb else B[:k, :n])
    ref = C.copy()
    ref[:m, :n] = 1.5 * opA @ opB - 0.5 * C[:m, :n]
    Cf = C.reshape(-1, order="F").copy()
    rc = _lib.load().rgf_debug_gemm(ta, tb, m, n, k, 1.5, dptr(A.reshape(-1, order="F").copy()), lda,
                                    dptr(B.reshape(-1, order="F").copy()), ldb, -0.5, dptr(Cf), ldc, tile << 12)
    assert rc == 0
    got = Cf.reshape(ldc, n, order="F")
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()) * k


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("tile", [1, 2, 3, 4])
def test_gemm_fast_loads_match_masked_loads(gpu, ta, tb, tile):
    """The FP64 kernel's full-k-step loads (loop-invariant lane offsets, edge rows clamped
    to the last valid row) give the same bits as per-element masked loads, on ragged
    shapes whose padding rows hold NaN (a read of them would reach the output)."""
    say(f"gemm fast loads tile={tile} ta={ta} tb={tb}")
    lib = _lib.load()
    rng = np.random.default_rng(11 + 4 * ta + 2 * tb + tile)
    m, n, k = 197, 131, 75
    lda, ldb, ldc = (k if ta else m) + 5, (n if tb else k) + 3, m
    A = np.full((lda, m if ta else k), np.nan, order="F")
    B = np.full((ldb, k if tb else n), np.nan, order="F")
    if ta:
        A[:k, :m] = rng.standard_normal((k, m))
    else:
        A[:m, :k] = rng.standard_normal((m, k))
    if tb:
        B[:n, :k] = rng.standard_normal((n, k))
    else:
        B[:k, :n] = rng.standard_normal((k, n))
    ref = (A[:k, :m].T if ta else A[:m, :k]) @ (B[:n, :k].T if tb else B[:k, :n])
    outs = []
    for fast in (1, 0):
        old = lib.rgf_debug_set_tuning(4, fast)
        try:
            Cf = np.zeros(ldc * n)
            rc = lib.rgf_debug_gemm(ta, tb, m, n, k, 1.0, dptr(A.reshape(-1, order="F").copy()), lda,
                                    dptr(B.reshape(-1, order="F").copy()), ldb, 0.0, dptr(Cf), ldc, tile << 12)
        finally:
            lib.rgf_debug_set_tuning(4, old)
        assert rc == 0
        outs.append(Cf.reshape(ldc, n, order="F"))
    assert np.isfinite(outs[0]).all()
    assert np.array_equal(outs[0], outs[1])
    assert np.abs(outs[0] - ref).max() <= 1e-13 * np.abs(ref).max() * k


@pytest.mark.parametrize("flags", [1, 2, 8, 16, 1 | 4 | 8])
@pytest.mark.parametrize("tile", [0, 2, 3, 4])
@pytest.mark.parametrize("size", [150, 100, 333])
def test_gemm_kernel_triangular_flags(gpu, flags, tile, size):
    """Triangle flags (k-range skipping, lower-stored C) against numpy, at sizes that are
    and are not a multiple of the 4-wide MFMA k step, up to 6 x 6 tiles (every lower tile
    of a round-robin-dealt launch written once: the others would keep the fill 7.0)."""
    say(f"gemm tile={tile} flags={flags} size={size}")
    rng = np.random.default_rng(flags)
    m = n = k = size
    L = np.tril(rng.standard_normal((k, k)))
    G0 = rng.standard_normal((k, k))
    # choose operands so that the flagged operand really is triangular
    if flags == 2:     # op(A) lower (N): A = L
        ta, tb, A, B = 0, 0, L, G0
    elif flags == 8:   # op(B) lower (N): B = L
        ta, tb, A, B = 0, 0, G0, L
    elif flags == 16:  # op(B) upper: op(B) = L^T with tb = 1
        ta, tb, A, B = 0, 1, G0, L
    elif flags == 1 | 4 | 8:   # L^T L, lower stores
        ta, tb, A, B = 1, 0, L, L
    else:              # lower stores only
        ta, tb, A, B = 0, 1, G0, G0
    opA = A.T if ta else A
    opB = B.T if tb else B
    C0 = np.full((m, n), 7.0)
    ref = opA @ opB
    if flags & 1:
        ref = np.where(np.tril(np.ones((m, n))) > 0, ref, 7.0)
    Cf = C0.reshape(-1, order="F").copy()
    rc = _lib.load().rgf_debug_gemm(ta, tb, m, n, k, 1.0, dptr(A.reshape(-1, order="F").copy()), k,
                                    dptr(B.reshape(-1, order="F").copy()), k, 0.0, dptr(Cf), m, flags | (tile << 12))
    assert rc == 0
    assert np.abs(Cf.reshape(m, n, order="F") - ref).max() < 1e-11


@pytest.mark.parametrize("ta,tb,flags", [(0, 0, 0), (0, 1, 1), (1, 0, 1 | 4 | 8), (0, 0, 8), (0, 1, 16), (0, 0, 2),
                                         (1, 1, 0)])
@pytest.mark.parametrize("m,n,k", [(333, 333, 333), (200, 130, 97), (64, 64, 128)])
def test_gemm_tile_shapes_same_bits(gpu, ta, tb, flags, m, n, k):
    """Every FP64 tile shape sums each output over the same k-steps in the same order (a
    tile only skips k-steps of exact zeros of a triangular operand), so the automatic
    choice of shape never changes a bit: all shapes give bitwise the same C."""
    say(f"gemm tile shapes same bits ta={ta} tb={tb} flags={flags} {m}x{n}x{k}")
    if flags & (1 | 2 | 4 | 8 | 16) and flags != 0:
        m = n = k = max(m, n, k)
    rng = np.random.default_rng(m + n + k + flags)
    lda, ldb, ldc = (k if ta else m), (n if tb else k), m
    A = rng.standard_normal((lda, m if ta else k))
    B = rng.standard_normal((ldb, k if tb else n))
    # triangular operands hold explicit zeros in their other triangle
    if flags & 2:
        A = np.tril(A) if not ta else np.triu(A)
    if flags & 4:
        A = np.triu(A) if not ta else np.tril(A)
    if flags & 8:
        B = np.tril(B) if not tb else np.triu(B)
    if flags & 16:
        B = np.triu(B) if not tb else np.tril(B)
    C0 = rng.standard_normal(ldc * n)
    outs = []
    for tile in (1, 2, 3, 4, 0):
        Cf = C0.copy()
        rc = _lib.load().rgf_debug_gemm(ta, tb, m, n, k, 0.75, dptr(A.reshape(-1, order="F").copy()), lda,
                                        dptr(B.reshape(-1, order="F").copy()), ldb, 0.5, dptr(Cf), ldc,
                                        flags | (tile << 12))
        assert rc == 0
        outs.append(Cf)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


@pytest.mark.parametrize("n", [1, 5, 64, 65, 150, 200, 300, 333, 1000])
def test_potrf_trtri_kernels(gpu, n):
    say(f"potrf/trtri n={n}")
    rng = np.random.default_rng(n)
    Gm = rng.standard_normal((n, n))
    A = Gm @ Gm.T / n + np.eye(n)
    lda = n + 5
    buf = np.zeros((lda, n), order="F")
    buf[:n, :n] = A
    flat = buf.reshape(-1, order="F").copy()
    import ctypes as C
    ld = C.c_double()
    assert _lib.load().rgf_debug_potrf(n, dptr(flat), lda, 0, C.byref(ld)) == 0
    L = np.tril(flat.reshape(lda, n, order="F")[:n])
    Lc = np.linalg.cholesky(A)
    assert np.abs(L - Lc).max() < 1e-12 * np.abs(Lc).max() * n
    assert ld.value == pytest.approx(np.linalg.slogdet(A)[1], rel=1e-12, abs=1e-12)
    flat = buf.reshape(-1, order="F").copy()
    assert _lib.load().rgf_debug_potrf(n, dptr(flat), lda, 1, C.byref(ld)) == 0
    X = flat.reshape(lda, n, order="F")[:n]
    assert np.abs(np.triu(X, 1)).max() == 0.0   # explicit zeros above the diagonal
    Xr = np.linalg.inv(Lc)
    assert np.abs(X - Xr).max() < 1e-11 * np.abs(Xr).max()


@pytest.mark.parametrize("bad", [70, 0, 20, 63, 99, 200, 311])
@pytest.mark.parametrize("invert", [0, 1])
def test_potrf_not_spd(gpu, bad, invert):
    """The first non-positive pivot is reported 1-based, wherever it falls in a tile (invert=1
    factors through the one-launch task graph, invert=0 through the tile launches)."""
    say(f"potrf not SPD at {bad} invert={invert}")
    n = 100 if bad < 100 else 330
    A = np.eye(n)
    A[bad, bad] = -1.0
    A[min(bad + 3, n - 1), min(bad + 3, n - 1)] = 0.0   # a later bad pivot is not reported
    flat = A.reshape(-1, order="F").copy()
    import ctypes as C
    ld = C.c_double()
    assert _lib.load().rgf_debug_potrf(n, dptr(flat), n, invert, C.byref(ld)) == rgf_amd.RGF_ENOTSPD
    assert f"column {bad + 1}".encode() in _lib.load().rgf_last_error(None)


@pytest.mark.parametrize("n", [130, 333, 1000, 2865])
def test_potrf_task_graph_matches_tile_launches(gpu, n):
    """L^-1 through the one-launch task-graph Cholesky against the tile-launch Cholesky: the
    same factorization to rounding (the panel updates are grouped differently)."""
    say(f"potrf task graph vs tile launches n={n}")
    rng = np.random.default_rng(7 + n)
    Gm = rng.standard_normal((n, n))
    A = Gm @ Gm.T / n + np.eye(n)
    import ctypes as C
    lib = _lib.load()
    outs, lds = [], []
    for dag in (1, 0):
        old = lib.rgf_debug_set_tuning(16, dag)
        try:
            flat = A.reshape(-1, order="F").copy()
            ld = C.c_double()
            assert lib.rgf_debug_potrf(n, dptr(flat), n, 1, C.byref(ld)) == 0
        finally:
            lib.rgf_debug_set_tuning(16, old)
        outs.append(flat.reshape(n, n, order="F"))
        lds.append(ld.value)
    X0, X1 = outs
    assert np.abs(np.triu(X0, 1)).max() == 0.0
    assert np.abs(X0 - X1).max() < 1e-12 * np.abs(X1).max()
    assert lds[0] == pytest.approx(lds[1], rel=1e-13)
    assert lds[0] == pytest.approx(np.linalg.slogdet(A)[1], rel=1e-11)


# ------------------------------------------------------------------ full path vs oracle
SHAPES = [(42, 3, 0), (8, 5, 3), (65, 4, 5), (130, 3, 8), (64, 1, 4), (100, 1, 0), (1, 5, 1), (3, 3, 2),
          (200, 6, 6), (129, 2, 0)]


@pytest.mark.parametrize("ns,nt,nd", SHAPES)
def test_generated_vs_oracle(gpu, ns, nt, nd):
    say(f"generated ns={ns} nt={nt} nd={nd}")
    with rgf_amd.RGF.generated(ns, nt, nd, seed=1234 + ns) as s:
        A_mf = s.export_matrix_mf()
        # device generator vs host restatement of the same definition
        H_mf = oracle.generate(1234 + ns, ns, nt, nd)
        assert np.abs(A_mf - H_mf).max() <= 1e-13 * np.abs(H_mf).max()
        L = oracle.factor(A_mf, ns, nt, nd)
        ld = s.logDet()
        ld0 = oracle.logdet(L, ns, nt, nd)
        assert abs(ld - ld0) <= 1e-11 * abs(ld0)
        # factor blocks: diagonal blocks hold L_ii^-1
        F = s.export_factor_mf()
        Ld = oracle.mf_to_dense(L, ns, nt, nd, symmetric=False)
        Fd = oracle.mf_to_dense(F, ns, nt, nd, symmetric=False)
        n = ns * nt + nd
        for t in range(nt):
            sl = slice(t * ns, (t + 1) * ns)
            Xr = np.linalg.inv(Ld[sl, sl])
            assert np.abs(Fd[sl, sl] - Xr).max() <= 1e-11 * np.abs(Xr).max()
        off = Ld.copy()
        offF = Fd.copy()
        for t in range(nt):
            sl = slice(t * ns, (t + 1) * ns)
            off[sl, sl] = 0
            offF[sl, sl] = 0
        off[nt * ns:, nt * ns:] = 0
        offF[nt * ns:, nt * ns:] = 0
        assert np.abs(offF - off).max() <= 1e-11 * max(1.0, np.abs(off).max())
        for nrhs in (1, 3, 16, 17):
            b = np.random.default_rng(nrhs).standard_normal((n, nrhs))
            x = s.solve(None, b, nrhs)
            x0 = oracle.solve(L, ns, nt, nd, b)
            assert np.abs(x - x0).max() <= 1e-10 * np.abs(x0).max(), nrhs
        d = s.RGFdiag()
        d0 = oracle.selinv(L, ns, nt, nd, exact=True)
        assert np.abs(d / d0 - 1).max() <= 1e-10
        # factor still valid after RGFdiag (solve again)
        b = np.ones((n, 1))
        assert np.abs(s.solve(None, b, 1) - oracle.solve(L, ns, nt, nd, b)).max() <= 1e-10 * np.abs(
            oracle.solve(L, ns, nt, nd, b)).max()


@pytest.mark.parametrize("ns,nt,nd", [(8, 5, 3), (65, 4, 5), (3, 3, 2)])
def test_reference_compat_selinv(gpu, ns, nt, nd):
    say(f"compat selinv ns={ns} nt={nt} nd={nd}")
    opts = rgf_amd.default_options(selinv_mode=rgf_amd.RGF_SELINV_REFERENCE)
    with rgf_amd.RGF.generated(ns, nt, nd, seed=7, options=opts) as s:
        A_mf = s.export_matrix_mf()
        d = s.RGFdiag()
    L = oracle.factor(A_mf, ns, nt, nd)
    d0 = oracle.selinv(L, ns, nt, nd, exact=False)
    assert np.abs(d / d0 - 1).max() <= 1e-10


def _csc_case(name):
    if name.startswith("closed"):
        ns, nt, nd = (3, 3, 0) if name == "closed_9x9" else (3, 3, 2)
        ia, ja, a = oracle.closed_form_csc(ns, nt, nd)
    elif name == "inla_126":   # the reference's own INLA matrix (data/A_126_126_ns42_nt3.dat)
        ia, ja, a, ns, nt, nd = fixtures.inla_126()
    else:
        A, ns, nt, nd = fixtures.spde_like(nd=G[name]["nd"])
        ia, ja, a = oracle.dense_to_csc_lower(A, ns, nt, nd)
    return ia, ja, a, ns, nt, nd


@pytest.mark.parametrize("name", ["closed_9x9", "closed_11x11", "inla_126", "spde_126", "spde_126_nd2"])
def test_csc_fixture_goldens(gpu, name):
    say(f"csc fixture {name}")
    ia, ja, a, ns, nt, nd = _csc_case(name)
    g = G[name]
    n = ns * nt + nd
    with rgf_amd.RGF(ia, ja, a, ns, nt, nd) as s:
        s.factorize()
        assert s.logDet() == pytest.approx(g["logdet"], rel=1e-11)
        b = fixtures.rhs_formula(n, 2)
        x = s.solve(None, b, 2)
        xg = np.array(g["x_rhs_formula_2"]).reshape(n, 2, order="F")
        assert np.abs(x - xg).max() <= 1e-9 * np.abs(xg).max()
        d = s.RGFdiag()
        assert np.abs(d / np.array(g["diag_inv"]) - 1).max() <= 1e-9
        r = s.residualNormNormalized(x[:, 0].copy(), b[:, 0].copy())
        assert r < 1e-12
        # at a perturbed x the residual is O(1e-3) of b: both norms against numpy to 1e-12
        A = oracle.csc_to_dense(ia, ja, a, n)
        xp = x[:, 0] + 1e-3 * np.abs(xg[:, 0]).max() * np.sin(np.arange(n) + 0.3)
        rp = np.linalg.norm(b[:, 0] - A @ xp)
        assert rp > 1e-4 * np.linalg.norm(b[:, 0])
        assert s.residualNorm(xp.copy(), b[:, 0].copy()) == pytest.approx(rp, rel=1e-12)
        assert s.residualNormNormalized(xp.copy(), b[:, 0].copy()) == pytest.approx(rp / np.linalg.norm(b[:, 0]),
                                                                                    rel=1e-12)
    opts = rgf_amd.default_options(selinv_mode=rgf_amd.RGF_SELINV_REFERENCE)
    with rgf_amd.RGF(ia, ja, a, ns, nt, nd, opts) as s:
        d = s.RGFdiag()
        assert np.abs(d / np.array(g["diag_reference"]) - 1).max() <= 1e-9


def test_in_place_mode(gpu):
    say("keep_matrix=0")
    opts = rgf_amd.default_options(keep_matrix=0)
    ns, nt, nd = 70, 4, 3
    with rgf_amd.RGF.generated(ns, nt, nd, seed=3, options=opts) as s, \
            rgf_amd.RGF.generated(ns, nt, nd, seed=3) as k:
        assert s.logDet() == pytest.approx(k.logDet(), rel=1e-14)
        n = ns * nt + nd
        b = np.ones((n, 1))
        assert np.abs(s.solve(None, b, 1) - k.solve(None, b, 1)).max() < 1e-13
        assert np.abs(s.RGFdiag() - k.RGFdiag()).max() < 1e-13
        with pytest.raises(rgf_amd.RGFError) as e:
            s.residualNorm(np.ones(n), np.ones(n))
        assert e.value.code == rgf_amd.RGF_ESTATE


def test_not_spd_reports_column(gpu):
    say("not SPD through the API")
    ns, nt, nd = 3, 3, 0
    ia, ja, a = oracle.closed_form_csc(ns, nt, nd)
    a = a.copy()
    # make A[4,4] tiny so the Schur complement of block 1 goes negative at its second pivot
    idx = [k for c in range(9) for k in range(int(ia[c]), int(ia[c + 1])) if c == 4 and ja[k] == 4][0]
    a[idx] = -50.0
    with rgf_amd.RGF(ia, ja, a, ns, nt, nd) as s:
        with pytest.raises(rgf_amd.RGFError) as e:
            s.factorize()
        assert e.value.code == rgf_amd.RGF_ENOTSPD
        assert "column 5" in str(e.value)


def test_device_entry_points_and_determinism(gpu):
    say("device entry points")
    ns, nt, nd = 150, 5, 4
    n = ns * nt + nd
    with rgf_amd.RGF.generated(ns, nt, nd, seed=77) as s:
        b = np.random.default_rng(1).standard_normal((n, 5))
        bd = rgf_amd.DeviceBuffer(n * 5)
        xd = rgf_amd.DeviceBuffer(n * 5)
        bd.upload(b.reshape(-1, order="F"))
        s.solve_device(xd, bd, 5)
        s.synchronize()
        x1 = xd.download().reshape(n, 5, order="F")
        assert np.array_equal(x1, s.solve(None, b, 5))
        dd = rgf_amd.DeviceBuffer(n)
        s.selinv_diag_device(dd)
        s.synchronize()
        d1 = dd.download()
        s.factorize()
        assert np.array_equal(d1, s.RGFdiag())   # bitwise reproducible
        info = s.info()
        F, S = rgf_amd.canonical_flops(ns, nt, nd)
        assert info.flops_factor == pytest.approx(F) and info.flops_selinv == pytest.approx(S)


def test_larger_block_properties(gpu):
    """ns = 1000 (the INLA-like single-matrix block size) on 5 blocks: parity with the
    oracle on the same matrix and a small residual."""
    ns, nt, nd = 1000, 5, 4
    say(f"larger ns={ns} nt={nt} nd={nd}")
    with rgf_amd.RGF.generated(ns, nt, nd, seed=2024) as s:
        A_mf = s.export_matrix_mf()
        n = s.n
        b = np.random.default_rng(5).standard_normal(n)
        x = s.solve(None, b, 1)
        assert s.residualNormNormalized(x, b) < 1e-14
        ld = s.logDet()
        d = s.RGFdiag()
    L = oracle.factor(A_mf, ns, nt, nd)
    assert abs(ld - oracle.logdet(L, ns, nt, nd)) <= 1e-12 * abs(ld)
    assert np.abs(d / oracle.selinv(L, ns, nt, nd, exact=True) - 1).max() <= 1e-11


def test_c1_full_size_vs_oracle(gpu):
    """The INLA-like single-matrix workload at its full size (ns 1000, nt 50, nd 4,
    n = 50 004): logdet, a solve and diag(A^-1) against the oracle on the same matrix."""
    ns, nt, nd = 1000, 50, 4
    say(f"c1 full size ns={ns} nt={nt} nd={nd}")
    with rgf_amd.RGF.generated(ns, nt, nd, seed=2024) as s:
        A_mf = s.export_matrix_mf()
        b = np.random.default_rng(9).standard_normal(s.n)
        x = s.solve(None, b, 1)
        ld = s.logDet()
        d = s.RGFdiag()
    L = oracle.factor(A_mf, ns, nt, nd)
    assert abs(ld - oracle.logdet(L, ns, nt, nd)) <= 1e-11 * abs(ld)
    x0 = oracle.solve(L, ns, nt, nd, b)
    assert np.abs(x - x0).max() <= 1e-10 * np.abs(x0).max()
    assert np.abs(d / oracle.selinv(L, ns, nt, nd, exact=True) - 1).max() <= 1e-10


def test_c2_full_size_properties(gpu):
    """The c2 workload at full size (ns 2865, nt 100, nd 6, n = 286 506), where a dense
    check is out of reach, through properties that do not depend on the size: a solve's
    residual against the resident matrix; diag(A^-1)_j equals x_j of A x = e_j (and the
    selected-inverse block entries equal the same solve's neighbours) for sampled j in
    the first, a middle and the last block and in the arrow; the mixed path against
    FP64 (logdet 1e-8 relative, diag 1e-4)."""
    ns, nt, nd = 2865, 100, 6
    say(f"c2 full size ns={ns} nt={nt} nd={nd}")
    n = ns * nt + nd
    js = [0, 17, ns - 1, 50 * ns + 1000, (nt - 1) * ns + 5, n - 3, n - 1]
    E = np.zeros((len(js), n))
    for c, j in enumerate(js):
        E[c, j] = 1.0
    rhs = E.reshape(-1)   # column-major n x len(js)
    with rgf_amd.RGF.generated(ns, nt, nd, seed=2024) as s:
        b = np.random.default_rng(11).standard_normal(n)
        x = s.solve(None, b, 1)
        assert s.residualNormNormalized(x, b) < 1e-14
        ld = s.logDet()
        d = s.RGFdiag()
        X = s.solve(None, rhs, len(js)).reshape(len(js), n)
        # only the sub-diagonal and arrow blocks (the diagonal blocks alone are 6.6 GB)
        sub = rgf_amd.DeviceBuffer((nt - 1) * ns * ns)
        arrow = rgf_amd.DeviceBuffer(nt * nd * ns)
        try:
            assert _lib.load().rgf_selinv_blocks_device(s._h, None, sub.ptr, arrow.ptr, None) == 0
            s.synchronize()
            sub0 = sub.download(ns * ns).reshape(ns, ns, order="F")
            arr = arrow.download().reshape(nt, ns, nd).transpose(0, 2, 1)
        finally:
            sub.free()
            arrow.free()
    for c, j in enumerate(js):
        assert abs(X[c, j] - d[j]) <= 1e-10 * abs(d[j]), j
    # Sigma_{1,0}[:, 17] is column 17 of A^-1 in rows ns..2ns-1
    assert np.abs(sub0[:, 17] - X[1, ns:2 * ns]).max() <= 1e-10 * np.abs(X[1]).max()
    # Sigma_{d,99}[:, 5] is column (nt-1)*ns+5 of A^-1 in the arrow rows
    assert np.abs(arr[nt - 1][:, 5] - X[4, nt * ns:]).max() <= 1e-10 * np.abs(X[4]).max()
    assert (d > 0).all()
    opts = rgf_amd.default_options(precision=rgf_amd.RGF_MIXED_BF16X3)
    with rgf_amd.RGF.generated(ns, nt, nd, seed=2024, options=opts) as m:
        assert abs(m.logDet() - ld) <= 1e-8 * abs(ld)
        assert np.abs(m.RGFdiag() / d - 1).max() <= 1e-4


def test_cli_matches_reference_driver_format(gpu, tmp_path):
    say("RGFSolver CLI")
    ia, ja, a = oracle.closed_form_csc(3, 3, 2)
    mpath = tmp_path / "A_11.dat"
    oracle.write_dat(str(mpath), 11, ia, ja, a)
    b = fixtures.rhs_formula(11)[:, 0]
    bpath = tmp_path / "b.dat"
    bpath.write_text("\n".join(repr(float(v)) for v in b) + "\n")
    out = subprocess.run([os.path.join(ROOT, "rgf_amd", "RGFSolver"), "3", "3", "2", str(mpath), str(bpath)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    full = [l for l in lines if l.startswith("logdet (full precision):")][0]
    assert float(full.split(":")[1]) == pytest.approx(G["closed_11x11"]["logdet"], rel=1e-12)
    for key in ("factorize time:", "solve time:", "logdet time:", "RGF time:", "Residual norm:",
                "Residual norm normalized:"):
        assert any(l.startswith(key) for l in lines), key
    x = np.array([float(l) for l in lines[-11:]])
    xg = np.array(G["closed_11x11"]["x_rhs_formula_2"])[:11]
    assert np.abs(x - xg).max() <= 1e-12 * np.abs(xg).max()


# ------------------------------------------------------------------ batched handles
@pytest.mark.parametrize("sched", [rgf_amd.RGF_CHOLESKY_TASK_GRAPH, rgf_amd.RGF_CHOLESKY_TILE_LAUNCHES])
@pytest.mark.parametrize("ns,nt,nd,count", [(100, 6, 5, 3), (65, 3, 0, 4), (8, 5, 3, 5), (200, 3, 4, 6)])
def test_batch_matches_single_and_oracle(gpu, ns, nt, nd, count, sched):
    """c3-style batch (per-instance theta): every matrix of the batch gives bitwise the
    same logdet and diag(A^-1) as that matrix alone on the same Cholesky schedule, and
    matches the oracle."""
    say(f"batch ns={ns} nt={nt} nd={nd} count={count} schedule={sched}")
    seed = 99
    opts = rgf_amd.default_options(cholesky=sched)
    with rgf_amd.RGF.generated(ns, nt, nd, seed=seed, matrix_id=10, count=count, theta_mode=1,
                               options=opts) as hb:
        assert hb.count == count
        lds = hb.logdet_batch()
        diags = hb.selinv_diag_batch()
        mats = []
        for b in range(count):
            hb.select(b)
            mats.append(hb.export_matrix_mf())
            assert hb.logDet() == lds[b]
    for b in range(count):
        td = 0.5 + 1.5 * oracle.gen_uniform(seed, 10 + b, 0)
        tc = 0.1 + 0.9 * oracle.gen_uniform(seed, 10 + b, 1)
        H = oracle.generate(seed, ns, nt, nd, matrix_id=10 + b, theta_d=td, theta_c=tc)
        assert np.abs(mats[b] - H).max() <= 1e-13 * np.abs(H).max()
        L = oracle.factor(mats[b], ns, nt, nd)
        assert abs(lds[b] - oracle.logdet(L, ns, nt, nd)) <= 1e-11 * abs(lds[b])
        assert np.abs(diags[b] / oracle.selinv(L, ns, nt, nd, exact=True) - 1).max() <= 1e-10
        with rgf_amd.RGF.generated(ns, nt, nd, seed=seed, matrix_id=10 + b, theta_d=td, theta_c=tc,
                                   options=opts) as h1:
            assert np.array_equal(h1.export_matrix_mf(), mats[b])
            assert h1.logDet() == lds[b]
            assert np.array_equal(h1.RGFdiag(), diags[b])


def test_batch_in_place_regenerate(gpu):
    say("batch keep_matrix=0 + regenerate")
    opts = rgf_amd.default_options(keep_matrix=0)
    with rgf_amd.RGF.generated(40, 4, 2, seed=5, count=3, theta_mode=1, options=opts) as h:
        l1 = h.logdet_batch()
        d1 = h.selinv_diag_batch()
        h.regenerate()
        h.factorize()
        assert np.array_equal(h.logdet_batch(), l1)
        assert np.array_equal(h.selinv_diag_batch(), d1)
    keep = rgf_amd.default_options(keep_matrix=1)
    assert rgf_amd.footprint_bytes(2865, 100, 6, 1, keep) > 2 * 8 * 100 * 5736 * 2865
    mixed = rgf_amd.default_options(keep_matrix=1, precision=rgf_amd.RGF_MIXED_BF16X3)
    # the mixed path adds bf16 hi/lo copies of both operands of its largest product
    assert rgf_amd.footprint_bytes(2865, 100, 6, 1, mixed) >= rgf_amd.footprint_bytes(2865, 100, 6, 1, keep) + \
        2 * 2 * 2 * 2871 * 2865


# ------------------------------------------------------------------ mixed precision (split-BF16 products)
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(16, 16, 32), (130, 70, 33), (64, 200, 300), (300, 260, 129)])
@pytest.mark.parametrize("kind", [256, 512])
def test_gemm_bf16x3_kernel(gpu, ta, tb, m, n, k, kind):
    """Split-BF16 product error is bounded by ~2^-16 of sum_k |a_ik||b_kj| (hi*hi + hi*lo +
    lo*hi kept, FP32 accumulation); the layout of the output is checked exactly by the
    same bound (a misplaced element would be off by O(1)).  kind 256 splits while staging,
    512 splits the operands into bf16 copies first."""
    say(f"gemm bf16x3 kind={kind} ta={ta} tb={tb} m={m} n={n} k={k}")
    rng = np.random.default_rng(m + 3 * n + 7 * k)
    lda, ldb, ldc = (k if ta else m) + 1, (n if tb else k) + 2, m + 3
    A = np.asfortranarray(rng.standard_normal((lda, m if ta else k)))
    B = np.asfortranarray(rng.standard_normal((ldb, k if tb else n)))
    C = np.asfortranarray(rng.standard_normal((ldc, n)))
    opA = (A[:k, :m].T if ta else A[:m, :k])
    opB = (B[:n, :k].T if tb else B[:k, :n])
    ref = C.copy()
    ref[:m, :n] = 1.5 * opA @ opB - 0.5 * C[:m, :n]
    bound = 1.5 * np.abs(opA) @ np.abs(opB)
    Cf = C.reshape(-1, order="F").copy()
    rc = _lib.load().rgf_debug_gemm(ta, tb, m, n, k, 1.5, dptr(A.reshape(-1, order="F").copy()), lda,
                                    dptr(B.reshape(-1, order="F").copy()), ldb, -0.5, dptr(Cf), ldc, kind)
    assert rc == 0
    got = Cf.reshape(ldc, n, order="F")
    err = np.abs(got[:m, :n] - ref[:m, :n])
    assert (err <= 4e-5 * bound + 1e-12).all(), (err / (bound + 1e-300)).max()
    assert np.array_equal(got[m:, :], C[m:, :])


@pytest.mark.parametrize("ta,tb,flags", [(0, 1, 1), (0, 0, 8), (0, 1, 16), (1, 0, 1), (1, 0, 1 | 4 | 8),
                                         (0, 0, 2), (1, 1, 0)])
@pytest.mark.parametrize("m,n,k", [(200, 200, 200), (333, 333, 333), (70, 70, 70)])
@pytest.mark.parametrize("kb", [32, 64])
@pytest.mark.parametrize("swz", [0, 1])
def test_gemm_bf16x3_presplit_matches_staged(gpu, ta, tb, flags, m, n, k, kb, swz):
    """The pre-split product runs the same MFMA sequence on the same hi/lo values as the
    staged one, over the same k ranges (triangle flags included), so both give the
    same bits -- with 32 or 64 k per LDS stage of the pre-split kernel, padded or
    XOR-swizzled LDS rows."""
    say(f"gemm bf16x3 presplit vs staged ta={ta} tb={tb} flags={flags} {m}x{n}x{k} kb={kb} swz={swz}")
    lib = _lib.load()
    old = lib.rgf_debug_set_tuning(1, kb)
    old_sw = lib.rgf_debug_set_tuning(3, swz)
    assert old in (32, 64) and old_sw in (0, 1)
    try:
        _presplit_vs_staged(ta, tb, flags, m, n, k)
    finally:
        lib.rgf_debug_set_tuning(1, old)
        lib.rgf_debug_set_tuning(3, old_sw)


def _presplit_vs_staged(ta, tb, flags, m, n, k):
    rng = np.random.default_rng(flags + m)
    lda, ldb, ldc = (k if ta else m), (n if tb else k), m + 1
    A = rng.standard_normal(lda * (m if ta else k))
    B = rng.standard_normal(ldb * (k if tb else n))
    C = rng.standard_normal(ldc * n)
    out = []
    for kind in (256, 512):
        Cf = C.copy()
        rc = _lib.load().rgf_debug_gemm(ta, tb, m, n, k, 1.0, dptr(A.copy()), lda, dptr(B.copy()), ldb, 1.0,
                                        dptr(Cf), ldc, kind | flags)
        assert rc == 0
        out.append(Cf)
    assert np.array_equal(out[0], out[1])


def _op_zero_triangles(X, ld, rows, k, trans, lower):
    """Zero op(X)[i][kk] for kk > i (lower) or kk < i (upper); op(X) = X (rows x k) or X^T."""
    for i in range(rows):
        for kk in range(k):
            if (kk > i) if lower else (kk < i):
                X[(kk + i * ld) if trans else (i + kk * ld)] = 0.0


@pytest.mark.parametrize("ta,tb,flags", [(0, 1, 1), (0, 0, 0), (0, 0, 8), (0, 1, 16), (1, 0, 1), (1, 0, 1 | 4 | 8),
                                         (0, 0, 2), (1, 1, 0)])
@pytest.mark.parametrize("m,n,k", [(333, 333, 333), (70, 70, 70), (300, 520, 260), (600, 200, 600)])
def test_gemm_bf16x3_wide_matches_staged(gpu, ta, tb, flags, m, n, k):
    """The 256 x 128 pre-split kernel (buffer loads into LDS, forced here for every launch
    size of the products without triangular operands; those with them keep the 128 x 128
    kernel) sums every output over the same k-blocks in the same MFMA order as the staged
    kernel, so both give the same bits; the operands hold the zeros their triangle flags
    promise."""
    say(f"gemm bf16x3 wide vs staged ta={ta} tb={tb} flags={flags} {m}x{n}x{k}")
    lib = _lib.load()
    rng = np.random.default_rng(flags + m + 7 * n)
    lda, ldb, ldc = (k if ta else m), (n if tb else k), m + 1
    A = rng.standard_normal(lda * (m if ta else k))
    B = rng.standard_normal(ldb * (k if tb else n))
    C = rng.standard_normal(ldc * n)
    # op(A) is m x k: ALOWER zeroes kk > i, AUPPER kk < i; op(B)^T is n x k: BLOWER
    # (op(B)[kk][j] == 0 for kk < j) zeroes kk < j of row j, BUPPER kk > j
    if flags & 2:
        _op_zero_triangles(A, lda, m, k, ta, True)
    if flags & 4:
        _op_zero_triangles(A, lda, m, k, ta, False)
    if flags & 8:
        _op_zero_triangles(B, ldb, n, k, not tb, False)
    if flags & 16:
        _op_zero_triangles(B, ldb, n, k, not tb, True)
    out = []
    for kind, wide in ((256, None), (512, 0), (512, 1 << 30)):
        old = lib.rgf_debug_set_tuning(14, wide) if wide is not None else None
        try:
            Cf = C.copy()
            rc = lib.rgf_debug_gemm(ta, tb, m, n, k, 1.0, dptr(A.copy()), lda, dptr(B.copy()), ldb, 1.0, dptr(Cf), ldc,
                                    kind | flags)
        finally:
            if old is not None:
                lib.rgf_debug_set_tuning(14, old)
        assert rc == 0
        out.append(Cf)
    assert np.array_equal(out[0], out[2])   # 128 / 64 tiles on the blocked split copies
    assert np.array_equal(out[0], out[1])   # 256 x 128 tiles


@pytest.mark.parametrize("count", [1, 3])
def test_mixed_sweep_wide_kernel_same_bits(gpu, count):
    """Whole mixed sweeps (Schur, trsm, Takahashi products with their split-copy epilogues,
    lower-stored and triangular launches) with the 256 x 128 product kernel forced for every
    pre-split launch give the bits of the sweeps without it."""
    say(f"mixed sweep wide kernel forced vs off, count={count}")
    lib = _lib.load()
    opts = rgf_amd.default_options(precision=rgf_amd.RGF_MIXED_BF16X3)
    res = []
    for wide in (0, 1 << 30):
        old = lib.rgf_debug_set_tuning(14, wide)
        try:
            with rgf_amd.RGF.generated(300, 4, 5, seed=12, count=count, theta_mode=1 if count > 1 else 0,
                                       options=opts) as h:
                res.append((h.logdet_batch(), h.selinv_diag_batch()))
        finally:
            lib.rgf_debug_set_tuning(14, old)
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])


def test_mixed_kept_splits_same_bits(gpu):
    """Sweeps that keep split copies (written by product epilogues, the mirror and the
    L_ii^-1 merges, read again by later products) give the bits of sweeps that split every
    operand where it is read.  At this size millions of split values are written into
    copies, so some land on a bf16 tie after the FP64 -> FP32 step: every split rounds the
    FP64 value once, wherever it is made (splits that rounded twice in the L_ii^-1 copies
    changed 12 diagonal entries of this case)."""
    say("mixed sweep keep_splits 1 vs 0, ns=1500 nt=4 nd=6 count=2")
    res = []
    for keep in (1, 0):
        opts = rgf_amd.default_options(precision=rgf_amd.RGF_MIXED_BF16X3, keep_splits=keep)
        with rgf_amd.RGF.generated(1500, 4, 6, seed=31, count=2, theta_mode=1, options=opts) as h:
            res.append((h.logdet_batch(), h.selinv_diag_batch()))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])


@pytest.mark.parametrize("prec", [rgf_amd.RGF_FP64, rgf_amd.RGF_MIXED_BF16X3])
def test_autotune_keeps_bits(gpu, prec, tmp_path):
    """Tile shapes tuned per launch signature (every candidate timed on scratch outputs the
    first time a signature is seen) leave logdet and diag(A^-1) bitwise as with the size
    thresholds; a saved cache, loaded again, gives the same bits too."""
    say(f"autotune bits prec={prec}")

    def run():
        with rgf_amd.RGF.generated(300, 5, 6, seed=5, count=3, theta_mode=1,
                                   options=rgf_amd.default_options(precision=prec)) as h:
            return h.logdet_batch(), h.selinv_diag_batch()

    prev = rgf_amd.set_autotune(0)
    try:
        rgf_amd.autotune_clear()
        base = run()
        rgf_amd.set_autotune(2)
        tuned = run()
        n = rgf_amd.autotune_entries()
        assert n > 0
        path = str(tmp_path / "tune.txt")
        rgf_amd.autotune_save(path)
        rgf_amd.autotune_clear()
        rgf_amd.autotune_load(path)
        assert rgf_amd.autotune_entries() == n
        rgf_amd.set_autotune(1)
        cached = run()
        assert rgf_amd.autotune_entries() == n      # nothing tuned again
    finally:
        rgf_amd.autotune_clear()
        rgf_amd.set_autotune(prev)
    for got in (tuned, cached):
        assert np.array_equal(got[0], base[0])
        assert np.array_equal(got[1], base[1])


def test_group_uneven_shapes_match_alone(gpu):
    """Handles of uneven shapes, batch counts and precisions factorized and selected-inverted
    as one group graph (captured, then replayed after new matrix values) give the bits of
    the same handles run alone; an in-place handle must hold a matrix again first."""
    say("group of uneven handles vs alone")
    shapes = [(300, 4, 5, 2, rgf_amd.RGF_FP64, 1), (200, 6, 0, 1, rgf_amd.RGF_MIXED_BF16X3, 1),
              (130, 3, 7, 3, rgf_amd.RGF_FP64, 0), (64, 5, 3, 1, rgf_amd.RGF_MIXED_BF16X3, 1)]

    def make(q, seed):
        ns, nt, nd, cnt, prec, keep = shapes[q]
        return rgf_amd.RGF.generated(ns, nt, nd, seed=seed + q, count=cnt, theta_mode=1,
                                     options=rgf_amd.default_options(precision=prec, keep_matrix=keep))

    def alone(seed):
        out = []
        for q in range(len(shapes)):
            with make(q, seed) as h:
                out.append((h.logdet_batch(), h.selinv_diag_batch()))
        return out

    hs = [make(q, 40) for q in range(len(shapes))]
    diags = [rgf_amd.DeviceBuffer(h.n * h.count) for h in hs]
    ref = alone(40)
    try:
        with rgf_amd.Group(hs) as g:
            for rnd in range(2):
                if rnd == 1:   # the captured graphs replayed
                    with pytest.raises(rgf_amd.RGFError) as e:
                        g.factorize()          # the in-place handle holds its factor
                    assert e.value.code == rgf_amd.RGF_ESTATE
                    for h in hs:
                        h.regenerate()         # the in-place handle's factor is its matrix again
                    for d in diags:
                        d.upload(np.full(d.count, np.nan))
                g.factorize()
                g.selinv_diag_device(diags)
                g.synchronize()
                for q, h in enumerate(hs):
                    assert np.array_equal(h.logdet_batch(), ref[q][0]), (rnd, q)
                    assert np.array_equal(diags[q].download(), ref[q][1].reshape(-1)), (rnd, q)
    finally:
        for d in diags:
            d.free()
        for h in hs:
            h.close()


@pytest.mark.parametrize("ns,nt,nd", [(200, 6, 6), (65, 4, 5), (300, 3, 0)])
def test_mixed_precision_vs_fp64(gpu, ns, nt, nd):
    """Mixed path (split-BF16 Schur / solve / Takahashi products, FP64 potrf, logdet and
    refinement): logdet within 1e-8 relative of FP64, refined solves at FP64 accuracy,
    diag(A^-1) within 1e-4 relative (stated, not refined)."""
    say(f"mixed ns={ns} nt={nt} nd={nd}")
    opts = rgf_amd.default_options(precision=rgf_amd.RGF_MIXED_BF16X3, refine_tol=1e-14)
    with rgf_amd.RGF.generated(ns, nt, nd, seed=31, options=opts) as s:
        A_mf = s.export_matrix_mf()
        ld = s.logDet()
        n = s.n
        b = np.random.default_rng(2).standard_normal((n, 16))
        x = s.solve(None, b, 16)
        its, res = s.refine_stats()
        d = s.RGFdiag()
    L = oracle.factor(A_mf, ns, nt, nd)
    ld0 = oracle.logdet(L, ns, nt, nd)
    assert abs(ld - ld0) <= 1e-8 * abs(ld0)
    x0 = oracle.solve(L, ns, nt, nd, b)
    assert np.abs(x - x0).max() <= 1e-12 * np.abs(x0).max()
    assert res <= 1e-14 and its >= 1
    d0 = oracle.selinv(L, ns, nt, nd, exact=True)
    assert np.abs(d / d0 - 1).max() <= 1e-4
    say(f"  logdet rel {abs(ld - ld0) / abs(ld0):.2e}  diag rel {np.abs(d / d0 - 1).max():.2e}  refine {its} steps")


def test_mixed_batch_matches_single(gpu):
    say("mixed batch vs single")
    opts = rgf_amd.default_options(precision=rgf_amd.RGF_MIXED_BF16X3)
    with rgf_amd.RGF.generated(150, 5, 4, seed=8, count=3, theta_mode=1, options=opts) as hb:
        lds = hb.logdet_batch()
        diags = hb.selinv_diag_batch()
    for b in range(3):
        td = 0.5 + 1.5 * oracle.gen_uniform(8, b, 0)
        tc = 0.1 + 0.9 * oracle.gen_uniform(8, b, 1)
        with rgf_amd.RGF.generated(150, 5, 4, seed=8, matrix_id=b, theta_d=td, theta_c=tc, options=opts) as h1:
            assert h1.logDet() == lds[b]
            assert np.array_equal(h1.RGFdiag(), diags[b])


# ------------------------------------------------------------------ selected-inverse blocks
@pytest.mark.parametrize("prec,tol", [(rgf_amd.RGF_FP64, 1e-10), (rgf_amd.RGF_MIXED_BF16X3, 1e-4)])
@pytest.mark.parametrize("ns,nt,nd", [(40, 4, 3), (64, 3, 0), (33, 1, 5), (130, 3, 7)])
def test_selinv_blocks_vs_dense_inverse(gpu, prec, tol, ns, nt, nd):
    """Every block of A^-1 on the pattern of A (Sigma_ii, Sigma_{i+1,i}, Sigma_{d,i},
    Sigma_dd) against numpy's dense inverse of the handle's own matrix; the diagonal of
    the blocks equals diag(A^-1) from RGFdiag."""
    say(f"selinv blocks prec={prec} ns={ns} nt={nt} nd={nd}")
    opts = rgf_amd.default_options(precision=prec, keep_matrix=1)
    with rgf_amd.RGF.generated(ns, nt, nd, seed=17, options=opts) as h:
        A = oracle.mf_to_dense(h.export_matrix_mf(), ns, nt, nd)
        S = np.linalg.inv(A)
        blk = h.selinv_blocks()
        d = h.RGFdiag()
    scale = np.abs(S).max()
    for i in range(nt):
        r = slice(i * ns, (i + 1) * ns)
        assert np.abs(blk["diag"][i] - S[r, r]).max() <= tol * scale, ("diag", i)
        if i < nt - 1:
            r1 = slice((i + 1) * ns, (i + 2) * ns)
            assert np.abs(blk["sub"][i] - S[r1, r]).max() <= tol * scale, ("sub", i)
        if nd > 0:
            assert np.abs(blk["arrow"][i] - S[nt * ns:, r]).max() <= tol * scale, ("arrow", i)
    if nd > 0:
        assert np.abs(blk["corner"] - S[nt * ns:, nt * ns:]).max() <= tol * scale
    diag_blocks = np.concatenate([np.diag(blk["diag"][i]) for i in range(nt)] +
                                 ([np.diag(blk["corner"])] if nd > 0 else []))
    assert np.array_equal(diag_blocks, d)


def test_selinv_blocks_need_exact_mode(gpu):
    opts = rgf_amd.default_options(selinv_mode=rgf_amd.RGF_SELINV_REFERENCE)
    with rgf_amd.RGF.generated(40, 3, 2, seed=1, options=opts) as h:
        with pytest.raises(rgf_amd.RGFError):
            h.selinv_blocks()


# ------------------------------------------------------------------ HIP graph replays
@pytest.mark.parametrize("prec", [rgf_amd.RGF_FP64, rgf_amd.RGF_MIXED_BF16X3])
def test_graph_replays_match_eager(gpu, prec):
    """Sweeps and solve pieces captured in HIP graphs (options.graphs = 1): every replay --
    with eager work (regeneration, solves, host copies) and synchronizations between the
    replays, and a re-capture after the kernel-timing configuration changed -- gives
    bitwise the eager handle's logdet, diag(A^-1), solutions (1 and 16 RHS, host and
    device buffers) and refinement statistics."""
    say(f"graph replays prec={prec}")
    ns, nt, nd, count = 150, 5, 4, 2
    base = dict(precision=prec, keep_matrix=1)
    eager = rgf_amd.RGF.generated(ns, nt, nd, seed=21, count=count, theta_mode=1,
                                  options=rgf_amd.default_options(**base))
    graph = rgf_amd.RGF.generated(ns, nt, nd, seed=21, count=count, theta_mode=1,
                                  options=rgf_amd.default_options(graphs=1, **base))
    n = eager.n
    b = np.random.default_rng(3).standard_normal(n)
    b16 = np.random.default_rng(4).standard_normal(n * 16)
    buf = rgf_amd.DeviceBuffer(n * count)
    bdev = rgf_amd.DeviceBuffer(n * 16)
    xdev = rgf_amd.DeviceBuffer(n * 16)
    bdev.upload(b16)
    try:
        for rnd in range(5):
            if rnd == 2:
                rgf_amd.kernel_timing_enable(0)    # the next capture records timing events
            if rnd == 4:
                rgf_amd.kernel_timing_enable(-1)   # and the one after that drops them again
            for h in (eager, graph):
                if rnd == 3:
                    h.regenerate()
                h.factorize()
            assert np.array_equal(graph.logdet_batch(), eager.logdet_batch())
            assert np.array_equal(graph.selinv_diag_batch(), eager.selinv_diag_batch())
            graph.selinv_diag_device(buf)   # asynchronous on the handle's stream
            graph.synchronize()
            assert np.array_equal(buf.download().reshape(count, n), eager.selinv_diag_batch())
            for m in range(count):
                eager.select(m)
                graph.select(m)
                assert np.array_equal(graph.solve(None, b, 1), eager.solve(None, b, 1))
                x16 = eager.solve(None, b16, 16)
                assert np.array_equal(graph.solve(None, b16, 16), x16)
                for rep in range(2):   # device buffers: the captured solve pieces replay
                    graph.solve_device(xdev, bdev, 16)
                    graph.synchronize()
                    assert np.array_equal(xdev.download(), x16)
                    assert graph.refine_stats() == eager.refine_stats()
    finally:
        rgf_amd.kernel_timing_enable(-1)
        bdev.free()
        xdev.free()
        buf.free()
        eager.close()
        graph.close()


# ------------------------------------------------------------------ variant switches keep the bits
@pytest.mark.parametrize("prec", [rgf_amd.RGF_FP64, rgf_amd.RGF_MIXED_BF16X3])
@pytest.mark.parametrize("key,alt", [(4, 0), (8, 0), (11, 0), (12, 0), (20, 0), (21, 0)])
def test_sweep_variant_switches_keep_bits(gpu, prec, key, alt):
    """The adopted kernel variants change how the same products are loaded, dealt over
    the XCDs or launched together, not what is summed: logdet and diag(A^-1) of a batch and
    of a single matrix are bitwise those of the alternative (key 4 masked loads, 8 tile
    ranges for triangular products, 11 separate launches of X^T X and W^T Z, 12 W of the next
    block in a launch of its own, 20 / 21 the task-graph Cholesky's claim order)."""
    say(f"switch {key}={alt} prec={prec}")
    lib = _lib.load()
    opts = rgf_amd.default_options(precision=prec)

    def run():
        with rgf_amd.RGF.generated(330, 4, 5, seed=21, count=3, theta_mode=1, options=opts) as hb:
            lb, db = hb.logdet_batch(), hb.selinv_diag_batch()
        with rgf_amd.RGF.generated(200, 5, 0, seed=22, options=opts) as h1:
            l1, d1 = h1.logDet(), h1.RGFdiag()
        return lb, db, l1, d1

    ref = run()
    old = lib.rgf_debug_set_tuning(key, alt)
    assert old >= 0
    try:
        got = run()
    finally:
        lib.rgf_debug_set_tuning(key, old)
    assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1])
    assert ref[2] == got[2] and np.array_equal(ref[3], got[3])


# ------------------------------------------------------------------ footprints and large batches
@pytest.mark.parametrize("prec", [rgf_amd.RGF_FP64, rgf_amd.RGF_MIXED_BF16X3])
@pytest.mark.parametrize("ns,nt,nd,count", [(130, 4, 5, 1), (200, 3, 0, 1), (64, 5, 3, 3)])
def test_footprint_covers_device_bytes(gpu, prec, ns, nt, nd, count):
    """rgf_footprint_bytes (the serving loop's admission figure) is at least what a handle
    holds after factorize, 16-RHS solve and selected inversion, up to the 256-byte
    rounding of each of its allocations."""
    say(f"footprint prec={prec} ns={ns} nt={nt} nd={nd} count={count}")
    opts = rgf_amd.default_options(precision=prec, keep_matrix=1)
    with rgf_amd.RGF.generated(ns, nt, nd, seed=4, count=count, theta_mode=1, options=opts) as h:
        h.factorize()
        h.solve(None, np.ones((h.n, 16)), 16)
        h.selinv_diag_batch()
        held = h.info().device_bytes
    fp = rgf_amd.footprint_bytes(ns, nt, nd, count, opts)
    assert fp + 256 * 40 >= held, (fp, held)
    assert fp <= 1.5 * held + (1 << 20)


def test_large_batch_count_with_arrow(gpu):
    """More matrices than 64 * nt in one handle with nd > 0: the arrow-row update's
    split-k partials for the whole batch fit its buffer; sampled instances equal
    single-matrix handles bitwise and the last matches the oracle."""
    ns, nt, nd, count = 48, 2, 3, 200
    say(f"large batch ns={ns} nt={nt} nd={nd} count={count}")
    tiles = rgf_amd.default_options(cholesky=rgf_amd.RGF_CHOLESKY_TILE_LAUNCHES)   # the batch's schedule
    with rgf_amd.RGF.generated(ns, nt, nd, seed=6, count=count, theta_mode=1) as hb:
        lds = hb.logdet_batch()
        diags = hb.selinv_diag_batch()
    for b in (0, 129, count - 1):
        td = 0.5 + 1.5 * oracle.gen_uniform(6, b, 0)
        tc = 0.1 + 0.9 * oracle.gen_uniform(6, b, 1)
        with rgf_amd.RGF.generated(ns, nt, nd, seed=6, matrix_id=b, theta_d=td, theta_c=tc, options=tiles) as h1:
            assert h1.logDet() == lds[b]
            assert np.array_equal(h1.RGFdiag(), diags[b])
            L = oracle.factor(h1.export_matrix_mf(), ns, nt, nd)
        assert abs(lds[b] - oracle.logdet(L, ns, nt, nd)) <= 1e-11 * abs(lds[b])
        assert np.abs(diags[b] / oracle.selinv(L, ns, nt, nd, exact=True) - 1).max() <= 1e-10


def test_generated_residual_at_perturbed_x(gpu):
    """residualNorm / residualNormNormalized (RGF.H:631-663) against numpy's dense A x at a
    perturbed solution, on a generated matrix with an arrowhead."""
    ns, nt, nd = 150, 4, 5
    say(f"residual norms ns={ns} nt={nt} nd={nd}")
    with rgf_amd.RGF.generated(ns, nt, nd, seed=41) as s:
        A = oracle.mf_to_dense(s.export_matrix_mf(), ns, nt, nd)
        b = np.random.default_rng(2).standard_normal(s.n)
        x = s.solve(None, b, 1)
        assert s.residualNormNormalized(x, b) < 1e-14
        for scale in (1e-3, 1.0):
            xp = x + scale * np.cos(np.arange(s.n) * 0.7)
            r = np.linalg.norm(b - A @ xp)
            assert s.residualNorm(xp, b) == pytest.approx(r, rel=1e-12)
            assert s.residualNormNormalized(xp, b) == pytest.approx(r / np.linalg.norm(b), rel=1e-12)


# ------------------------------------------------------------------ solves through the kept W_i
@pytest.mark.parametrize("ns,nt,nd,count", [(42, 3, 0, 1), (65, 4, 5, 1), (130, 3, 8, 1), (200, 6, 6, 1),
                                            (1000, 5, 4, 1), (130, 4, 3, 3)])
def test_solve_with_kept_w(gpu, ns, nt, nd, count):
    """keep_w = 1: after a selected inversion, solves chain x through the kept W_i = [B_i; D_i] X_i
    (one product per block and direction) -- the same solution as the block solve to rounding,
    for 1, 3, 16 and 17 right-hand sides, through the oracle on small shapes; after a new
    factorization (the kept W_i stale) the block solve again, bitwise that of keep_w = 0; HIP
    graph replays keyed apart for the two sequences."""
    say(f"kept W solve ns={ns} nt={nt} nd={nd} count={count}")
    gen = dict(seed=77 + ns, count=count, theta_mode=1 if count > 1 else 0)
    with rgf_amd.RGF.generated(ns, nt, nd, options=rgf_amd.default_options(), **gen) as h0, \
         rgf_amd.RGF.generated(ns, nt, nd, options=rgf_amd.default_options(keep_w=1, graphs=1), **gen) as h1:
        n = h0.n
        rng = np.random.default_rng(ns)
        b = rng.standard_normal((n, 17))
        for mb in range(count):
            if count > 1:
                h0.select(mb)
                h1.select(mb)
            if mb == 0:   # before any selected inversion: the block solve
                x_before = h1.solve(None, b[:, :3].copy(), 3)
                assert np.array_equal(x_before, h0.solve(None, b[:, :3].copy(), 3))
            h0.RGFdiag()
            d1 = h1.RGFdiag()
            assert np.array_equal(d1, h0.RGFdiag())           # keeping W changes no selected-inverse bit
            for nr in (1, 3, 16, 17):
                bb = b[:, :nr].copy()
                xw = h1.solve(None, bb, nr)
                x0 = h0.solve(None, bb, nr)
                assert np.abs(xw - x0).max() <= 1e-12 * np.abs(x0).max(), (mb, nr)
                assert np.array_equal(h1.solve(None, bb, nr), xw)   # graph replay
                assert h1.residualNormNormalized(xw[:, :1].copy(), bb[:, :1].copy()) <= 1e-13
            if ns * nt <= 2000 and count == 1:
                A_mf = oracle.generate(77 + ns, ns, nt, nd)
                L = oracle.factor(A_mf, ns, nt, nd)
                xo = oracle.solve(L, ns, nt, nd, b[:, :1])
                assert np.abs(h1.solve(None, b[:, :1].copy(), 1) - xo).max() <= 1e-10 * np.abs(xo).max()
        # a new factorization makes the kept W_i stale: the block solve again, bitwise keep_w = 0's
        h1.factorize()
        h0.factorize()
        assert np.array_equal(h1.solve(None, b[:, :1].copy(), 1), h0.solve(None, b[:, :1].copy(), 1))


def test_mixed_ignores_keep_w(gpu):
    """keep_w applies to FP64 handles: a mixed-precision handle with it solves (refined) and
    inverts bitwise as without it."""
    say("mixed handle with keep_w")
    ns, nt, nd = 300, 5, 4
    b = None
    out = []
    for kw in (0, 1):
        opts = rgf_amd.default_options(precision=rgf_amd.RGF_MIXED_BF16X3, keep_w=kw)
        with rgf_amd.RGF.generated(ns, nt, nd, seed=5, options=opts) as h:
            d = h.RGFdiag()
            if b is None:
                b = np.random.default_rng(1).standard_normal((h.n, 16))
            out.append((d, h.solve(None, b, 16), h.solve(None, b[:, :1].copy(), 1)))
    for u, v in zip(out[0], out[1]):
        assert np.array_equal(u, v)
