"""Pins the CPU oracle (oracle/rgf_oracle.c) before it is trusted as the checker.

Each check is independent of the restatement it pins: dense numpy factorizations,
inverses and solves, the reference's own fixtures and derived goldens, published
known-answer vectors, and a planted mistake that must be caught.
"""
import os

import numpy as np
import pytest

from oracle import oracle
import fixtures

REF = "/root/reference/data"
G = fixtures.goldens()


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10 (Salmon et al., SC'11)
    assert oracle.philox4x32_10([0, 0, 0, 0], [0, 0]).tolist() == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert oracle.philox4x32_10([0xffffffff] * 4, [0xffffffff] * 2).tolist() == [0x408f276d, 0x41c83b0e,
                                                                                0xa20bc7c6, 0x6d5451fd]
    assert oracle.philox4x32_10([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                                [0xa4093822, 0x299f31d0]).tolist() == [0xd16cfe09, 0x94fdcceb, 0x5001e420,
                                                                       0x24126ea1]


def test_generator_normals_statistics():
    z = np.array([oracle.gen_normal(3, 0, 1, 7, e) for e in range(20000)])
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.03
    # pairs share a counter: distinct values
    assert z[0] != z[1]
    g = G["generator"]["gen_4_3_2_s1"]
    assert [oracle.gen_normal(1, 0, 1, 0, e) for e in range(4)] == pytest.approx(g["first_normals"], rel=0, abs=1e-15)


@pytest.mark.parametrize("key", ["gen_4_3_2_s1", "gen_42_3_0_s0"])
def test_generator_definition(key):
    g = G["generator"][key]
    ns, nt, nd = g["ns"], g["nt"], g["nd"]
    A = oracle.mf_to_dense(oracle.generate(g["seed"], ns, nt, nd), ns, nt, nd)
    assert np.allclose(A, A.T)
    off = np.abs(A).sum(1) - np.abs(np.diag(A))
    # dominance: a_kk - sum_{j!=k}|a_kj| equals the pre-dominance diagonal + 1 > 1
    assert np.all(np.diag(A) - off > 1.0)
    assert np.linalg.slogdet(A)[1] == pytest.approx(g["logdet"], rel=1e-13)
    assert np.trace(A) == pytest.approx(g["trace"], rel=1e-13)


@pytest.mark.parametrize("name,ns,nt,nd,fn", [("closed_9x9", 3, 3, 0, "A_9_9_ns3_nt3.dat"),
                                              ("closed_11x11", 3, 3, 2, "A_11_11_ns3_nt3_nd2.dat")])
def test_closed_form_regenerates_reference_fixture(name, ns, nt, nd, fn):
    ia, ja, a = oracle.closed_form_csc(ns, nt, nd)
    assert G[name]["matches_reference_file"]
    path = os.path.join(REF, fn)
    if os.path.exists(path):
        _, ia2, ja2, a2 = oracle.read_dat(path)
        assert np.array_equal(ia, ia2) and np.array_equal(ja, ja2) and np.array_equal(a, a2)


def _check_all(A, ns, nt, nd, g=None, tol=1e-12):
    n = A.shape[0]
    mf = oracle.dense_to_mf(A, ns, nt, nd)
    assert np.array_equal(oracle.mf_to_dense(mf, ns, nt, nd), np.tril(A) + np.tril(A, -1).T)
    L = oracle.factor(mf, ns, nt, nd)
    Lc = np.linalg.cholesky(A)
    # no fill in the block-arrow pattern: every block of L equals the dense factor's
    assert np.abs(oracle.mf_to_dense(L, ns, nt, nd, symmetric=False) - Lc).max() <= tol * np.abs(Lc).max() * 10
    ld = oracle.logdet(L, ns, nt, nd)
    assert ld == pytest.approx(np.linalg.slogdet(A)[1], rel=tol)
    b = fixtures.rhs_formula(n, 2)
    x = oracle.solve(L, ns, nt, nd, b)
    xd = np.linalg.solve(A, b)
    assert np.abs(x - xd).max() <= 1e-9 * np.abs(xd).max()
    de = oracle.selinv(L, ns, nt, nd, exact=True)
    assert np.abs(de / np.diag(np.linalg.inv(A)) - 1).max() < 1e-10
    dc = oracle.selinv(L, ns, nt, nd, exact=False)
    assert np.abs(dc / oracle.reference_selinv_dense(A, ns, nt, nd) - 1).max() < 1e-10
    if g is not None:
        assert ld == pytest.approx(g["logdet"], rel=1e-12)
        assert np.abs(de / np.array(g["diag_inv"]) - 1).max() < 1e-10
        assert np.abs(dc / np.array(g["diag_reference"]) - 1).max() < 1e-10
        xg = np.array(g["x_rhs_formula_2"]).reshape(n, 2, order="F")
        assert np.abs(x - xg).max() <= 1e-9 * np.abs(xg).max()
    return L, de, dc


def test_closed_form_11x11_goldens_and_reference_deviation():
    ia, ja, a = oracle.closed_form_csc(3, 3, 2)
    A = oracle.csc_to_dense(ia, ja, a, 11)
    L, de, dc = _check_all(A, 3, 3, 2, G["closed_11x11"])
    # the reference's selinv drops the arrow cross terms: 5.36e-5 relative on block 0
    dev = np.abs(dc / de - 1)
    assert 5.0e-5 < dev.max() < 5.5e-5
    assert dev[6:].max() < 1e-14   # last block and corner are exact in both


def test_closed_form_9x9_goldens():
    ia, ja, a = oracle.closed_form_csc(3, 3, 0)
    A = oracle.csc_to_dense(ia, ja, a, 9)
    _, de, dc = _check_all(A, 3, 3, 0, G["closed_9x9"])
    assert np.abs(dc / de - 1).max() < 1e-14   # nd = 0: both recursions are exact


def test_inla_126_committed_fixture():
    """The committed CSC arrays of data/A_126_126_ns42_nt3.dat (the GPU suite's input)
    are the file the goldens were derived from, and reproduce them through the oracle."""
    ia, ja, a, ns, nt, nd = fixtures.inla_126()
    assert fixtures.inla_126_sha256() == G["inla_126"]["reference_file_sha256"]
    assert int(ia[-1]) == G["inla_126"]["nnz"] and (ns, nt, nd) == (42, 3, 0)
    A = oracle.csc_to_dense(ia, ja, a, ns * nt + nd)
    _check_all(A, ns, nt, nd, G["inla_126"], tol=1e-11)
    path = os.path.join(REF, "A_126_126_ns42_nt3.dat")
    if os.path.exists(path):
        _, ia2, ja2, a2 = oracle.read_dat(path)
        assert np.array_equal(ia, ia2) and np.array_equal(ja, ja2) and np.array_equal(a, a2)


def test_inla_126_reference_fixture():
    path = os.path.join(REF, "A_126_126_ns42_nt3.dat")
    if not os.path.exists(path):
        pytest.skip("reference data not present")
    n, ia, ja, a = oracle.read_dat(path)
    A = oracle.csc_to_dense(ia, ja, a, n)
    mf = oracle.csc_to_mf(ia, ja, a, 42, 3, 0)
    assert np.array_equal(oracle.mf_to_dense(mf, 42, 3, 0), A)
    _check_all(A, 42, 3, 0, G["inla_126"], tol=1e-11)
    assert G["inla_126"]["logdet"] == pytest.approx(150.65253749112748, rel=1e-15)


@pytest.mark.parametrize("name", ["spde_126", "spde_126_nd2"])
def test_spde_like_goldens(name):
    A, ns, nt, nd = fixtures.spde_like(nd=G[name]["nd"])
    _check_all(A, ns, nt, nd, G[name], tol=1e-11)


@pytest.mark.parametrize("ns,nt,nd", [(8, 5, 3), (6, 1, 2), (5, 1, 0), (7, 4, 0), (16, 3, 7), (1, 6, 1),
                                      (2, 2, 2)])
def test_generated_shapes(ns, nt, nd):
    A = oracle.mf_to_dense(oracle.generate(5, ns, nt, nd), ns, nt, nd)
    _check_all(A, ns, nt, nd)


def test_nt1_with_arrow_fixed():
    """nt = 1, nd > 0: the reference skips the arrow trsm (RGF.H:258); the oracle does not."""
    A = oracle.mf_to_dense(oracle.generate(9, 5, 1, 3), 5, 1, 3)
    L = oracle.factor(oracle.dense_to_mf(A, 5, 1, 3), 5, 1, 3)
    assert np.abs(oracle.mf_to_dense(L, 5, 1, 3, symmetric=False) - np.linalg.cholesky(A)).max() < 1e-14


def test_getpos_matches_dense_layout():
    ns, nt, nd = 4, 3, 2
    n = ns * nt + nd
    A = np.arange(1, n * n + 1, dtype=float).reshape(n, n)
    A = np.tril(A) + np.tril(A, -1).T
    mf = oracle.dense_to_mf(A, ns, nt, nd)
    for c in range(n):
        for r in range(c, n):
            br, bc = r // ns, c // ns
            if c >= ns * nt or r >= ns * nt or br in (bc, bc + 1):
                assert mf[oracle.getpos(r, c, ns, nt, nd)] == A[r, c]
    dp = oracle.diag_pos(ns, nt, nd)
    assert np.array_equal(mf[dp], np.diag(A))


def test_planted_mistake_is_caught():
    """A sign error planted in a copy of the oracle's factor must fail the pins."""
    A = oracle.mf_to_dense(oracle.generate(2, 6, 3, 2), 6, 3, 2)
    L = oracle.factor(oracle.dense_to_mf(A, 6, 3, 2), 6, 3, 2)
    bad = L.copy()
    bad[6 + 2] = -bad[6 + 2]     # one coupling entry of B_0 (supernode 0, row ns+2, col 0)
    b = fixtures.rhs_formula(A.shape[0])
    x = oracle.solve(bad, 6, 3, 2, b)
    assert np.abs(A @ x - b).max() > 1e-6
    d = oracle.selinv(bad, 6, 3, 2, exact=True)
    assert np.abs(d / np.diag(np.linalg.inv(A)) - 1).max() > 1e-8


def test_residual_norm_matches_numpy():
    ia, ja, a = oracle.closed_form_csc(3, 3, 2)
    A = oracle.csc_to_dense(ia, ja, a, 11)
    x = np.linspace(-1, 1, 11)
    b = fixtures.rhs_formula(11)[:, 0].copy()
    r = oracle.residual_norm(ia, ja, a, 11, x, b)
    assert r == pytest.approx(np.linalg.norm(b - A @ x), rel=1e-14)
    rn = oracle.residual_norm(ia, ja, a, 11, x, b, normalized=True)
    assert rn == pytest.approx(np.linalg.norm(b - A @ x) / np.linalg.norm(b), rel=1e-14)


def test_csc_pattern_violation_rejected():
    ia = np.array([0, 2, 3, 4, 5, 6, 7, 8, 9, 10], dtype=np.uint64)
    ja = np.array([0, 7, 1, 2, 3, 4, 5, 6, 7, 8], dtype=np.uint64)   # (7,0) is outside the pattern
    a = np.ones(10)
    with pytest.raises(ValueError):
        oracle.csc_to_mf(ia, ja, a, 3, 3, 0)
