"""Pins of the complex restatement (oracle.zfactor_blocks / zsolve / zlogdet / zselinv,
RGF<CPX> of RGF.H:756-780 in complex-symmetric algebra) before it is used as a checker:
L L^T reconstructs A, solves and the diagonal of the inverse agree with numpy's dense
complex solve and inverse, exp(logdet) with numpy's determinant, real-valued input with
the real oracle, and a conjugated (Hermitian) variant planted in a copy fails the pins."""
import numpy as np
import pytest

import fixtures
from oracle import oracle

SHAPES = [(5, 4, 3), (6, 3, 0), (4, 1, 2), (7, 2, 7), (3, 5, 1)]


@pytest.mark.parametrize("ns,nt,nd", SHAPES)
def test_complex_oracle_against_dense(ns, nt, nd):
    A = fixtures.complex_symmetric(ns, nt, nd, seed=ns + 10 * nt + nd)
    assert np.array_equal(A, A.T) and np.abs(A.imag).max() > 0.1
    F = oracle.zfactor_blocks(A, ns, nt, nd)
    L = oracle.zblocks_to_dense(F)
    assert np.abs(L @ L.T - A).max() <= 1e-13 * np.abs(A).max()
    assert (np.diag(L).real > 0).all()                       # principal square roots
    b = np.random.default_rng(1).standard_normal((A.shape[0], 3)) + 1j
    x = oracle.zsolve(F, b)
    assert np.abs(x - np.linalg.solve(A, b)).max() <= 1e-12 * np.abs(x).max()
    d = oracle.zselinv(F, exact=True)
    di = np.diag(np.linalg.inv(A))
    assert np.abs(d - di).max() <= 1e-12 * np.abs(di).max()
    ld = oracle.zlogdet(F)
    sign, logabs = np.linalg.slogdet(A)
    assert ld.real == pytest.approx(logabs, rel=1e-13, abs=1e-13)
    k = (ld.imag - np.angle(sign)) / (2 * np.pi)             # equal modulo 2 pi
    assert abs(k - round(k)) < 1e-10


def test_complex_reference_formula_without_cross_terms():
    """exact=False keeps the reference's formula (RGF.H:476-503): it differs from the dense
    inverse on blocks 0..nt-2 when nd > 0 and agrees on the last block and the corner."""
    ns, nt, nd = 5, 4, 3
    A = fixtures.complex_symmetric(ns, nt, nd, seed=2)
    F = oracle.zfactor_blocks(A, ns, nt, nd)
    di = np.diag(np.linalg.inv(A))
    dr = oracle.zselinv(F, exact=False)
    assert np.abs(dr[(nt - 1) * ns:] - di[(nt - 1) * ns:]).max() <= 1e-12 * np.abs(di).max()
    assert np.abs(dr[:(nt - 1) * ns] - di[:(nt - 1) * ns]).max() > 1e-6
    # nd = 0: both recursions are exact
    A0 = fixtures.complex_symmetric(ns, nt, 0, seed=2)
    F0 = oracle.zfactor_blocks(A0, ns, nt, 0)
    assert np.abs(oracle.zselinv(F0, exact=False) - np.diag(np.linalg.inv(A0))).max() <= 1e-12


def test_real_valued_complex_equals_real_oracle():
    ns, nt, nd = 6, 3, 2
    mf = oracle.generate(4, ns, nt, nd)
    A = oracle.mf_to_dense(mf, ns, nt, nd)
    L = oracle.factor(mf, ns, nt, nd)
    F = oracle.zfactor_blocks(A.astype(np.complex128), ns, nt, nd)
    assert oracle.zlogdet(F) == pytest.approx(oracle.logdet(L, ns, nt, nd), rel=1e-14)
    b = fixtures.rhs_formula(A.shape[0], 2)
    assert np.abs(oracle.zsolve(F, b) - oracle.solve(L, ns, nt, nd, b)).max() <= 1e-13
    assert np.abs(oracle.zselinv(F) - oracle.selinv(L, ns, nt, nd, exact=True)).max() <= 1e-14


def test_complex_csc_roundtrip():
    ns, nt, nd = 4, 3, 2
    A = fixtures.complex_symmetric(ns, nt, nd, seed=9)
    ia, ja, a = fixtures.complex_csc_lower(A, ns, nt, nd)
    R = np.zeros_like(A)
    for c in range(A.shape[0]):
        for k in range(int(ia[c]), int(ia[c + 1])):
            R[int(ja[k]), c] = a[k]
    R = np.tril(R) + np.tril(R, -1).T
    assert np.array_equal(R, A)


def test_planted_conjugation_is_caught():
    """Hermitian algebra (conjugate transposes) in place of the transposes must fail."""
    ns, nt, nd = 5, 3, 2
    A = fixtures.complex_symmetric(ns, nt, nd, seed=5)
    F = oracle.zfactor_blocks(A, ns, nt, nd)
    bad = dict(F)
    bad["B"] = [B.conj() for B in F["B"]]
    L = oracle.zblocks_to_dense(bad)
    assert np.abs(L @ L.T - A).max() > 1e-3
    b = np.ones(A.shape[0], dtype=np.complex128)
    assert np.abs(A @ oracle.zsolve(bad, b) - b).max() > 1e-6
