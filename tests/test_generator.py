"""The benchmark matrices' definition, pinned independently of the oracle's C generator.

The workloads' text: diagonal blocks B = G G^T/ns + I with G ~ N(0, 1), coupling blocks
~ N(0, 0.1), arrow rows ~ N(0, 0.05), corner H H^T + nd*I with H ~ N(0, 1), diagonal
dominance by adding the row's off-diagonal abs-sum + 1, and per-instance scales
theta_d ~ U(0.5, 2) on the diagonal blocks and theta_c ~ U(0.1, 1) on the couplings.
Here the counter-based stream (Philox4x32-10, Salmon et al. SC'11, then Box-Muller) is
restated in numpy and the dense matrix is built from it with numpy alone; the oracle's
generator (and through the GPU suite the device generator, which is compared with it)
must produce the same matrix, and the draws must have the stated standard deviations.
"""
import numpy as np
import pytest

from oracle import oracle

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)


def philox(ctr, key):
    """Philox4x32-10 on arrays of counters: ctr (N, 4) uint64 holding 32-bit words, key (2,)."""
    c = [ctr[:, i].astype(np.uint64) for i in range(4)]
    k0, k1 = int(key[0]), int(key[1])
    for r in range(10):
        if r:
            k0 = (k0 + W0) & 0xFFFFFFFF
            k1 = (k1 + W1) & 0xFFFFFFFF
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
    return c


def normals(seed, mid, kind, block, e0, count):
    """Element e of a (kind, block) stream: Box-Muller on the 53-bit uniforms of counter
    (e >> 1, kind << 24 | block, matrix id); cos for even e, sin for odd e."""
    e = np.arange(e0, e0 + count, dtype=np.uint64)
    p = e >> np.uint64(1)
    ctr = np.stack([p & MASK, p >> np.uint64(32), np.full_like(p, (kind << 24) | block),
                    np.full_like(p, mid)], axis=1)
    w = philox(ctr, (seed & 0xFFFFFFFF, seed >> 32))
    u1 = ((w[0] >> np.uint64(5)) << np.uint64(26) | (w[1] >> np.uint64(6))).astype(np.float64) / 2.0 ** 53
    u2 = ((w[2] >> np.uint64(5)) << np.uint64(26) | (w[3] >> np.uint64(6))).astype(np.float64) / 2.0 ** 53
    r = np.sqrt(-2.0 * np.log1p(-u1))
    t = 2.0 * np.pi * u2
    return np.where(e & np.uint64(1), r * np.sin(t), r * np.cos(t))


def uniform(seed, mid, idx):
    ctr = np.array([[idx, 0, 5 << 24, mid]], dtype=np.uint64)
    w = philox(ctr, (seed & 0xFFFFFFFF, seed >> 32))
    return float(((w[0] >> np.uint64(5)) << np.uint64(26) | (w[1] >> np.uint64(6)))[0]) / 2.0 ** 53


def dense_generated(seed, ns, nt, nd, mid=0, sigma_c=0.1, sigma_a=0.05, theta_d=1.0, theta_c=1.0):
    n = ns * nt + nd
    A = np.zeros((n, n))
    for t in range(nt):
        G = normals(seed, mid, 1, t, 0, ns * ns).reshape(ns, ns, order="F")
        A[t * ns:(t + 1) * ns, t * ns:(t + 1) * ns] = theta_d * (G @ G.T / ns + np.eye(ns))
        if t < nt - 1:
            C = theta_c * sigma_c * normals(seed, mid, 2, t, 0, ns * ns).reshape(ns, ns, order="F")
            A[(t + 1) * ns:(t + 2) * ns, t * ns:(t + 1) * ns] = C
            A[t * ns:(t + 1) * ns, (t + 1) * ns:(t + 2) * ns] = C.T
        if nd > 0:
            D = sigma_a * normals(seed, mid, 3, 0, t * ns * nd, ns * nd).reshape(nd, ns, order="F")
            A[nt * ns:, t * ns:(t + 1) * ns] = D
            A[t * ns:(t + 1) * ns, nt * ns:] = D.T
    if nd > 0:
        H = normals(seed, mid, 4, 0, 0, nd * nd).reshape(nd, nd, order="F")
        A[nt * ns:, nt * ns:] = H @ H.T + nd * np.eye(nd)
    off = np.abs(A).sum(axis=1) - np.abs(np.diag(A))
    A[np.diag_indices(n)] += off + 1.0
    return A


def test_numpy_stream_matches_oracle_draws():
    for (seed, mid, kind, block) in ((1, 0, 1, 0), (2024, 7, 3, 0), (2 ** 40 + 5, 3, 2, 9)):
        z = normals(seed, mid, kind, block, 11, 9)
        o = [oracle.gen_normal(seed, mid, kind, block, e) for e in range(11, 20)]
        assert np.allclose(z, o, rtol=1e-13, atol=1e-15)
    assert uniform(99, 10, 1) == pytest.approx(oracle.gen_uniform(99, 10, 1), abs=1e-16)


@pytest.mark.parametrize("ns,nt,nd,td,tc", [(5, 3, 2, 1.0, 1.0), (7, 4, 0, 1.7, 0.3), (6, 2, 6, 0.6, 0.9),
                                            (3, 1, 2, 1.0, 1.0)])
def test_oracle_generator_matches_numpy_definition(ns, nt, nd, td, tc):
    A = dense_generated(17, ns, nt, nd, mid=2, theta_d=td, theta_c=tc)
    H = oracle.mf_to_dense(oracle.generate(17, ns, nt, nd, matrix_id=2, theta_d=td, theta_c=tc), ns, nt, nd)
    assert np.abs(A - H).max() <= 1e-13 * np.abs(A).max()


def test_stated_standard_deviations():
    """sigma_c = 0.1 and sigma_a = 0.05 are standard deviations of the coupling and
    arrow entries; the diagonal-block draws G are standard normal."""
    ns, nt, nd = 160, 2, 8
    A = oracle.mf_to_dense(oracle.generate(5, ns, nt, nd), ns, nt, nd)
    C = A[ns:2 * ns, :ns]
    D = A[nt * ns:, :nt * ns]
    assert abs(C.mean()) < 0.003 and C.std() == pytest.approx(0.1, rel=0.02)
    assert abs(D.mean()) < 0.003 and D.std() == pytest.approx(0.05, rel=0.04)
    G = normals(5, 0, 1, 0, 0, ns * ns)
    assert abs(G.mean()) < 0.02 and G.std() == pytest.approx(1.0, rel=0.02)
    # off-diagonal of a diagonal block is (G G^T / ns)_ij, variance 1/ns
    B = A[:ns, :ns] - np.diag(np.diag(A[:ns, :ns]))
    assert B[np.tril_indices(ns, -1)].std() == pytest.approx(1.0 / np.sqrt(ns), rel=0.05)


def test_per_instance_theta_ranges():
    td = np.array([0.5 + 1.5 * uniform(3, m, 0) for m in range(200)])
    tc = np.array([0.1 + 0.9 * uniform(3, m, 1) for m in range(200)])
    assert 0.5 <= td.min() and td.max() < 2.0 and 0.1 <= tc.min() and tc.max() < 1.0
    assert td.mean() == pytest.approx(1.25, abs=0.1) and tc.mean() == pytest.approx(0.55, abs=0.06)
    for m in (0, 5):
        assert 0.5 + 1.5 * oracle.gen_uniform(3, m, 0) == pytest.approx(td[m], abs=1e-15)
