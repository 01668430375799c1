"""The batched and serving workloads at their real shapes, on the HIP path.

c3: 64 matrices of (ns 1500, nt 60, nd 5) with per-instance theta in one batched handle,
FP64 and mixed: sampled instances bitwise equal to single-matrix handles of the same
seed, id and Cholesky schedule (a batch this large takes the tile-launch schedule; a
single matrix by default the task graph, which agrees to rounding), one instance against
the oracle on the same matrix, every mixed logdet within 1e-8 of FP64.
c4: requests drawn from the serving distribution (ns ~ U{300..3000}, nt ~ U{10..120},
nd ~ U{0..8}) through the continuous-batching server, FP64 and mixed: every result
bitwise equal to the same request run alone, every x with a normalized residual
<= 1e-12 against the matrix, and logdet against the oracle where the oracle is quick.
"""
import sys

import numpy as np
import pytest

import rgf_amd
from rgf_amd import serve
from oracle import oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


C3 = (1500, 60, 5, 64)
C3_SEED = 2024
_C3_FP64 = {}   # FP64 logdets of the c3 batch, for the mixed case of the same session


def _theta(seed, mid):
    return 0.5 + 1.5 * oracle.gen_uniform(seed, mid, 0), 0.1 + 0.9 * oracle.gen_uniform(seed, mid, 1)


@pytest.mark.parametrize("prec", [rgf_amd.RGF_FP64, rgf_amd.RGF_MIXED_BF16X3])
def test_c3_batch(gpu, prec):
    ns, nt, nd, count = C3
    say(f"c3 batch of {count} x ({ns}, {nt}, {nd}) precision={prec}")
    opts = rgf_amd.default_options(precision=prec, keep_matrix=0)
    with rgf_amd.RGF.generated(ns, nt, nd, seed=C3_SEED, count=count, theta_mode=1, options=opts) as hb:
        lds = hb.logdet_batch()
        diags = hb.selinv_diag_batch()
        n = hb.n
        b = np.random.default_rng(3).standard_normal(n)
        hb.select(17)
        x17 = hb.solve(None, b, 1)
    say(f"  batch done; logdet range {lds.min():.6e} .. {lds.max():.6e}")
    assert np.isfinite(lds).all() and np.isfinite(diags).all() and (diags > 0).all()
    tiles = rgf_amd.default_options(precision=prec, keep_matrix=0, cholesky=rgf_amd.RGF_CHOLESKY_TILE_LAUNCHES)
    for mid in (0, 17, count - 1):
        td, tc = _theta(C3_SEED, mid)
        with rgf_amd.RGF.generated(ns, nt, nd, seed=C3_SEED, matrix_id=mid, theta_d=td, theta_c=tc,
                                   options=tiles) as h1:
            assert h1.logDet() == lds[mid], mid
            assert np.array_equal(h1.RGFdiag(), diags[mid]), mid
            if mid == 17:
                assert np.array_equal(h1.solve(None, b, 1), x17)
    say("  sampled instances equal single-matrix handles of the same schedule bitwise")
    # the same matrix alone on the default (task-graph) schedule: the same result to rounding
    td, tc = _theta(C3_SEED, 17)
    with rgf_amd.RGF.generated(ns, nt, nd, seed=C3_SEED, matrix_id=17, theta_d=td, theta_c=tc,
                               options=opts) as h1:
        ld_g, d_g = h1.logDet(), h1.RGFdiag()
    tol_d = 1e-12 if prec == rgf_amd.RGF_FP64 else 1e-5
    assert abs(ld_g - lds[17]) <= 1e-13 * abs(lds[17])
    assert np.abs(d_g / diags[17] - 1).max() <= tol_d
    say(f"  task-graph schedule vs batch: logdet rel {abs(ld_g - lds[17]) / abs(lds[17]):.1e}, "
        f"diag rel {np.abs(d_g / diags[17] - 1).max():.1e}")
    # instance 17 against the oracle on the same matrix (host generator = device generator)
    td, tc = _theta(C3_SEED, 17)
    A_mf = oracle.generate(C3_SEED, ns, nt, nd, matrix_id=17, theta_d=td, theta_c=tc)
    L = oracle.factor(A_mf, ns, nt, nd)
    ld0 = oracle.logdet(L, ns, nt, nd)
    x0 = oracle.solve(L, ns, nt, nd, b)
    d0 = oracle.selinv(L, ns, nt, nd, exact=True)
    if prec == rgf_amd.RGF_FP64:
        assert abs(lds[17] - ld0) <= 1e-10 * abs(ld0)
        assert np.abs(x17 - x0).max() <= 1e-9 * np.abs(x0).max()
        assert np.abs(diags[17] / d0 - 1).max() <= 1e-9
        _C3_FP64["logdets"] = lds.copy()
    else:
        assert abs(lds[17] - ld0) <= 1e-8 * abs(ld0)
        # the batch holds factors only (keep_matrix=0: 64 matrices of 2.2 GB), so the mixed
        # solve is not refined against A here; refined mixed solves: c4 test and the parity suite
        assert np.abs(x17 - x0).max() <= 1e-6 * np.abs(x0).max()
        assert np.abs(diags[17] / d0 - 1).max() <= 1e-4
        ref = _C3_FP64.get("logdets")
        if ref is None:   # the FP64 case did not run in this session: compute it
            with rgf_amd.RGF.generated(ns, nt, nd, seed=C3_SEED, count=count, theta_mode=1,
                                       options=rgf_amd.default_options(keep_matrix=0)) as h64:
                ref = h64.logdet_batch()
        ref = np.asarray(ref)
        err = np.abs(lds / ref - 1)
        say(f"  mixed logdet max rel err vs FP64 over {count}: {err.max():.2e}")
        assert err.max() <= 1e-8
    say(f"  oracle on instance 17: logdet rel {abs(lds[17] - ld0) / abs(ld0):.2e}")


C4_SEED = 8   # 40 draws with 4 follow-ups and 4 requests small enough for the oracle


@pytest.mark.parametrize("prec", [rgf_amd.RGF_FP64, rgf_amd.RGF_MIXED_BF16X3])
def test_c4_serving_sample(gpu, prec):
    reqs = serve.poisson_workload(40, rate=400.0, seed=C4_SEED, followup_frac=0.1)
    say(f"c4 serving sample: {len(reqs)} requests precision={prec}")
    ex = serve.GPUExecutor(prec)
    srv = serve.Server(ex, slots=8, budget_bytes=100 << 30)
    stats = srv.run(reqs)
    say(f"  served: {stats}")
    assert stats["completed"] == len(reqs) and stats["failed"] == 0
    assert srv.max_concurrent >= 2
    keep = rgf_amd.default_options(keep_matrix=1)
    checked = 0
    for r in reqs:
        alone = serve.Request(r.rid, 0.0, r.ns, r.nt, r.nd, r.seed)
        h = ex.run(alone, None)
        ex.release(h)
        assert alone.logdet == r.logdet, r.rid
        assert np.array_equal(alone.x, r.x), r.rid
        b = ex.rhs(r)
        with rgf_amd.RGF.generated(r.ns, r.nt, r.nd, seed=r.seed, options=keep) as hk:
            res = hk.residualNormNormalized(r.x.copy(), b)
        assert res <= 1e-12, (r.rid, res)
        if r.ns * r.nt <= 20000 and r.followup_of is None:
            A_mf = oracle.generate(r.seed, r.ns, r.nt, r.nd)
            L = oracle.factor(A_mf, r.ns, r.nt, r.nd)
            ld0 = oracle.logdet(L, r.ns, r.nt, r.nd)
            tol = 1e-10 if prec == rgf_amd.RGF_FP64 else 1e-8
            assert abs(r.logdet - ld0) <= tol * abs(ld0), r.rid
            checked += 1
    say(f"  every request equal alone and in the mix; {checked} logdets against the oracle")
    assert checked >= 3
