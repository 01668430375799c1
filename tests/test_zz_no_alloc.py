"""rgf_solve_device_noalloc (the solve for callers that cannot allocate: buffers in a static
device array) against rgf_solve_device, through the C ABI: bit for bit, on two calls in a
row with the second input smaller, FP64 and refined mixed precision, host-side rejections."""
import sys

import numpy as np
import pytest

import rgf_amd
from rgf_amd import _lib

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _solve_both(h, b, nrhs, in_place=False):
    """(x from rgf_solve_device, x from rgf_solve_device_noalloc, refine stats of each)."""
    lib = _lib.load()
    n = h.n
    bd = rgf_amd.DeviceBuffer(n * nrhs)
    xd = rgf_amd.DeviceBuffer(n * nrhs)
    try:
        bd.upload(b.reshape(-1, order="F"))
        if in_place:
            assert lib.rgf_solve_device(h._h, bd.ptr, bd.ptr, nrhs) == 0, lib.rgf_last_error(h._h)
        else:
            assert lib.rgf_solve_device(h._h, xd.ptr, bd.ptr, nrhs) == 0, lib.rgf_last_error(h._h)
        h.synchronize()
        ref = (bd if in_place else xd).download()
        st_ref = h.refine_stats()
        bd.upload(b.reshape(-1, order="F"))
        xd.upload(np.full(n * nrhs, np.nan))
        if in_place:
            rc = lib.rgf_solve_device_noalloc(h._h, bd.ptr, bd.ptr, nrhs)
        else:
            rc = lib.rgf_solve_device_noalloc(h._h, xd.ptr, bd.ptr, nrhs)
        assert rc == 0, lib.rgf_last_error(h._h)
        h.synchronize()
        got = (bd if in_place else xd).download()
        return ref, got, st_ref, h.refine_stats()
    finally:
        bd.free()
        xd.free()


@pytest.mark.parametrize("prec", [rgf_amd.RGF_FP64, rgf_amd.RGF_MIXED_BF16X3])
def test_noalloc_solve_matches_solve_device(gpu, prec):
    opts = rgf_amd.default_options(precision=prec, keep_matrix=1, refine_tol=1e-14)
    # first call: 20 right-hand sides (two passes of the refined path); second call: a
    # smaller matrix with 3 (the scratch keeps the first call's contents)
    for (ns, nt, nd, nrhs) in ((150, 5, 4, 20), (60, 3, 2, 3)):
        say(f"no-alloc solve prec={prec} ns={ns} nt={nt} nd={nd} nrhs={nrhs}")
        with rgf_amd.RGF.generated(ns, nt, nd, seed=91 + ns, options=opts) as h:
            h.factorize()
            b = np.random.default_rng(ns).standard_normal((h.n, nrhs))
            ref, got, st_ref, st = _solve_both(h, b, nrhs)
            assert np.isfinite(ref).all()
            assert np.array_equal(got, ref)
            assert st == st_ref
            ref, got, st_ref, st = _solve_both(h, b, nrhs, in_place=True)
            assert np.array_equal(got, ref)
            assert st == st_ref
            x = h.solve(None, b, nrhs)   # the handle's own path afterwards is unaffected
            assert np.array_equal(x.reshape(-1, order="F"), ref)


def test_noalloc_solve_inla_size(gpu):
    """ns 1000, nt 50, nd 4 (n = 50 004), 16 right-hand sides in one pass."""
    ns, nt, nd, nrhs = 1000, 50, 4, 16
    say(f"no-alloc solve ns={ns} nt={nt} nd={nd} nrhs={nrhs}")
    with rgf_amd.RGF.generated(ns, nt, nd, seed=5) as h:
        h.factorize()
        b = np.random.default_rng(1).standard_normal((h.n, nrhs))
        ref, got, _, _ = _solve_both(h, b, nrhs)
        assert np.array_equal(got, ref)
        assert h.residualNormNormalized(got[:h.n].copy(), b[:, 0].copy()) < 1e-14


def test_noalloc_solve_rejections(gpu):
    lib = _lib.load()
    say("no-alloc solve rejections")
    with rgf_amd.RGF.generated(40, 3, 2, seed=1) as h:
        bd = rgf_amd.DeviceBuffer(h.n)
        try:
            assert lib.rgf_solve_device_noalloc(h._h, bd.ptr, bd.ptr, 1) == rgf_amd.RGF_ESTATE   # not factorized
            assert b"factorized" in lib.rgf_last_error(h._h)
            assert lib.rgf_solve_device_noalloc(h._h, None, bd.ptr, 1) == rgf_amd.RGF_EINVAL
        finally:
            bd.free()
    # refined mixed precision keeps 16 columns per pass: 3 x 16 x n doubles exceed the
    # 128 MiB scratch at n = 360 000 (checked before any device work)
    opts = rgf_amd.default_options(precision=rgf_amd.RGF_MIXED_BF16X3, keep_matrix=1)
    with rgf_amd.RGF.generated(3000, 120, 0, seed=2, options=opts) as h:
        bd = rgf_amd.DeviceBuffer(h.n * 16)
        xd = rgf_amd.DeviceBuffer(h.n * 16)
        try:
            assert lib.rgf_solve_device_noalloc(h._h, xd.ptr, bd.ptr, 16) == rgf_amd.RGF_EINVAL
            assert b"scratch" in lib.rgf_last_error(h._h)
        finally:
            bd.free()
            xd.free()
