"""CPU checks of the C ABI: the library loads, exports every declared symbol, and
rejects bad input before touching a device."""
import os
import re

import numpy as np
import pytest

import rgf_amd
from rgf_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rgf_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for header in ("rgf.h", "rgf_testing.h"):
        names = _declared(header)
        assert names
        for name in names:
            assert hasattr(lib, name), f"{name} declared in {header} but not exported"
    assert set(_declared("rgf.h")) == set(_lib.EXPORTS)
    assert set(_declared("rgf_testing.h")) == set(_lib.TEST_EXPORTS)


def test_defaults():
    o = rgf_amd.default_options()
    assert (o.precision, o.selinv_mode, o.keep_matrix) == (rgf_amd.RGF_FP64, rgf_amd.RGF_SELINV_EXACT, 1)
    assert (o.refine_iters, o.refine_tol, o.graphs) == (10, 1e-14, 0)
    assert (o.keep_splits, o.cholesky, o.keep_w) == (1, rgf_amd.RGF_CHOLESKY_AUTO, 0)
    g = rgf_amd.default_gen_params()
    assert (g.sigma_c, g.sigma_a, g.theta_d, g.theta_c) == (0.1, 0.05, 1.0, 1.0)


def test_bad_pattern_rejected_without_device():
    # (7, 0) lies outside the block-arrow pattern of ns=3, nt=3, nd=0
    ia = np.array([0, 2, 3, 4, 5, 6, 7, 8, 9, 10], dtype=np.uint64)
    ja = np.array([0, 7, 1, 2, 3, 4, 5, 6, 7, 8], dtype=np.uint64)
    with pytest.raises(rgf_amd.RGFError) as e:
        rgf_amd.RGF(ia, ja, np.ones(10), 3, 3, 0)
    assert e.value.code == rgf_amd.RGF_EINVAL
    assert "outside" in str(e.value)


def test_bad_shapes_rejected():
    ia = np.zeros(3 * 3 + 4 + 1, dtype=np.uint64)
    with pytest.raises(rgf_amd.RGFError) as e:
        rgf_amd.RGF(ia, np.zeros(0, np.uint64), np.zeros(0), 3, 3, 4)   # nd > ns
    assert e.value.code == rgf_amd.RGF_EINVAL
    with pytest.raises(rgf_amd.RGFError):
        rgf_amd.RGF(np.zeros(5, np.uint64), np.zeros(0, np.uint64), np.zeros(0), 3, 3, 0)   # wrong ia length


def test_bad_cholesky_schedule_rejected_without_device():
    ia = np.arange(3 + 1, dtype=np.uint64)   # the 3 x 3 identity, ns = 3, nt = 1
    ja = np.arange(3, dtype=np.uint64)
    for bad in (-1, 3):
        with pytest.raises(rgf_amd.RGFError) as e:
            rgf_amd.RGF(ia, ja, np.ones(3), 3, 1, 0, options=rgf_amd.default_options(cholesky=bad))
        assert e.value.code == rgf_amd.RGF_EINVAL and "cholesky schedule" in str(e.value)
    for bad in (-1, 2):
        with pytest.raises(rgf_amd.RGFError) as e:
            rgf_amd.RGF(ia, ja, np.ones(3), 3, 1, 0, options=rgf_amd.default_options(keep_w=bad))
        assert e.value.code == rgf_amd.RGF_EINVAL and "keep_w" in str(e.value)


def test_canonical_flops_closed_form():
    F, S = rgf_amd.canonical_flops(2865, 100, 6)
    assert F + S == pytest.approx(1.636e13, rel=2e-3)
    F1, S1 = rgf_amd.canonical_flops(1000, 50, 4)
    assert F1 + S1 == pytest.approx(3.456e11, rel=2e-3)
    assert rgf_amd.mf_size(42, 3, 0) == 42 * 42 * 5


def test_footprint_bytes_host_only():
    """Admission footprints come from the shape and options alone (no device call)."""
    keep = rgf_amd.default_options(keep_matrix=1)
    inplace = rgf_amd.default_options(keep_matrix=0)
    mixed = rgf_amd.default_options(keep_matrix=1, precision=rgf_amd.RGF_MIXED_BF16X3)
    f = rgf_amd.footprint_bytes(500, 10, 4, 1, keep)
    assert f == rgf_amd.footprint_bytes(500, 10, 4, 1, None)
    ldp = 2 * 500 + 8
    # the factor and the kept matrix: two copies of nt panels of ldp x ns doubles
    assert f - rgf_amd.footprint_bytes(500, 10, 4, 1, inplace) >= 8 * 10 * ldp * 500
    assert rgf_amd.footprint_bytes(500, 10, 4, 1, mixed) > f
    assert rgf_amd.footprint_bytes(500, 10, 4, 3, keep) > 2 * f
    # the task-graph Cholesky (default for few matrices) holds L_ii^-1 tiles as they form
    tiles = rgf_amd.default_options(keep_matrix=1, cholesky=rgf_amd.RGF_CHOLESKY_TILE_LAUNCHES)
    assert f - rgf_amd.footprint_bytes(500, 10, 4, 1, tiles) == 8 * 512 * 448
    assert rgf_amd.footprint_bytes(500, 10, 4, 8, keep) == rgf_amd.footprint_bytes(500, 10, 4, 8, tiles)
    # the kept W_i of the selected inversion: nt blocks of r8(ns + nd) x ns doubles per matrix
    kw = rgf_amd.default_options(keep_matrix=1, keep_w=1)
    assert rgf_amd.footprint_bytes(500, 10, 4, 2, kw) - rgf_amd.footprint_bytes(500, 10, 4, 2, keep) == 8 * 2 * 10 * 504 * 500
    mixed_kw = rgf_amd.default_options(keep_matrix=1, precision=rgf_amd.RGF_MIXED_BF16X3, keep_w=1)
    assert rgf_amd.footprint_bytes(500, 10, 4, 1, mixed_kw) == rgf_amd.footprint_bytes(500, 10, 4, 1, mixed)
    assert rgf_amd.footprint_bytes(0, 10, 4, 1, keep) == 0
    assert rgf_amd.footprint_bytes(5, 10, 6, 1, keep) == 0     # nd > ns


def test_tuning_switches_validate_without_device():
    """rgf_debug_set_tuning returns the previous value for a known key and value and -1
    (changing nothing) otherwise; it touches no device."""
    lib = _lib.load()
    for key, good, bad in ((1, (32, 64), (16, 48)), (3, (0, 1), (2, -1)), (4, (0, 1), (2, 7)), (8, (0, 1), (2, 3)),
                           (9, (0, 1), (2, -1)), (11, (0, 1), (2, 3)),
                           (12, (0, 1), (2, 9))):
        cur = lib.rgf_debug_set_tuning(key, good[0])
        assert cur in good
        for v in bad:
            assert lib.rgf_debug_set_tuning(key, v) == -1
        assert lib.rgf_debug_set_tuning(key, good[1]) == good[0]
        assert lib.rgf_debug_set_tuning(key, cur) == good[1]
    for key, good, bad in ((16, (0, 1), (2, -1)), (21, (1, 0), (2, -1)), (20, (2, 4), (17, -1)),
                           (22, (512, 2048), (63, 2049))):
        cur = lib.rgf_debug_set_tuning(key, good[1])
        assert cur in good and lib.rgf_debug_set_tuning(key, bad[0]) == -1
        assert lib.rgf_debug_set_tuning(key, cur) == good[1]
    for key in (0, 2, 5, 99):
        assert lib.rgf_debug_set_tuning(key, 1) == -1
    cur = lib.rgf_debug_set_tuning(13, 0)   # threshold of the 32 x 32 FP64 tiles: any value >= 0
    assert cur >= 0 and lib.rgf_debug_set_tuning(13, -1) == -1
    assert lib.rgf_debug_set_tuning(13, cur) == 0
    cur = lib.rgf_debug_set_tuning(14, 0)   # threshold of the 256 x 128 split-BF16 tiles: any value >= 0
    assert cur >= 0 and lib.rgf_debug_set_tuning(14, -1) == -1
    assert lib.rgf_debug_set_tuning(14, cur) == 0


def test_autotune_cache_without_device(tmp_path):
    """The tile-shape cache: mode validation, save / load round trip, malformed files
    rejected with the cache unchanged (no device call)."""
    prev = rgf_amd.set_autotune(0)
    try:
        with pytest.raises(rgf_amd.RGFError):
            rgf_amd.set_autotune(3)
        assert rgf_amd.set_autotune(1) == 0
        rgf_amd.autotune_clear()
        assert rgf_amd.autotune_entries() == 0
        p = tmp_path / "tune.txt"
        # signature: kind, ta, tb, launches' products, then m, n, k, flags, entries per product
        p.write_text("rgf-autotune 1 gfx950\n"
                     "9 0 0 1 1 2871 2865 2871 0 1 3 0.512\n"
                     "14 1 1 0 2 2871 2871 2865 1 8 2871 2865 2865 16 8 2 1.5\n")
        rgf_amd.autotune_load(str(p))
        assert rgf_amd.autotune_entries() == 2
        q = tmp_path / "again.txt"
        rgf_amd.autotune_save(str(q))
        rgf_amd.autotune_clear()
        rgf_amd.autotune_load(str(q))
        assert rgf_amd.autotune_entries() == 2
        assert sorted(q.read_text().splitlines()[1:]) == sorted(
            line.replace("0.512", "0.512000").replace("1.5", "1.500000") for line in p.read_text().splitlines()[1:])
        for bad in ("rgf-autotune 1 gfx950\n9 0 0 1 1 10 10 10 0 1 7 0.5\n",      # no FP64 shape 7
                    "rgf-autotune 1 gfx950\n9 1 0 1 1 10 10 10 0 1 3 0.5\n",      # no split-BF16 shape 3
                    "rgf-autotune 1 gfx950\n9 0 0 1 2 10 10 10 0 1 0 0.5\n",      # length != 4 + 5 products
                    "rgf-autotune 1 gfx950\n9 0 0 1 1 10 10\n",                    # truncated
                    "something else\n"):
            p.write_text(bad)
            with pytest.raises(rgf_amd.RGFError):
                rgf_amd.autotune_load(str(p))
            assert rgf_amd.autotune_entries() == 2
        with pytest.raises(rgf_amd.RGFError):
            rgf_amd.autotune_load(str(tmp_path / "missing.txt"))
    finally:
        rgf_amd.autotune_clear()
        rgf_amd.set_autotune(prev)


def test_group_rejects_bad_arguments_without_device():
    lib = _lib.load()
    import ctypes as C
    g = C.c_void_p()
    assert lib.rgf_group_create(None, 0, C.byref(g)) == rgf_amd.RGF_EINVAL
    arr = (C.c_void_p * 2)(None, None)
    assert lib.rgf_group_create(arr, 2, C.byref(g)) == rgf_amd.RGF_EINVAL
    assert b"null" in lib.rgf_last_error(None)
    assert lib.rgf_group_factorize(None) == rgf_amd.RGF_EINVAL
    assert lib.rgf_group_destroy(None) == rgf_amd.RGF_OK
