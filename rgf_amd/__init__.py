"""rgf_amd -- MI355X-native RGF solver (block-tridiagonal-with-arrowhead SPD matrices).

The compute path is librgf_hip.so (hand-written gfx950 HIP kernels behind the C
ABI of include/rgf.h).  This module is the Python face of the reference's
``template<class T> class RGF`` (rjanalik/RGF, RGF.H:11-80): the class below
keeps the reference's method names and argument meaning so that callers and
tests read like the reference's own driver (main.C:86-121).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import (RGF_CHOLESKY_AUTO, RGF_CHOLESKY_TASK_GRAPH, RGF_CHOLESKY_TILE_LAUNCHES, RGF_EDEVICE,
                   RGF_EINVAL, RGF_ENOMEM, RGF_ENOTSPD, RGF_ESTATE, RGF_FP64, RGF_MIXED_BF16X3, RGF_OK,
                   RGF_SELINV_EXACT, RGF_SELINV_REFERENCE, DeviceBuffer, GenParams, Info, Options, gpu_count)

__all__ = ["RGF", "ZRGF", "Group", "RGFError", "Options", "GenParams", "Info", "DeviceBuffer", "default_options",
           "default_gen_params", "gpu_count", "RGF_FP64", "RGF_MIXED_BF16X3",
           "RGF_SELINV_EXACT", "RGF_SELINV_REFERENCE", "RGF_CHOLESKY_AUTO", "RGF_CHOLESKY_TASK_GRAPH",
           "RGF_CHOLESKY_TILE_LAUNCHES", "RGF_OK", "RGF_EINVAL", "RGF_ENOTSPD",
           "RGF_ENOMEM", "RGF_EDEVICE", "RGF_ESTATE", "mf_size", "canonical_flops"]


class RGFError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_lib.STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def _check(rc: int, h) -> None:
    if rc != RGF_OK:
        msg = _lib.load().rgf_last_error(h)
        raise RGFError(rc, msg.decode() if msg else "")


def default_options(**kw) -> Options:
    o = Options()
    _lib.load().rgf_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def default_gen_params(**kw) -> GenParams:
    g = GenParams()
    _lib.load().rgf_default_gen_params(C.byref(g))
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def mf_size(ns: int, nt: int, nd: int) -> int:
    """Length of the reference's packed MF array (RGF.H:95)."""
    return 2 * (nt - 1) * ns * ns + ns * ns + (ns * nt + nd) * nd


def canonical_flops(ns: int, nt: int, nd: int) -> tuple[float, float]:
    """Canonical algorithmic flops (F_alg, S_alg) of factorize and selected inversion:
    per block potrf ns^3/3, trsm ns^3, syrk ns^3 (factor); trtri + lauum 2ns^3/3,
    trmm ns^3, gemm 2ns^3, gemmt ns^3 (selected inversion); plus the arrow terms."""
    P = ns ** 3 / 3.0
    Pd = nd ** 3 / 3.0
    F = nt * P + (nt - 1) * 2 * ns ** 3 + nd * (nt * ns ** 2 + 2 * (nt - 1) * ns ** 2 + nt * nd * ns) + Pd
    S = nt * 2 * P + (nt - 1) * 4 * ns ** 3 + nd * (nt * ns ** 2 + 4 * (nt - 1) * ns ** 2) + 2 * Pd
    return float(F), float(S)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class RGF:
    """RGF<double> of the reference (RGF.H:11-80) on one MI355X.

    ``RGF(ia, ja, a, ns, nt, nd)`` takes the lower-triangle CSC of main.C:50-70
    (0-based, size_t indices); ``RGF.generated(ns, nt, nd, seed=...)`` builds the
    benchmark matrix on the device instead.
    """

    def __init__(self, ia, ja, a, ns: int, nt: int, nd: int, options: Options | None = None, *,
                 _handle=None):
        lib = _lib.load()
        self.ns, self.nt, self.nd = int(ns), int(nt), int(nd)
        self.n = self.ns * self.nt + self.nd
        self._h = C.c_void_p()
        if _handle is not None:
            self._h = _handle
            return
        ia = np.ascontiguousarray(ia, dtype=np.uint64)
        ja = np.ascontiguousarray(ja, dtype=np.uint64)
        a = np.ascontiguousarray(a, dtype=np.float64)
        if ia.size != self.n + 1:
            raise RGFError(RGF_EINVAL, f"ia has {ia.size} entries, expected n+1 = {self.n + 1}")
        szp = C.POINTER(C.c_size_t)
        rc = lib.rgf_create_csc(ia.ctypes.data_as(szp), ja.ctypes.data_as(szp), _dptr(a), self.ns, self.nt,
                                self.nd, C.byref(options) if options is not None else None, C.byref(self._h))
        _check(rc, None)

    @classmethod
    def generated(cls, ns: int, nt: int, nd: int, seed: int = 0, matrix_id: int = 0,
                  options: Options | None = None, count: int = 1, **gen) -> "RGF":
        """Matrix (or `count` matrices of one structure, ids matrix_id + b) generated on
        the device; gen keys: sigma_c, sigma_a, theta_d, theta_c, theta_mode."""
        lib = _lib.load()
        g = default_gen_params(seed=seed, matrix_id=matrix_id, **gen)
        h = C.c_void_p()
        opt = C.byref(options) if options is not None else None
        if count == 1:
            rc = lib.rgf_create_generated(ns, nt, nd, C.byref(g), opt, C.byref(h))
        else:
            rc = lib.rgf_create_generated_batch(ns, nt, nd, count, C.byref(g), opt, C.byref(h))
        _check(rc, None)
        return cls(None, None, None, ns, nt, nd, _handle=h)

    @property
    def count(self) -> int:
        return int(_lib.load().rgf_batch_count(self._h))

    def select(self, index: int) -> None:
        _check(_lib.load().rgf_select(self._h, index), self._h)

    def regenerate(self) -> None:
        _check(_lib.load().rgf_regenerate(self._h), self._h)

    def logdet_batch(self) -> np.ndarray:
        out = np.empty(self.count)
        _check(_lib.load().rgf_logdet_batch(self._h, _dptr(out)), self._h)
        return out

    def refine_stats(self) -> tuple[int, float]:
        """(steps, max relative residual) of the last mixed-precision solve's refinement."""
        it, res = C.c_int(), C.c_double()
        _check(_lib.load().rgf_refine_stats(self._h, C.byref(it), C.byref(res)), self._h)
        return it.value, res.value

    def selinv_diag_batch(self) -> np.ndarray:
        out = np.empty(self.count * self.n)
        _check(_lib.load().rgf_selinv_diag_batch(self._h, _dptr(out)), self._h)
        return out.reshape(self.count, self.n)

    # ------------------------------------------------------------ reference API
    def factorize(self) -> None:                      # RGF.H:510-546
        _check(_lib.load().rgf_factorize(self._h), self._h)

    def solve(self, x, b=None, nrhs: int | None = None):
        """solve(b, nrhs) in place (RGF.H:550-554) or solve(x, b, nrhs) (RGF.H:558-578).
        Arrays are column-major n x nrhs (a 2-D numpy array is taken as n x nrhs)."""
        if b is None or isinstance(b, (int, np.integer)):
            if isinstance(b, (int, np.integer)):
                nrhs = int(b)
            b = x
        bb = np.asfortranarray(b, dtype=np.float64)
        if nrhs is None:
            nrhs = 1 if bb.ndim == 1 else bb.shape[1]
        if bb.size != self.n * nrhs:
            raise RGFError(RGF_EINVAL, f"rhs has {bb.size} values, expected n*nrhs = {self.n * nrhs}")
        out = np.empty_like(bb)
        _check(_lib.load().rgf_solve(self._h, _dptr(out.reshape(-1, order="F")), _dptr(bb.reshape(-1, order="F")),
                                     nrhs), self._h)
        if x is not None and isinstance(x, np.ndarray) and x.flags.writeable and x.size == out.size:
            x.reshape(-1, order="F")[:] = out.reshape(-1, order="F")
        return out

    def RGFdiag(self, diag: np.ndarray | None = None) -> np.ndarray:   # RGF.H:582-612
        out = np.empty(self.n) if diag is None else diag
        assert out.dtype == np.float64 and out.flags.c_contiguous and out.size == self.n
        _check(_lib.load().rgf_selinv_diag(self._h, _dptr(out)), self._h)
        return out

    def logDet(self) -> float:                        # RGF.H:616-627
        v = C.c_double()
        _check(_lib.load().rgf_logdet(self._h, C.byref(v)), self._h)
        return v.value

    def residualNorm(self, x, b) -> float:            # RGF.H:631-655
        xx = np.ascontiguousarray(x, dtype=np.float64)
        bb = np.ascontiguousarray(b, dtype=np.float64)
        v = C.c_double()
        _check(_lib.load().rgf_residual_norm(self._h, _dptr(xx), _dptr(bb), C.byref(v)), self._h)
        return v.value

    def residualNormNormalized(self, x, b) -> float:  # RGF.H:659-663
        xx = np.ascontiguousarray(x, dtype=np.float64)
        bb = np.ascontiguousarray(b, dtype=np.float64)
        v = C.c_double()
        _check(_lib.load().rgf_residual_norm_normalized(self._h, _dptr(xx), _dptr(bb), C.byref(v)), self._h)
        return v.value

    # ------------------------------------------------------------ extensions
    def solve_device(self, x: DeviceBuffer, b: DeviceBuffer, nrhs: int) -> None:
        _check(_lib.load().rgf_solve_device(self._h, x.ptr, b.ptr, nrhs), self._h)

    def selinv_diag_device(self, diag: DeviceBuffer) -> None:
        _check(_lib.load().rgf_selinv_diag_device(self._h, diag.ptr), self._h)

    def selinv_blocks(self) -> dict:
        """Blocks of A^-1 on the block pattern of A for the selected matrix (exact selected
        inversion): {"diag": (nt, ns, ns) Sigma_ii, "sub": (nt-1, ns, ns) Sigma_{i+1,i},
        "arrow": (nt, nd, ns) Sigma_{d,i}, "corner": (nd, nd) Sigma_dd}."""
        ns, nt, nd = self.ns, self.nt, self.nd
        sizes = {"diag": nt * ns * ns, "sub": (nt - 1) * ns * ns, "arrow": nt * nd * ns, "corner": nd * nd}
        bufs = {k: DeviceBuffer(v) if v > 0 else None for k, v in sizes.items()}
        try:
            ptr = [bufs[k].ptr if bufs[k] is not None else None for k in ("diag", "sub", "arrow", "corner")]
            _check(_lib.load().rgf_selinv_blocks_device(self._h, *ptr), self._h)
            self.synchronize()
            out = {}
            shapes = {"diag": (nt, ns, ns), "sub": (nt - 1, ns, ns), "arrow": (nt, nd, ns), "corner": (1, nd, nd)}
            for k, (nb, r, c) in shapes.items():
                if bufs[k] is None:
                    out[k] = np.zeros((nb, r, c))
                else:   # column-major blocks
                    out[k] = bufs[k].download().reshape(nb, c, r).transpose(0, 2, 1).copy()
            out["corner"] = out["corner"][0]
            return out
        finally:
            for b in bufs.values():
                if b is not None:
                    b.free()

    def synchronize(self) -> None:
        _check(_lib.load().rgf_synchronize(self._h), self._h)

    def info(self) -> Info:
        i = Info()
        _check(_lib.load().rgf_get_info(self._h, C.byref(i)), self._h)
        return i

    def export_matrix_mf(self) -> np.ndarray:
        out = np.empty(mf_size(self.ns, self.nt, self.nd))
        _check(_lib.load().rgf_export_matrix_mf(self._h, _dptr(out)), self._h)
        return out

    def export_factor_mf(self) -> np.ndarray:
        out = np.empty(mf_size(self.ns, self.nt, self.nd))
        _check(_lib.load().rgf_export_factor_mf(self._h, _dptr(out)), self._h)
        return out

    def close(self) -> None:
        if self._h:
            _lib.load().rgf_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Group:
    """Handles of uneven shapes whose factorizations and selected inversions run as one HIP
    graph launch (one branch per handle, overlapping on the device; captured on the first
    call, replayed afterwards).  Results equal those of the handles run alone, bit for bit.
    The handles must outlive the group."""

    def __init__(self, handles: list["RGF"]):
        self.handles = list(handles)
        arr = (C.c_void_p * len(self.handles))(*[h._h.value for h in self.handles])
        self._g = C.c_void_p()
        _check(_lib.load().rgf_group_create(arr, len(self.handles), C.byref(self._g)), None)

    def _checkg(self, rc: int) -> None:
        _check(rc, None)   # a group call's message is the process-wide one

    def factorize(self) -> None:
        self._checkg(_lib.load().rgf_group_factorize(self._g))

    def selinv_diag_device(self, diags: list[DeviceBuffer]) -> None:
        """diag(A^-1) of every matrix of handle q into diags[q] (count_q * n_q values)."""
        arr = (C.c_void_p * len(diags))(*[d.ptr.value for d in diags])
        self._checkg(_lib.load().rgf_group_selinv_diag_device(self._g, arr))

    def synchronize(self) -> None:
        self._checkg(_lib.load().rgf_group_synchronize(self._g))

    def close(self) -> None:
        if self._g:
            _lib.load().rgf_group_destroy(self._g)
            self._g = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _zcheck(rc: int, h) -> None:
    if rc != RGF_OK:
        msg = _lib.load().rgf_z_last_error(h)
        raise RGFError(rc, msg.decode() if msg else "")


def _zptr(a: np.ndarray):
    assert a.dtype == np.complex128 and (a.flags.c_contiguous or a.flags.f_contiguous)
    return a.ctypes.data_as(C.POINTER(C.c_double))


class ZRGF:
    """RGF<std::complex<double>> of the reference (RGF.H:756-780) on one MI355X, in the
    complex-symmetric algebra of its block updates (A = A^T = L L^T).  Same constructor
    (lower-triangle CSC, complex values) and methods as RGF; logDet returns a complex."""

    def __init__(self, ia, ja, a, ns: int, nt: int, nd: int, options: Options | None = None):
        lib = _lib.load()
        self.ns, self.nt, self.nd = int(ns), int(nt), int(nd)
        self.n = self.ns * self.nt + self.nd
        self._h = C.c_void_p()
        ia = np.ascontiguousarray(ia, dtype=np.uint64)
        ja = np.ascontiguousarray(ja, dtype=np.uint64)
        a = np.ascontiguousarray(a, dtype=np.complex128)
        if ia.size != self.n + 1:
            raise RGFError(RGF_EINVAL, f"ia has {ia.size} entries, expected n+1 = {self.n + 1}")
        szp = C.POINTER(C.c_size_t)
        rc = lib.rgf_z_create_csc(ia.ctypes.data_as(szp), ja.ctypes.data_as(szp), _zptr(a), self.ns, self.nt,
                                  self.nd, C.byref(options) if options is not None else None, C.byref(self._h))
        _zcheck(rc, None)

    def factorize(self) -> None:                      # RGF.H:510-546
        _zcheck(_lib.load().rgf_z_factorize(self._h), self._h)

    def solve(self, x, b=None, nrhs: int | None = None):
        """solve(b, nrhs) or solve(x, b, nrhs) (RGF.H:550-578); column-major n x nrhs."""
        if b is None or isinstance(b, (int, np.integer)):
            if isinstance(b, (int, np.integer)):
                nrhs = int(b)
            b = x
        bb = np.asfortranarray(b, dtype=np.complex128)
        if nrhs is None:
            nrhs = 1 if bb.ndim == 1 else bb.shape[1]
        if bb.size != self.n * nrhs:
            raise RGFError(RGF_EINVAL, f"rhs has {bb.size} values, expected n*nrhs = {self.n * nrhs}")
        out = np.empty_like(bb)
        _zcheck(_lib.load().rgf_z_solve(self._h, _zptr(out), _zptr(bb), nrhs), self._h)
        if x is not None and isinstance(x, np.ndarray) and x.flags.writeable and x.size == out.size:
            x.reshape(-1, order="F")[:] = out.reshape(-1, order="F")
        return out

    def RGFdiag(self, diag: np.ndarray | None = None) -> np.ndarray:   # RGF.H:582-612
        out = np.empty(self.n, dtype=np.complex128) if diag is None else diag
        _zcheck(_lib.load().rgf_z_selinv_diag(self._h, _zptr(out)), self._h)
        return out

    def logDet(self) -> complex:                      # RGF.H:616-627
        v = np.zeros(2)
        _zcheck(_lib.load().rgf_z_logdet(self._h, _dptr(v)), self._h)
        return complex(v[0], v[1])

    def residualNorm(self, x, b) -> float:            # RGF.H:631-655
        v = C.c_double()
        _zcheck(_lib.load().rgf_z_residual_norm(self._h, _zptr(np.ascontiguousarray(x, dtype=np.complex128)),
                                                _zptr(np.ascontiguousarray(b, dtype=np.complex128)), C.byref(v)),
                self._h)
        return v.value

    def residualNormNormalized(self, x, b) -> float:  # RGF.H:659-663
        v = C.c_double()
        _zcheck(_lib.load().rgf_z_residual_norm_normalized(
            self._h, _zptr(np.ascontiguousarray(x, dtype=np.complex128)),
            _zptr(np.ascontiguousarray(b, dtype=np.complex128)), C.byref(v)), self._h)
        return v.value

    def close(self) -> None:
        if self._h:
            _lib.load().rgf_z_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def footprint_bytes(ns: int, nt: int, nd: int, count: int = 1, options: Options | None = None) -> int:
    """Device bytes a handle of this shape and options holds (admission control)."""
    opt = C.byref(options) if options is not None else None
    return int(_lib.load().rgf_footprint_bytes(ns, nt, nd, count, opt))


def set_device_cache_limit(nbytes: int) -> int:
    """Idle device bytes the library's arena keeps per device for reuse; returns the old limit."""
    prev = C.c_size_t()
    _check(_lib.load().rgf_set_device_cache_limit(int(nbytes), C.byref(prev)), None)
    return prev.value


def trim_device_cache(device: int = -1) -> None:
    """Give every wholly free block of the arena back to the device (-1: all devices)."""
    _check(_lib.load().rgf_trim_device_cache(int(device)), None)


def set_autotune(mode: int) -> int:
    """Tile-shape choice of the product launches: 0 size thresholds, 1 cached per-signature
    choices, 2 the same and time every candidate for signatures not cached yet (results
    are bitwise the same whichever shape runs).  Returns the previous mode."""
    prev = C.c_int()
    _check(_lib.load().rgf_set_autotune(int(mode), C.byref(prev)), None)
    return prev.value


def autotune_entries() -> int:
    """Launch signatures with a cached tile-shape choice."""
    return int(_lib.load().rgf_autotune_entries())


def autotune_clear() -> None:
    _check(_lib.load().rgf_autotune_clear(), None)


def autotune_save(path: str) -> None:
    """Write the tile-shape cache to a text file."""
    _check(_lib.load().rgf_autotune_save(os.fsencode(path)), None)


def autotune_load(path: str) -> None:
    """Add the choices of a file written by autotune_save to the cache."""
    _check(_lib.load().rgf_autotune_load(os.fsencode(path)), None)


def kernel_timing_enable(op_class: int) -> None:
    _lib.load().rgf_kernel_timing_enable(op_class)


def kernel_timing_read() -> tuple[float, float, int]:
    ms, fl, n = C.c_double(), C.c_double(), C.c_longlong()
    rc = _lib.load().rgf_kernel_timing_read(C.byref(ms), C.byref(fl), C.byref(n))
    if rc != 0:
        raise RGFError(rc, "kernel timing read failed")
    return ms.value, fl.value, n.value
