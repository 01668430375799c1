"""RGF<std::complex<double>> on the HIP path (through the C ABI) against the complex oracle
restatement and numpy's dense complex solve / inverse / determinant: complex-symmetric
block-arrow matrices of several shapes, real-valued input against the FP64 real path,
residual norms, zero pivots, and a C++ caller of RGF<CPX> built against RGF.hpp.
Bar: x and diag(A^-1) to 1e-10 relative, logdet real part to 1e-11 and imaginary part
equal modulo 2 pi."""
import os
import subprocess
import sys

import numpy as np
import pytest

import rgf_amd
import fixtures
from oracle import oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


SHAPES = [(5, 4, 3), (42, 3, 0), (65, 4, 5), (130, 3, 8), (64, 1, 4), (100, 1, 0), (1, 5, 1), (200, 5, 6),
          (129, 2, 0)]


@pytest.mark.parametrize("ns,nt,nd", SHAPES)
def test_complex_vs_oracle_and_dense(gpu, ns, nt, nd):
    say(f"complex ns={ns} nt={nt} nd={nd}")
    A = fixtures.complex_symmetric(ns, nt, nd, seed=ns + nt + nd)
    n = A.shape[0]
    ia, ja, a = fixtures.complex_csc_lower(A, ns, nt, nd)
    rng = np.random.default_rng(n)
    b = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    with rgf_amd.ZRGF(ia, ja, a, ns, nt, nd) as z:
        z.factorize()
        ld = z.logDet()
        x = z.solve(None, b, 3)
        d = z.RGFdiag()
        x2 = z.solve(None, b[:, :1].copy(), 1)          # the factor stays valid after RGFdiag
        r = z.residualNormNormalized(x[:, 0].copy(), b[:, 0].copy())
    F = oracle.zfactor_blocks(A, ns, nt, nd)
    x0 = oracle.zsolve(F, b)
    d0 = oracle.zselinv(F, exact=True)
    ld0 = oracle.zlogdet(F)
    assert np.abs(x - x0).max() <= 1e-10 * np.abs(x0).max()
    assert np.abs(x2[:, 0] - x0[:, 0]).max() <= 1e-10 * np.abs(x0).max()
    assert np.abs(x - np.linalg.solve(A, b)).max() <= 1e-10 * np.abs(x0).max()
    di = np.diag(np.linalg.inv(A))
    assert np.abs(d - d0).max() <= 1e-10 * np.abs(d0).max()
    assert np.abs(d - di).max() <= 1e-10 * np.abs(di).max()
    assert abs(ld.real - ld0.real) <= 1e-11 * abs(ld0.real)
    # every log L_jj is a principal logarithm, so the imaginary parts agree exactly, not
    # only modulo 2 pi
    assert abs(ld.imag - ld0.imag) <= 1e-9 * max(1.0, abs(ld0.imag))
    sign, logabs = np.linalg.slogdet(A)
    k = (ld.imag - np.angle(sign)) / (2 * np.pi)
    assert abs(ld.real - logabs) <= 1e-11 * abs(logabs) and abs(k - round(k)) < 1e-8
    assert r < 1e-13


def test_complex_reference_formula(gpu):
    ns, nt, nd = 8, 5, 3
    say("complex reference selinv formula")
    A = fixtures.complex_symmetric(ns, nt, nd, seed=12)
    ia, ja, a = fixtures.complex_csc_lower(A, ns, nt, nd)
    opts = rgf_amd.default_options(selinv_mode=rgf_amd.RGF_SELINV_REFERENCE)
    with rgf_amd.ZRGF(ia, ja, a, ns, nt, nd, opts) as z:
        d = z.RGFdiag()
    d0 = oracle.zselinv(oracle.zfactor_blocks(A, ns, nt, nd), exact=False)
    assert np.abs(d - d0).max() <= 1e-10 * np.abs(d0).max()


@pytest.mark.parametrize("ns,nt,nd", [(70, 4, 3), (33, 3, 0)])
def test_real_valued_complex_equals_real_path(gpu, ns, nt, nd):
    say(f"real-valued complex ns={ns} nt={nt} nd={nd}")
    mf = oracle.generate(8, ns, nt, nd)
    A = oracle.mf_to_dense(mf, ns, nt, nd)
    ia, ja, av = oracle.dense_to_csc_lower(A, ns, nt, nd)
    b = fixtures.rhs_formula(A.shape[0], 2)
    with rgf_amd.RGF(ia, ja, av, ns, nt, nd) as r, rgf_amd.ZRGF(ia, ja, av.astype(np.complex128), ns, nt, nd) as z:
        ldr, ldz = r.logDet(), z.logDet()
        xr, xz = r.solve(None, b, 2), z.solve(None, b.astype(np.complex128), 2)
        dr, dz = r.RGFdiag(), z.RGFdiag()
    assert ldz.imag == 0.0 and abs(ldz.real - ldr) <= 1e-13 * abs(ldr)
    assert np.abs(xz.imag).max() == 0.0 and np.abs(xz.real - xr).max() <= 1e-12 * np.abs(xr).max()
    assert np.abs(dz.imag).max() == 0.0 and np.abs(dz.real / dr - 1).max() <= 1e-12


def test_complex_residual_norms(gpu):
    ns, nt, nd = 60, 4, 4
    say("complex residual norms")
    A = fixtures.complex_symmetric(ns, nt, nd, seed=3)
    ia, ja, a = fixtures.complex_csc_lower(A, ns, nt, nd)
    n = A.shape[0]
    b = np.cos(np.arange(n)) + 1j * np.sin(0.3 * np.arange(n))
    with rgf_amd.ZRGF(ia, ja, a, ns, nt, nd) as z:
        x = z.solve(None, b, 1)
        for scale in (1e-3, 1.0):
            xp = x + scale * (np.sin(np.arange(n)) - 0.5j)
            rr = np.linalg.norm(b - A @ xp)
            assert z.residualNorm(xp, b) == pytest.approx(rr, rel=1e-12)
            assert z.residualNormNormalized(xp, b) == pytest.approx(rr / np.linalg.norm(b), rel=1e-12)


def test_complex_zero_pivot(gpu):
    say("complex zero pivot")
    ns, nt, nd = 3, 3, 0
    A = np.diag(np.arange(1, 10, dtype=float) + 0.5j)
    A[4, 4] = 0.0                                    # left out of the CSC: a zero on the device
    A[7, 4] = A[4, 7] = 0.25
    ia, ja, vals = fixtures.complex_csc_lower(A, ns, nt, nd)
    with rgf_amd.ZRGF(ia, ja, vals, ns, nt, nd) as z:
        with pytest.raises(rgf_amd.RGFError) as e:
            z.factorize()
    assert e.value.code == rgf_amd.RGF_ENOTSPD and "column 5" in str(e.value)


def test_complex_cpp_caller(gpu, tmp_path):
    """RGF<CPX> through RGF.hpp: a caller with the reference driver's complex typedef."""
    say("complex C++ caller")
    exe = tmp_path / "dropin_complex"
    r = subprocess.run(["g++", "-std=c++11", "-O1", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "native", "dropin_complex.cpp"), "-o", str(exe),
                        "-L", os.path.join(ROOT, "rgf_amd"), "-lrgf_hip",
                        "-Wl,-rpath," + os.path.join(ROOT, "rgf_amd")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    ns, nt, nd = 6, 3, 2
    A = fixtures.complex_symmetric(ns, nt, nd, seed=21)
    n = A.shape[0]
    ia, ja, a = fixtures.complex_csc_lower(A, ns, nt, nd)
    m = tmp_path / "A.dat"
    with open(m, "w") as f:
        f.write(f"{n}\n{n}\n{len(a)}\n")
        f.write("\n".join(str(int(v)) for v in ia) + "\n")
        f.write("\n".join(str(int(v)) for v in ja) + "\n")
        f.write("\n".join(f"{float(v.real)!r} {float(v.imag)!r}" for v in a) + "\n")
    b = np.arange(1, n + 1) - 0.5j
    bp = tmp_path / "b.dat"
    bp.write_text("\n".join(f"{float(v.real)!r} {float(v.imag)!r}" for v in b) + "\n")
    out = subprocess.run([str(exe), str(ns), str(nt), str(nd), str(m), str(bp)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    F = oracle.zfactor_blocks(A, ns, nt, nd)
    lr, li = map(float, lines[0].split()[1:3])
    ld0 = oracle.zlogdet(F)
    assert abs(lr - ld0.real) <= 1e-12 * abs(ld0.real) and abs(li - ld0.imag) <= 1e-10
    x = np.array([complex(float(l.split()[1]), float(l.split()[2])) for l in lines if l.startswith("x ")])
    d = np.array([complex(float(l.split()[1]), float(l.split()[2])) for l in lines if l.startswith("diag ")])
    assert np.abs(x - np.linalg.solve(A, b)).max() <= 1e-12 * np.abs(x).max()
    assert np.abs(d - np.diag(np.linalg.inv(A))).max() <= 1e-12 * np.abs(d).max()
