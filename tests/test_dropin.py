"""RGF.hpp as a drop-in for the reference's RGF.H: a caller with the reference driver's
include set and call sequence (tests/native/dropin_main.cpp, after main.C:1-134) compiles
and links against RGF.hpp + librgf_hip.so alone (CPU), and on the GPU prints the
reference driver's fields with the closed-form 11 x 11 fixture's solution and diagonal."""
import os
import subprocess
import time

import numpy as np
import pytest

import fixtures
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "rgf_amd")
SRC = os.path.join(ROOT, "tests", "native", "dropin_main.cpp")


def _build(tmp_path):
    exe = tmp_path / "dropin_main"
    r = subprocess.run(["g++", "-std=c++11", "-O1", "-Wall", "-Wno-unused-result", "-I", INC, SRC, "-o", str(exe),
                        "-L", LIBDIR, "-lrgf_hip", "-Wl,-rpath," + LIBDIR], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("the reference-style caller does not build against RGF.hpp:\n" + r.stderr)
    return exe


def test_reference_style_caller_builds_against_rgf_hpp(tmp_path):
    exe = _build(tmp_path)
    assert exe.exists()


def test_get_time_semantics(tmp_path):
    """get_time(t0) = wall seconds since the epoch minus t0 (Utilities.C:26-32)."""
    src = tmp_path / "gt.cpp"
    src.write_text('#include "RGF.hpp"\n#include <cstdio>\n#include <unistd.h>\n'
                   'int main(){ double a = get_time(0.0); double t = get_time(0.0); usleep(200000);'
                   ' t = get_time(t); std::printf("%.6f %.6f\\n", a, t); return 0; }\n')
    exe = tmp_path / "gt"
    r = subprocess.run(["g++", "-std=c++11", "-I", INC, str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    now = time.time()
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert abs(float(out[0]) - now) < 5.0
    assert 0.19 <= float(out[1]) < 2.0


@pytest.mark.gpu
def test_reference_style_caller_runs(gpu, tmp_path):
    print("drop-in caller on the 11 x 11 closed form", flush=True)
    exe = _build(tmp_path)
    ia, ja, a = oracle.closed_form_csc(3, 3, 2)
    mpath = tmp_path / "A_11.dat"
    oracle.write_dat(str(mpath), 11, ia, ja, a)
    b = fixtures.rhs_formula(11)[:, 0]
    bpath = tmp_path / "b.dat"
    bpath.write_text("\n".join(repr(float(v)) for v in b) + "\n")
    out = subprocess.run([str(exe), "3", "3", "2", str(mpath), str(bpath)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    g = fixtures.goldens()["closed_11x11"]
    ld = [ln for ln in lines if ln.startswith("logdet:")][0]
    assert float(ld.split(":")[1]) == pytest.approx(g["logdet"], abs=1e-6)   # printed with %f
    for key in ("factorize time:", "solve time:", "logdet time:", "RGF time:"):
        v = float([ln for ln in lines if ln.startswith(key)][0].split(":")[1])
        assert 0.0 <= v < 60.0, key
    res = float([ln for ln in lines if ln.startswith("Residual norm normalized:")][0].split(":")[1])
    assert res < 1e-14
    x = np.array([float(ln) for ln in lines if ln.strip() and ln.strip()[0] in "-0123456789" and
                  not ln.startswith(("logdet", "diag"))][-11:])
    xg = np.array(g["x_rhs_formula_2"])[:11]
    assert np.abs(x - xg).max() <= 1e-12 * np.abs(xg).max()
    d = np.array([float(ln.split()[1]) for ln in lines if ln.startswith("diag ")])
    assert np.abs(d / np.array(g["diag_inv"]) - 1).max() <= 1e-12


def test_complex_caller_builds_against_rgf_hpp(tmp_path):
    """RGF<CPX> (RGF.H:756-780) through RGF.hpp: a complex caller compiles and links."""
    exe = tmp_path / "dropin_complex"
    r = subprocess.run(["g++", "-std=c++11", "-O1", "-Wall", "-Wno-unused-result", "-I", INC,
                        os.path.join(ROOT, "tests", "native", "dropin_complex.cpp"), "-o", str(exe),
                        "-L", LIBDIR, "-lrgf_hip", "-Wl,-rpath," + LIBDIR], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
