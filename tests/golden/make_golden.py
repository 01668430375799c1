"""Writes tests/golden/goldens.json: derived expected outputs for the path.

Run from the repository root: ``python tests/golden/make_golden.py``.
Every number is computed here with dense numpy FP64 (slogdet, inv, solve), which
is independent of both the oracle's LAPACK restatement and the HIP path:

* closed_9x9 / closed_11x11: the formula behind the reference's
  data/A_9_9_ns3_nt3.dat and data/A_11_11_ns3_nt3_nd2.dat (regenerated, not
  copied); when /root/reference is present the script also records that the
  regenerated CSC equals the file value for value, and its SHA-256;
* inla_126: derived numbers of data/A_126_126_ns42_nt3.dat (the file itself is
  not stored); used by the CPU suite when /root/reference is present;
* spde_126 / spde_126_nd2: the formula-generated analogue (tests/fixtures.py);
* the reference's own selinv formula (RGF.H:422-506, no cross terms) evaluated
  densely for the nd > 0 cases ("diag_reference").
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle import oracle  # noqa: E402
import fixtures  # noqa: E402

REF = "/root/reference/data"


def entry(A, ns, nt, nd, extra=None):
    n = A.shape[0]
    b = fixtures.rhs_formula(n, 2)
    e = {
        "ns": ns, "nt": nt, "nd": nd, "n": n,
        "logdet": float(np.linalg.slogdet(A)[1]),
        "diag_inv": np.diag(np.linalg.inv(A)).tolist(),
        "diag_reference": oracle.reference_selinv_dense(A, ns, nt, nd).tolist(),
        "x_rhs_formula_2": np.linalg.solve(A, b).reshape(-1, order="F").tolist(),
        "cond": float(np.linalg.cond(A)),
    }
    if extra:
        e.update(extra)
    return e


def main():
    out = {}
    for name, (ns, nt, nd), fn in [("closed_9x9", (3, 3, 0), "A_9_9_ns3_nt3.dat"),
                                  ("closed_11x11", (3, 3, 2), "A_11_11_ns3_nt3_nd2.dat")]:
        ia, ja, a = oracle.closed_form_csc(ns, nt, nd)
        A = oracle.csc_to_dense(ia, ja, a, ns * nt + nd)
        extra = {"nnz": int(ia[-1])}
        path = os.path.join(REF, fn)
        if os.path.exists(path):
            n2, ia2, ja2, a2 = oracle.read_dat(path)
            extra["matches_reference_file"] = bool(np.array_equal(ia, ia2) and np.array_equal(ja, ja2)
                                                   and np.array_equal(a, a2))
            extra["reference_file_sha256"] = hashlib.sha256(open(path, "rb").read()).hexdigest()
        out[name] = entry(A, ns, nt, nd, extra)
    path = os.path.join(REF, "A_126_126_ns42_nt3.dat")
    if os.path.exists(path):
        n, ia, ja, a = oracle.read_dat(path)
        A = oracle.csc_to_dense(ia, ja, a, n)
        out["inla_126"] = entry(A, 42, 3, 0, {"nnz": int(ia[-1]),
                                              "reference_file_sha256": hashlib.sha256(open(path, "rb").read()).hexdigest()})
    for name, nd in [("spde_126", 0), ("spde_126_nd2", 2)]:
        A, ns, nt, nd = fixtures.spde_like(nd=nd)
        out[name] = entry(A, ns, nt, nd)
    # host generator pins (definition of the synthetic workloads)
    gens = {}
    for (ns, nt, nd, seed) in [(4, 3, 2, 1), (42, 3, 0, 0)]:
        mf = oracle.generate(seed, ns, nt, nd)
        A = oracle.mf_to_dense(mf, ns, nt, nd)
        gens[f"gen_{ns}_{nt}_{nd}_s{seed}"] = {
            "ns": ns, "nt": nt, "nd": nd, "seed": seed,
            "first_normals": [oracle.gen_normal(seed, 0, 1, 0, e) for e in range(4)],
            "logdet": float(np.linalg.slogdet(A)[1]),
            "trace": float(np.trace(A)),
        }
    out["generator"] = gens
    with open(os.path.join(fixtures.GOLDEN_DIR, "goldens.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
