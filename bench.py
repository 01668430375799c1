"""Benchmark: factorize + logdet + selected inversion (diag(A^-1)) per matrix on one MI355X.

A step is one pass of the hot path over one synthetic matrix resident in HBM:
factorize (block Cholesky with L_ii^-1), logdet, and the Takahashi recursion for
diag(A^-1), output left in device memory.  Default workload: ns=2865, nt=100, nd=6
(n = 286506), FP64, generated on the device from a seed (no dataset).

Prints one JSON line with value = canonical algorithmic TFLOP/s of the step,
the FP64-MFMA roofline of the dominant GEMM class measured with HIP events on the
library's stream, and the CPU baseline (the oracle's LAPACK restatement of the
reference's block algorithm on the host cores, on a bounded sample of blocks).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # name: (ns, nt, nd)
    "c2": (2865, 100, 6),
    "c1": (1000, 50, 4),
    "c0": (42, 3, 0),
}
PEAK_FP64_MFMA = 78.6   # TFLOP/s, MI355X dense FP64 matrix
PEAK_BF16_MFMA = 2500.0  # TFLOP/s, MI355X dense BF16 matrix (no sparsity)
PEAK_HBM = 8000.0       # GB/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-blocks", type=int, default=16, help="blocks of the CPU baseline sample")
    p.add_argument("--no-kernel-timing", action="store_true")
    p.add_argument("--no-mixed", action="store_true")
    p.add_argument("--batch", type=int, default=8, help="matrices per batched launch (0/1: skip)")
    p.add_argument("--serve", action="store_true", help="c4: Poisson request stream through the serving loop")
    p.add_argument("--requests", type=int, default=2000)
    p.add_argument("--rate", type=float, default=20.0, help="requests per second (Poisson)")
    p.add_argument("--slots", type=int, default=16)
    p.add_argument("--hw-queues", type=int, default=16,
                   help="c4: GPU_MAX_HW_QUEUES for the serving process (1..32; HIP's default is 4)")
    p.add_argument("--budget-gb", type=float, default=160.0,
                   help="serving: HBM budget of live handles (the arena keeps up to --cache-gb more)")
    p.add_argument("--cache-gb", type=float, default=115.0,
                   help="serving: idle device memory the library's arena keeps for reuse (GiB)")
    p.add_argument("--serve-cholesky", choices=["tiles", "graph"], default="tiles",
                   help="serving: Cholesky schedule of the request handles (tile launches or task graph)")
    p.add_argument("--c3", action="store_true", help="c3: 64 matrices (1500, 60, 5), per-instance theta")
    p.add_argument("--only-batched", choices=["fp64", "mixed"], default=None,
                   help="run only the batched measurement of one precision (for profiling)")
    p.add_argument("--profile-step", action="store_true",
                   help="only generate the matrix and run one step (for profiler counter runs)")
    p.add_argument("--profile-solve", action="store_true",
                   help="only generate the matrix, factorize it and run three 16-RHS solves (the block solve; "
                        "refined with --profile-precision mixed), for profiler runs")
    p.add_argument("--profile-precision", choices=["fp64", "mixed"], default="fp64",
                   help="precision of the --profile-step / --profile-solve run")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only run a plain device copy of MIB MiB three times (scale for the HBM "
                        "byte counters FETCH_SIZE / WRITE_SIZE of a profiler run)")
    return p.parse_args()


def copy_calibration(mib):
    """A plain device-to-device copy kernel of known size (reads and writes mib MiB per
    launch), to scale the profiler's FETCH_SIZE / WRITE_SIZE readings to bytes."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_launch": n * 8, "launches": 3}), flush=True)


def pmc_traffic():
    """HBM bytes per launch of the NN GEMM class from the committed profiler summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_gemm_nn_c2.json")
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def _cpu_sample(ns, nts, nd, seed):
    """Child-process body: time the oracle on nts blocks (threads from MKL_NUM_THREADS)."""
    from oracle import oracle
    oracle.factor(oracle.generate(seed + 1, ns, 2, nd), ns, 2, nd)   # warm-up: thread pool, pages
    mf = oracle.generate(seed, ns, nts, nd)
    t0 = time.perf_counter()
    L = oracle.factor(mf, ns, nts, nd)
    oracle.logdet(L, ns, nts, nd)
    oracle.selinv(L, ns, nts, nd, exact=True)
    print(json.dumps({"s": time.perf_counter() - t0}), flush=True)


def cpu_baseline(ns, nt, nd, seed, blocks):
    """The oracle (reference block algorithm on host MKL) on `blocks` of the workload's
    nt blocks with the box's threads, and on an eighth of them with one thread (the
    reference links sequential MKL); each sample runs in a child process whose MKL
    thread count is fixed by its environment.  Throughput in the GPU value's unit."""
    import subprocess
    import rgf_amd

    def run(nts, threads):
        env = dict(os.environ, MKL_NUM_THREADS=str(threads), OMP_NUM_THREADS=str(threads))
        code = (f"import sys; sys.path.insert(0, {ROOT!r}); import bench; "
                f"bench._cpu_sample({ns}, {nts}, {nd}, {seed})")
        out = subprocess.run([sys.executable, "-c", code], env=env, check=True, capture_output=True, text=True)
        dt = json.loads(out.stdout.strip().splitlines()[-1])["s"]
        F, S = rgf_amd.canonical_flops(ns, nts, nd)
        return dt, (F + S) / dt / 1e12

    threads = min(16, os.cpu_count() or 1)
    nts = max(2, min(blocks, nt))
    dt, tf = run(nts, threads)
    nts1 = max(2, nts // 8)
    dt1, tf1 = run(nts1, 1)
    # the two smaller workloads in full (every block), same threads
    full = {}
    for name in ("c0", "c1"):
        fns, fnt, fnd = WORKLOADS[name]
        env = dict(os.environ, MKL_NUM_THREADS=str(threads), OMP_NUM_THREADS=str(threads))
        code = (f"import sys; sys.path.insert(0, {ROOT!r}); import bench; "
                f"bench._cpu_sample({fns}, {fnt}, {fnd}, {seed})")
        out = subprocess.run([sys.executable, "-c", code], env=env, check=True, capture_output=True, text=True)
        fdt = json.loads(out.stdout.strip().splitlines()[-1])["s"]
        fF, fS = rgf_amd.canonical_flops(fns, fnt, fnd)
        full[name] = {"ns": fns, "nt": fnt, "nd": fnd, "s_per_matrix": round(fdt, 5),
                      "TFLOPs_canonical": round((fF + fS) / fdt / 1e12, 4)}
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": round(tf, 4),
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": f"factorize+logdet+selinv(exact) of ns={ns} nt={nts} nd={nd} ({nts} of {nt} blocks) "
                  f"on host MKL, {threads} threads, {dt:.2f} s; s/matrix extrapolated to nt={nt}: "
                  f"{dt / nts * nt:.2f}",
        "s_per_matrix_extrapolated": round(dt / nts * nt, 3),
        "one_thread": {"value": round(tf1, 4), "blocks": nts1, "s": round(dt1, 2),
                       "s_per_matrix_extrapolated": round(dt1 / nts1 * nt, 2)},
        "full_workloads": full,
        "host": {"nproc": os.cpu_count(), "model": model},
    }


def time_batch(rgf_amd, ns, nt, nd, count, precision, seed, steps, warmup):
    """Batched handle of `count` c2-shaped matrices (in place, regenerated before every
    step outside the timed region); returns (s per step, logdets, diag rows)."""
    import numpy as np
    opts = rgf_amd.default_options(precision=precision, keep_matrix=0, graphs=1 if warmup > 0 else 0)
    h = rgf_amd.RGF.generated(ns, nt, nd, seed=seed, count=count, options=opts)
    diag = rgf_amd.DeviceBuffer(h.n * count)
    times = []
    for it in range(warmup + steps):
        if it > 0:
            h.regenerate()
        h.synchronize()
        t0 = time.perf_counter()
        h.factorize()
        lds = h.logdet_batch()
        h.selinv_diag_device(diag)
        h.synchronize()
        if it >= warmup:
            times.append(time.perf_counter() - t0)
    d = diag.download().reshape(count, h.n)
    diag.free()
    h.close()
    return float(np.median(times)), lds, d


def run_serve(args):
    # hardware queues of the process (set before the HIP runtime starts): with HIP's default
    # of 4, the slots' streams share queues and their kernels run one after another.  At
    # 40 requests/s (1000 requests, 16 slots) 16 queues served FP64 at 15.2 req/s against
    # 12.9 with 4, mixed at 17.6 against 18.1.
    os.environ["GPU_MAX_HW_QUEUES"] = str(min(max(args.hw_queues, 1), 32))
    import rgf_amd
    from rgf_amd import serve
    out = {"metric": "serving req/s + p50/p99 latency (factorize+logdet+solve per request)",
           "workload": f"c4: {args.requests} requests, Poisson rate {args.rate}/s, ns~U{{300..3000}}, "
                       f"nt~U{{10..120}}, nd~U{{0..8}}, 10% follow-up solves, {args.slots} slots, "
                       f"{os.environ.get('GPU_MAX_HW_QUEUES')} HW queues, Cholesky {args.serve_cholesky}",
           "data": "synthetic (device generator, per-request seed)"}
    import ctypes as C
    lib = rgf_amd._lib.load()

    def cache_stats():
        h, m, f, g = C.c_longlong(), C.c_longlong(), C.c_longlong(), C.c_double()
        lib.rgf_debug_cache_stats(C.byref(h), C.byref(m), C.byref(f), C.byref(g))
        return h.value, m.value, f.value, g.value

    for name, prec in (("fp64", rgf_amd.RGF_FP64), ("mixed_bf16x3", rgf_amd.RGF_MIXED_BF16X3)):
        reqs = serve.poisson_workload(args.requests, args.rate, seed=args.seed, followup_frac=0.1)
        sched = (rgf_amd.RGF_CHOLESKY_TILE_LAUNCHES if args.serve_cholesky == "tiles"
                 else rgf_amd.RGF_CHOLESKY_TASK_GRAPH)
        srv = serve.Server(serve.GPUExecutor(prec, device_cache_gb=args.cache_gb, cholesky=sched), slots=args.slots,
                           budget_bytes=int(args.budget_gb * (1 << 30)),
                           progress=lambda d, t, nm=name: print(f"serve {nm}: {d}/{t} done", file=sys.stderr,
                                                                flush=True))
        c0 = cache_stats()
        stats = srv.run(reqs)
        c1 = cache_stats()
        out[name] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in stats.items()}
        out[name]["device_cache"] = {"reuses": c1[0] - c0[0], "hipMalloc": c1[1] - c0[1], "hipFree": c1[2] - c0[2],
                                     "held_GiB": round(c1[3], 1)}
        print(f"serve {name}: {out[name]}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


def run_c3(args):
    import numpy as np
    import rgf_amd
    ns, nt, nd, count = 1500, 60, 5, 64
    F, S = rgf_amd.canonical_flops(ns, nt, nd)
    res = {"metric": "batched factorize+logdet+selinv s/matrix & TFLOP/s",
           "workload": f"c3: {count} matrices ns={ns} nt={nt} nd={nd}, theta_d~U(0.5,2), theta_c~U(0.1,1)"}
    lds = {}
    diags = {}
    for name, prec in (("fp64", rgf_amd.RGF_FP64), ("mixed_bf16x3", rgf_amd.RGF_MIXED_BF16X3)):
        opts = rgf_amd.default_options(precision=prec, keep_matrix=0, graphs=1 if args.warmup > 0 else 0)
        h = rgf_amd.RGF.generated(ns, nt, nd, seed=args.seed, count=count, theta_mode=1, options=opts)
        dg = rgf_amd.DeviceBuffer(h.n * count)
        times = []
        for it in range(args.warmup + args.steps):
            if it > 0:
                h.regenerate()
            h.synchronize()
            t0 = time.perf_counter()
            h.factorize()
            lds[name] = h.logdet_batch()
            h.selinv_diag_device(dg)
            h.synchronize()
            if it >= args.warmup:
                times.append(time.perf_counter() - t0)
        t = float(np.median(times))
        diags[name] = dg.download().reshape(count, h.n)
        res[name] = {"s_per_batch": round(t, 4), "s_per_matrix": round(t / count, 5),
                     "TFLOPs_canonical": round(count * (F + S) / t / 1e12, 3),
                     "device_GB": round(h.info().device_bytes / 1e9, 2)}
        dg.free()
        h.close()
        print(f"c3 {name}: {res[name]}", file=sys.stderr, flush=True)
    res["mixed_speedup"] = round(res["fp64"]["s_per_batch"] / res["mixed_bf16x3"]["s_per_batch"], 3)
    res["logdet_max_rel_err"] = float(np.max(np.abs(lds["mixed_bf16x3"] / lds["fp64"] - 1)))
    res["diag_max_rel_err"] = float(np.max(np.abs(diags["mixed_bf16x3"] / diags["fp64"] - 1)))
    print(json.dumps(res), flush=True)


def note(msg):
    """Progress of a long run on stderr (the JSON line alone goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the solver runs one "
              "handle per GPU and has no multi-GPU path to launch.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.serve:
        return run_serve(args)
    if args.c3:
        return run_c3(args)
    import numpy as np
    import rgf_amd

    ns, nt, nd = WORKLOADS[args.workload]
    F, S = rgf_amd.canonical_flops(ns, nt, nd)
    if args.only_batched:
        prec = rgf_amd.RGF_FP64 if args.only_batched == "fp64" else rgf_amd.RGF_MIXED_BF16X3
        tb, lds, _ = time_batch(rgf_amd, ns, nt, nd, args.batch, prec, args.seed, args.steps, args.warmup)
        print(json.dumps({"only_batched": args.only_batched, "count": args.batch, "s_per_matrix": tb / args.batch,
                          "TFLOPs_canonical": args.batch * (F + S) / tb / 1e12}), flush=True)
        return
    # The timed FP64 steps launch their kernels one by one: the roofline's HIP events
    # around each launch of the NN class cannot be read back from a graph replay.  The
    # same steps replayed from HIP graphs are timed separately (graph_replay), and the
    # mixed and batched measurements replay graphs (captured in their untimed warm-up).
    graphs = 1 if args.warmup > 0 else 0
    main_graphs = graphs if (args.no_kernel_timing and not (args.profile_step or args.profile_solve)) else 0
    t_gen = time.perf_counter()
    if (args.profile_step or args.profile_solve) and args.profile_precision == "mixed":
        s = rgf_amd.RGF.generated(ns, nt, nd, seed=args.seed,
                                  options=rgf_amd.default_options(precision=rgf_amd.RGF_MIXED_BF16X3,
                                                                  refine_tol=1e-13))
    else:
        # keep_w: the selected inversion's W_i stay for the solves after it (no extra work in the step)
        s = rgf_amd.RGF.generated(ns, nt, nd, seed=args.seed,
                                  options=rgf_amd.default_options(graphs=main_graphs, keep_w=1))
    s.synchronize()
    t_gen = time.perf_counter() - t_gen
    diag = rgf_amd.DeviceBuffer(s.n)

    def step():
        s.factorize()
        ld = s.logDet()
        s.selinv_diag_device(diag)
        return ld

    if args.profile_solve:   # factorize, then 16-RHS solves (the solve after a new factorization)
        s.factorize()
        bsolve = rgf_amd.DeviceBuffer(s.n * 16)
        bsolve.upload(np.random.default_rng(args.seed).standard_normal(s.n * 16))
        xsolve = rgf_amd.DeviceBuffer(s.n * 16)
        times = []
        for _ in range(3):
            s.synchronize()
            t0 = time.perf_counter()
            s.solve_device(xsolve, bsolve, 16)
            s.synchronize()
            times.append(round((time.perf_counter() - t0) * 1e3, 3))
        out = {"precision": args.profile_precision, "solve_16rhs_ms": times}
        if args.profile_precision == "mixed":
            out["refine_steps"], out["refine_rel_residual"] = s.refine_stats()
        print(json.dumps(out), flush=True)
        s.close()
        return

    if args.profile_step:   # one step, NN class timed with HIP events (compare with a profiler)
        rgf_amd.kernel_timing_enable(0 if args.profile_precision == "fp64" else 4)
        t0 = time.perf_counter()
        step()
        s.synchronize()
        dt = time.perf_counter() - t0
        kms, kfl, kn = rgf_amd.kernel_timing_read()
        rgf_amd.kernel_timing_enable(-1)
        print(json.dumps({"precision": args.profile_precision, "profile_step_s": round(dt, 4), "nn_launches": kn,
                          "nn_avg_launch_ms": round(kms / max(kn, 1), 4),
                          "nn_TFLOPs": round(kfl / (kms * 1e-3) / 1e12, 3) if kms > 0 else None}), flush=True)
        s.close()
        return

    # NN class (W = F X, Z = Sb W, trtri merges) timed with HIP events around each launch;
    # enabled before the warm-up so the captured sweep records the events (a replay
    # re-records them: the figures are then those of the last timed step)
    if not args.no_kernel_timing:
        rgf_amd.kernel_timing_enable(0)
    for _ in range(args.warmup):
        step()
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ld = step()
    s.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    kms, kfl, kn = rgf_amd.kernel_timing_read() if not args.no_kernel_timing else (0.0, 0.0, 0)
    rgf_amd.kernel_timing_enable(-1)
    note(f"fp64 steps: {dt:.4f} s per step")
    if main_graphs:   # the timing change invalidates the captured sweeps: capture them again
        step()
        s.synchronize()

    # per-phase times (one extra pass, outside the timed region)
    ph = {}
    t = time.perf_counter(); s.factorize(); s.synchronize(); ph["factorize_ms"] = (time.perf_counter() - t) * 1e3
    t = time.perf_counter(); s.logDet(); ph["logdet_ms"] = (time.perf_counter() - t) * 1e3
    t = time.perf_counter(); s.selinv_diag_device(diag); s.synchronize(); ph["selinv_ms"] = (time.perf_counter() - t) * 1e3
    b = rgf_amd.DeviceBuffer(s.n * 16)
    x = rgf_amd.DeviceBuffer(s.n * 16)
    b.upload(np.random.default_rng(1).standard_normal(s.n * 16))
    s.solve_device(x, b, 1)
    s.synchronize()
    # solves after the selected inversion: through the kept W_i (one dependent product per block
    # and direction); then after a new factorization, the block solve (two per block)
    t = time.perf_counter(); s.solve_device(x, b, 1); s.synchronize(); ph["solve_1rhs_ms"] = (time.perf_counter() - t) * 1e3
    t = time.perf_counter(); s.solve_device(x, b, 16); s.synchronize(); ph["solve_16rhs_ms"] = (time.perf_counter() - t) * 1e3
    info = s.info()
    ph["solve_1rhs_GBps"] = info.bytes_solve_per_rhs / (ph["solve_1rhs_ms"] * 1e-3) / 1e9
    s.factorize()
    s.solve_device(x, b, 1)
    s.synchronize()
    t = time.perf_counter(); s.solve_device(x, b, 1); s.synchronize(); ph["solve_1rhs_block_ms"] = (time.perf_counter() - t) * 1e3
    t = time.perf_counter(); s.solve_device(x, b, 16); s.synchronize(); ph["solve_16rhs_block_ms"] = (time.perf_counter() - t) * 1e3
    ph["solve_1rhs_block_GBps"] = info.bytes_solve_per_rhs / (ph["solve_1rhs_block_ms"] * 1e-3) / 1e9
    ph["factorize_TFLOPs"] = F / (ph["factorize_ms"] * 1e-3) / 1e12
    ph["selinv_TFLOPs"] = S / (ph["selinv_ms"] * 1e-3) / 1e12
    # canonical F and S put the triangular inverses (trtri, nt ns^3/3 + nd^3/3) in the
    # selected inversion; here they run inside factorize (the factor stores L_ii^-1), so
    # the per-phase rates of the work as it runs move that term across
    trtri = nt * ns ** 3 / 3.0 + nd ** 3 / 3.0
    ph["factorize_TFLOPs_as_run"] = (F + trtri) / (ph["factorize_ms"] * 1e-3) / 1e12
    ph["selinv_TFLOPs_as_run"] = (S - trtri) / (ph["selinv_ms"] * 1e-3) / 1e12
    ph = {k: round(v, 4) for k, v in ph.items()}
    ph["solve_path"] = ("solve_*: after the selected inversion, through its kept W_i (keep_w=1); "
                        "solve_*_block_*: after a new factorization, the block solve")
    x64 = x.download()
    fp64_diag = diag.download()
    s.close()
    diag.free()

    # the same FP64 steps replayed from HIP graphs
    graph_replay = None
    if graphs and not main_graphs:
        g = rgf_amd.RGF.generated(ns, nt, nd, seed=args.seed, options=rgf_amd.default_options(graphs=1, keep_w=1))
        gd = rgf_amd.DeviceBuffer(g.n)
        for _ in range(args.warmup):
            g.factorize(); g.logDet(); g.selinv_diag_device(gd)
        g.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.factorize(); g.logDet(); g.selinv_diag_device(gd)
        g.synchronize()
        dtg = (time.perf_counter() - t0) / args.steps
        # solves of the same buffers replay their captured pieces (the first call captures)
        gs = {}
        for nr in (1, 16):
            g.solve_device(x, b, nr)
            g.synchronize()
            t = time.perf_counter(); g.solve_device(x, b, nr); g.synchronize(); gs[nr] = time.perf_counter() - t
        graph_replay = {"s_per_matrix": round(dtg, 4), "TFLOPs_canonical": round((F + S) / dtg / 1e12, 4),
                        "speedup_vs_launches": round(dt / dtg, 3),
                        "solve_1rhs_ms": round(gs[1] * 1e3, 3), "solve_16rhs_ms": round(gs[16] * 1e3, 3),
                        "solve_1rhs_GBps": round(info.bytes_solve_per_rhs / gs[1] / 1e9, 1)}
        g.close(); gd.free()
        note(f"graph replay: {dtg:.4f} s per step")

    # mixed precision, same matrix: split-BF16 products, FP64 potrf / logdet / refinement
    mixed = None
    if not args.no_mixed:
        opts = rgf_amd.default_options(precision=rgf_amd.RGF_MIXED_BF16X3, refine_tol=1e-13, graphs=graphs)
        m = rgf_amd.RGF.generated(ns, nt, nd, seed=args.seed, options=opts)
        md = rgf_amd.DeviceBuffer(m.n)
        for _ in range(args.warmup):
            m.factorize(); m.logDet(); m.selinv_diag_device(md)
        m.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            m.factorize(); ldm = m.logDet(); m.selinv_diag_device(md)
        m.synchronize()
        dtm = (time.perf_counter() - t0) / args.steps
        # split-BF16 products of one extra (untimed) step, HIP events around each launch on
        # the library's stream (an eager handle: events are not read back from graph replays)
        bf = None
        if not args.no_kernel_timing:
            me = rgf_amd.RGF.generated(ns, nt, nd, seed=args.seed,
                                       options=rgf_amd.default_options(precision=rgf_amd.RGF_MIXED_BF16X3))
            me.factorize(); me.selinv_diag_device(md); me.synchronize()
            rgf_amd.kernel_timing_enable(8)
            me.factorize(); me.logDet(); me.selinv_diag_device(md); me.synchronize()
            bms, bfl, bn = rgf_amd.kernel_timing_read()
            rgf_amd.kernel_timing_enable(-1)
            me.close()
            if bms > 0:
                bf = {"kernel": "gemm_bf16x3s_kernel / gemm_bf16x3_kernel (split-BF16 products, all transposes)",
                      "bound": "mfma", "unit": "TFLOP/s",
                      # three bf16 MFMA products (hi*hi, hi*lo, lo*hi) per FP64 product
                      "achieved": round(3 * bfl / (bms * 1e-3) / 1e12, 2), "peak": PEAK_BF16_MFMA,
                      "frac": round(3 * bfl / (bms * 1e-3) / 1e12 / PEAK_BF16_MFMA, 4),
                      "launches": bn, "avg_launch_ms": round(bms / bn, 4),
                      "share_of_step": round(bms * 1e-3 / dtm, 3)}
        xm = rgf_amd.DeviceBuffer(m.n * 16)
        m.solve_device(xm, b, 16)   # with graphs, the first call captures the solve pieces
        m.synchronize()
        t = time.perf_counter(); m.solve_device(xm, b, 16); m.synchronize(); t_solve16 = time.perf_counter() - t
        its, res = m.refine_stats()
        xmh = xm.download()
        dmix = md.download()
        mixed = {
            "s_per_matrix": round(dtm, 4),
            "TFLOPs_canonical": round((F + S) / dtm / 1e12, 4),
            "roofline": {
                # (i) the step's canonical flops (the FP64 path's numerator) against the FP64
                # MFMA peak, directly comparable with the FP64 path's fraction
                "step": {"bound": "mfma", "unit": "TFLOP/s", "achieved": round((F + S) / dtm / 1e12, 3),
                         "peak": PEAK_FP64_MFMA, "frac": round((F + S) / dtm / 1e12 / PEAK_FP64_MFMA, 4)},
                # (ii) the split-BF16 product kernels' executed MFMA work against the BF16 peak
                "split_bf16_products": bf,
            },
            "speedup_vs_fp64": round(dt / dtm, 3),
            "logdet_rel_err_vs_fp64": abs(ldm - ld) / abs(ld),
            "diag_max_rel_err_vs_fp64": float(np.max(np.abs(dmix / fp64_diag - 1))),
            "solve_16rhs_refined_ms": round(t_solve16 * 1e3, 3),
            "refine_steps": its, "refine_rel_residual": res,
            "x_16rhs_max_rel_err_vs_fp64": float(np.max(np.abs(xmh - x64)) / np.max(np.abs(x64))),
        }
        m.close(); md.free(); xm.free()
        note(f"mixed steps: {dtm:.4f} s per step")
    b.free(); x.free()

    # batched: B matrices of the c2 shape per launch, FP64 and mixed
    batched = None
    if args.batch > 1:
        tb64, ld64, d64 = time_batch(rgf_amd, ns, nt, nd, args.batch, rgf_amd.RGF_FP64, args.seed, 2, 1)
        batched = {"count": args.batch, "fp64_s_per_matrix": round(tb64 / args.batch, 4),
                   "fp64_TFLOPs": round(args.batch * (F + S) / tb64 / 1e12, 4)}
        note(f"fp64 batch of {args.batch}: {tb64 / args.batch:.4f} s per matrix")
        if not args.no_mixed:
            tbm, ldm_b, dm_b = time_batch(rgf_amd, ns, nt, nd, args.batch, rgf_amd.RGF_MIXED_BF16X3, args.seed, 2, 1)
            batched.update({
                "mixed_s_per_matrix": round(tbm / args.batch, 4),
                "mixed_TFLOPs_canonical": round(args.batch * (F + S) / tbm / 1e12, 4),
                "mixed_speedup_vs_fp64_batched": round(tb64 / tbm, 3),
                "mixed_speedup_vs_fp64_single": round(dt * args.batch / tbm, 3),
                "logdet_max_rel_err": float(np.max(np.abs(ldm_b - ld64) / np.abs(ld64))),
                "diag_max_rel_err": float(np.max(np.abs(dm_b / d64 - 1))),
            })

    achieved = (kfl / (kms * 1e-3) / 1e12) if kms > 0 else None
    roofline = {
        "bound": "mfma",
        "kernel": "gemm_f64_kernel<false,false> (NN: W = F L^-1, Z = S_below W, trtri merges)",
        "achieved": round(achieved, 3) if achieved else None,
        "peak": PEAK_FP64_MFMA,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP64_MFMA, 4) if achieved else None,
        "traffic": pmc_traffic(),
        "launches": kn,
        "avg_launch_ms": round(kms / kn, 4) if kn else None,
    }
    if batched and "mixed_s_per_matrix" in batched:
        note(f"mixed batch of {args.batch}: {batched['mixed_s_per_matrix']:.4f} s per matrix")
    cpu = None if args.no_cpu_baseline else cpu_baseline(ns, nt, nd, args.seed, args.cpu_blocks)
    if cpu:
        note(f"cpu baseline: {cpu}")
    out = {
        "metric": "factorize+logdet+selinv s/matrix & TFLOP/s; batched req/s + p99, mixed vs FP64",
        "value": round((F + S) / dt / 1e12, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (device Philox generator, seed %d)" % args.seed,
        "config": {"workload": f"{args.workload}: FP64 factorize+logdet+selinv(diag), ns={ns} nt={nt} nd={nd} "
                               f"n={ns * nt + nd}", "s_per_matrix": round(dt, 4),
                   "canonical_flops": F + S, "precision": "fp64", "selinv": "exact",
                   "logdet": ld, "device_bytes": info.device_bytes, "generate_s": round(t_gen, 3)},
        "phases": ph,
        "graph_replay": graph_replay,
        "mixed_bf16x3": mixed,
        "batched": batched,
        # the north-star comparison like for like: mixed batch against the FP64 batch of the
        # same count (and single matrix against single matrix)
        "mixed_vs_fp64": {
            "batched_like_for_like": batched.get("mixed_speedup_vs_fp64_batched") if batched else None,
            "single_like_for_like": mixed["speedup_vs_fp64"] if mixed else None,
            "batch_count": args.batch,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
