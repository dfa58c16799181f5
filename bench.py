#!/usr/bin/env python3
"""256^3 Jacobi-PCG time-to-1e-10 on one MI355X: mixed precision vs plain FP64,
and SpMV bandwidth against the HBM roofline.

    python bench.py [--steps K] [--warmup W] [--size N] [--format csr|jad|dia]

One step = one full solve of the 7-point Poisson system on an N^3 grid
(diag 6, off-diagonal -1, Dirichlet outside), b = A x_true with
x_true ~ U[0,1) (numpy default_rng(1)), x0 = 0, stopped at
||b - A x|| <= tol * ||b||.  Inputs are resident in HBM before the timed
region starts.  Prints one JSON line on stdout; progress goes to stderr.

`value` is the median mixed-precision solve (FP32 inner iteration, FP64
residual replacement) in the --format form (CSR by default, the reference's
format); the plain FP64 solve of the same form is timed beside it.  The
other two forms of the same block run the same solves (JAD, converted on
the device; DIA, bitwise the CSR solves), with fewer steps.  Then, untimed:
per-launch HIP events around the K1 kernel of one solve of each form and
precision (`roofline*`), the one-subdomain model of the reference's
distributed SpMV (six halo faces, faces packed every SpMV, received values
zero) in the CSR and DIA forms, the 384^3 precision workload (x_true ~
N(0,1), seed 2; mixed vs FP64, CSR), and the CPU side: the oracle's OpenMP
restatement of the reference loop timed on the host cores (`cpu_baseline`),
whose 256^3 and 384^3 FP64 solutions are the oracle the GPU solutions are
compared with.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "256^3 PCG time-to-1e-10: mixed vs plain FP64 GPU; SpMV GB/s as % HBM peak"
HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "hbm_traffic_256.txt")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def face_blocks(N):
    """Halo blocks of the centre rank of a 3x3x3 decomposition with N^3 cells
    per rank, in ascending neighbour rank order (z-, y-, x-, x+, y+, z+).
    Returns [(proc, rows, cols)] with 1-based local ids."""
    idx = np.arange(N * N, dtype=np.int64)
    a, b = idx % N, idx // N
    n2 = N * N

    def cell(x, y, z):
        return x + N * (y + N * z) + 1

    faces = [
        (4, cell(a, b, 0), cell(a, b, N - 1)),        # z-: neighbour's top plane
        (10, cell(a, 0, b), cell(a, N - 1, b)),       # y-
        (12, cell(0, a, b), cell(N - 1, a, b)),       # x-
        (14, cell(N - 1, a, b), cell(0, a, b)),       # x+
        (16, cell(a, N - 1, b), cell(a, 0, b)),       # y+
        (22, cell(a, b, N - 1), cell(a, b, 0)),       # z+
    ]
    out = []
    for proc, rows, cols in faces:
        order = np.argsort(rows, kind="stable")
        out.append((proc, rows[order].astype(np.int32), cols[order].astype(np.int32)))
    del n2
    return out


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (lines "traffic_bytes_per_launch <kernel> = <bytes>"), or None."""
    try:
        with open(PROFILE_TRAFFIC) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch") and "=" in line:
                    key, val = line.split("=", 1)
                    if key.split(None, 1)[1].strip() == kernel:
                        return float(val)
    except (OSError, IndexError, ValueError):
        pass
    return None


def cpu_baseline(N, tol):
    """The oracle's OpenMP restatement of the reference loop (thread 0 idle,
    4-aligned worker chunks) on the host cores: the full FP64 solve of the
    same 256^3 system (x0 = 0, stop at ||r|| <= tol ||b||, the reference's
    absolute |alpha| test kept).  Returns the timing and the solution, the
    FP64 oracle the GPU solutions are compared with."""
    import oracle as O

    threads = _cpu_threads()
    workers = threads - 1
    cpu_model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    t0 = time.time()
    A = O.gen_poisson7(N, N, N)
    xt = np.random.default_rng(1).random(A.n)
    b = O.spmv_csr(A, xt, workers=workers)
    jac = O.set_jacobi(O.csr_diagonal(A))
    log(f"cpu baseline: matrix ready in {time.time() - t0:.1f}s, {threads} threads")
    t0 = time.time()
    r = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=tol * float(np.linalg.norm(b)),
                     iter_max=20000, workers=workers)
    dt = time.time() - t0
    log(f"cpu baseline: {r['iters']} iterations in {dt:.1f}s ({dt / max(r['iters'], 1) * 1e3:.1f} ms/iter)")
    # the reference's own thread count of the CPU-runnable case: 4 threads
    # (3 computing workers), on a bounded sample of iterations
    t0 = time.time()
    r4 = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=0.0, iter_max=60, workers=3)
    dt4 = time.time() - t0
    per4 = dt4 / max(r4["iters"], 1)
    log(f"cpu baseline, 4 threads: {r4['iters']} iterations in {dt4:.1f}s ({per4 * 1e3:.1f} ms/iter)")
    return {
        "value": dt,
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": (f"the full FP64 solve of the {N}^3 system to {tol:g} relative residual, "
                   f"{r['iters']} Jacobi-PCG iterations (oracle/cel_oracle.c, reference op order, "
                   f"{workers} workers + idle master thread, OMP_PROC_BIND=close "
                   f"OMP_PLACES=cores), solve only"),
        "iters": r["iters"],
        "per_iter_ms": dt / max(r["iters"], 1) * 1e3,
        "host_cpu": cpu_model,
        "host_logical_cpus": os.cpu_count(),
        "four_thread_point": {"threads": 4, "workers": 3, "iters": r4["iters"],
                              "per_iter_ms": per4 * 1e3,
                              "extrapolated_solve_s": per4 * r["iters"]},
    }, r["x"]


def batch_bench(cel, steps):
    """64 variable-coefficient 7-point systems (edge ~ U{24..96}, kappa =
    exp(g), g ~ N(0, 0.5^2), b ~ U[0,1)) plus the 128^3 Poisson system with
    16 right-hand sides ~ U[-1,1], FP64 to 1e-9 relative: one tuned batch
    against the same 80 right-hand sides solved one by one (device-resident
    inputs, single-system solver).  Same generator stream as the batch test."""
    rng = np.random.default_rng(3)
    edges = rng.integers(24, 97, size=64)
    mats, bs, nrhs = [], [], []
    for e in edges:
        n = int(e) ** 3
        kappa = np.exp(rng.normal(0.0, 0.5, n))
        b = rng.random(n)
        M = cel.Matrix.diffusion7(int(e), int(e), int(e), kappa)
        M.set_jacobi()
        mats.append(M)
        bs.append(b)
        nrhs.append(1)
    P = cel.Matrix.poisson7(128, 128, 128)
    P.set_jacobi()
    mats.append(P)
    bs.append(rng.uniform(-1.0, 1.0, (16, 128 ** 3)))
    nrhs.append(16)
    batch = cel.Batch(mats, nrhs)
    t0 = time.perf_counter()
    info = batch.tune(bs, iters=32)
    tune_s = time.perf_counter() - t0
    log(f"batch tuned in {tune_s:.1f} s: {info}")
    kw = dict(tol=1e-9, tol_mode=cel.TOL_REL_B, alpha_stop=0, iter_max=20000)
    times = []
    for k in range(max(1, steps)):
        out = batch.solve(bs, **kw)
        times.append(out["total"]["solve_ms"])
        log(f"batch solve {k + 1}: {out['total']['solve_ms']:.1f} ms, max {out['total']['iters']} iterations")
    lanes = [(m, b if k == 1 else b[j]) for m, b, k in zip(mats, bs, nrhs) for j in range(k)]
    seq_ms, seq_iters = 0.0, 0
    dev = []
    for m, b in lanes:
        dev.append((m, cel.DeviceArray.from_host(b)))
    cel.lib().cel_gpu_synchronize()
    t0 = time.perf_counter()
    for m, d_b in dev:
        r = m.pcg(d_b, x0_mode=cel.X0_ZERO, **kw)
        seq_ms += r["solve_ms"]
        seq_iters += r["iters"]
    cel.lib().cel_gpu_synchronize()
    seq_wall = time.perf_counter() - t0
    log(f"one by one: {seq_ms:.1f} ms device, {seq_wall:.2f} s wall, {seq_iters} iterations in total")
    return {"workload": "64 diffusion7 systems (edge 24..96) + poisson7 128^3 x 16 rhs, tol 1e-9",
            "lanes": batch.lanes, "elements": info["elements"], "tuned": info,
            "batch_ms": times, "batch_ms_median": float(np.median(times)),
            "one_by_one_ms": seq_ms, "one_by_one_wall_s": seq_wall,
            "speedup": seq_ms / float(np.median(times)), "tune_s": tune_s}


def batch16_bench(cel, steps):
    """The batch's multi-right-hand-side system alone: 128^3 Poisson with 16
    right-hand sides ~ U[-1,1] (numpy default_rng(3)), FP64 to 1e-9
    relative, one tuned batch; time per iteration and the algorithmic bytes
    of its SpMV (CSR matrix once for the 16 lanes, x and y as 16-lane rows)."""
    P = cel.Matrix.poisson7(128, 128, 128)
    P.set_jacobi()
    bs = [np.random.default_rng(3).uniform(-1.0, 1.0, (16, 128 ** 3))]
    batch = cel.Batch([P], [16])
    info = batch.tune(bs, iters=32)
    kw = dict(tol=1e-9, tol_mode=cel.TOL_REL_B, alpha_stop=0, iter_max=20000)
    times, iters = [], 0
    for k in range(max(1, steps)):
        out = batch.solve(bs, **kw)
        times.append(out["total"]["solve_ms"])
        iters = out["total"]["iters"]
        log(f"128^3 x 16 solve {k + 1}: {times[-1]:.1f} ms, {iters} iterations, "
            f"{times[-1] / iters * 1e3:.1f} us per iteration")
    n = P.n
    spmv_bytes = 12 * P.nnz + 4 * (n + 1) + 2 * 16 * 8 * n
    return {"workload": "poisson7 128^3 x 16 rhs, tol 1e-9", "tuned": info, "solve_ms": times,
            "iters": iters, "us_per_iter": float(np.median(times)) / iters * 1e3,
            "spmv_algorithmic_bytes": spmv_bytes}


def gmres_bench(cel, N, steps, iter_cap):
    """Jacobi-GMRES with the element solver's settings (m = 3, eps 1e-8 on
    the true residual norm, x0 = 1) on the N^3 7-point block, b ~ U[0,1)
    (numpy default_rng(55)), inputs resident in HBM, in the DIA and the CSR
    form.  iter_cap > 0 caps the outer iterations (for profiler runs)."""
    M = cel.Matrix.poisson7(N, N, N)
    M.set_jacobi()
    M.to_dia()
    b = np.random.default_rng(55).random(M.n)
    d_b = cel.DeviceArray.from_host(b)
    cap = iter_cap if iter_cap > 0 else 5000
    out = {"workload": f"poisson7_{N}^3_jacobi_gmres3_eps1e-8", "n": M.n, "nnz": M.nnz}
    for name, f in (("dia", cel.FMT_DIA), ("csr", cel.FMT_CSR)):
        M.gmres(d_b, iter_max=min(cap, 2), fmt=f)  # warm-up (graph capture, first touch)
        runs = [M.gmres(d_b, iter_max=cap, fmt=f) for _ in range(max(1, steps))]
        r = runs[-1]
        res = d_b.to_host() - M.spmv(r["x"].to_host())
        ms = [x["solve_ms"] for x in runs]
        out[name] = {"solve_ms": ms, "iters": r["iters"], "cycles": r["cycles"],
                     "us_per_cycle": float(np.median(ms)) / max(r["cycles"], 1) * 1e3,
                     "final_norm": r["final_norm"], "true_residual_norm": float(np.linalg.norm(res)),
                     "status": r["status"]}
        log(f"gmres {name}: {np.median(ms):.1f} ms, {r['iters']} outer iterations, {r['cycles']} cycles, "
            f"{out[name]['us_per_cycle']:.1f} us per cycle, |r| {r['final_norm']:.3e}")
    return out


def sort_bench(cel, N, steps, warmup, skip_cpu):
    """COO ordering in HBM: the N^3 7-point matrix in shuffled entry order
    (numpy default_rng(7) permutation; the order a Matrix Market file need
    not keep) sorted with the reference's order (CEL_SORT_CEL_COMPAT: row
    sort + the column-loop segments) and with the full (row, col) order,
    arrays resident in HBM, in place.  Each step restores the shuffled input
    by a device copy outside the timed call.  Algorithmic bytes: the entry
    arrays read once and written once, 2 x (4 + 4 + 8) = 32 B per entry."""
    M = cel.Matrix.poisson7(N, N, N)
    xadj, cols, vals = M.get_csr()
    M.destroy()
    rows = np.repeat(np.arange(1, N ** 3 + 1, dtype=np.int32), np.diff(xadj))
    perm = np.random.default_rng(7).permutation(rows.size)
    rows, cols, vals = rows[perm], cols[perm], vals[perm]
    nnz = rows.size
    src = [cel.DeviceArray.from_host(a) for a in (rows, cols, vals)]
    work = [cel.DeviceArray(a.n, a.dtype) for a in src]
    n = N ** 3
    out = {"workload": f"poisson7 {N}^3 COO, shuffled, in place in HBM", "entries": int(nnz)}
    for name, mode in (("cel_compat", cel.SORT_CEL_COMPAT), ("row_col", cel.SORT_ROW_COL)):
        times = []
        for k in range(warmup + steps):
            for w, s_ in zip(work, src):
                w.copy_from(s_)
            cel.lib().cel_gpu_synchronize()
            t0 = time.perf_counter()
            cel.coo_sort(n, *work, mode=mode)
            dt = time.perf_counter() - t0
            if k >= warmup:
                times.append(dt)
        ms = float(np.median(times)) * 1e3
        gbs = 32.0 * nnz / (ms * 1e-3) / 1e9
        log(f"sort {name}: {ms:.2f} ms median of {steps}, {nnz / ms / 1e6:.2f} G entries/s, "
            f"{gbs:.0f} GB/s algorithmic")
        out[name] = {"ms": ms, "ms_all": [t * 1e3 for t in times], "entries_per_s": nnz / (ms * 1e-3),
                     "algorithmic_GBs": gbs, "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}
    # correctness of the last (row, col) sort: the generator's CSR order
    r_s, c_s = work[0].to_host(), work[1].to_host()
    want_rows = np.repeat(np.arange(1, n + 1, dtype=np.int32), np.diff(xadj))
    out["row_col_matches_generator"] = bool(np.array_equal(r_s, want_rows) and
                                            np.array_equal(c_s, cols[np.argsort(perm)]))
    if not skip_cpu:
        out["cpu_baseline"] = sort_cpu_baseline()
        log(f"cpu baseline (oracle sort): {out['cpu_baseline']}")
    return out


def bandwidth_bench(cel):
    """HBM read rate of 2 GiB read as 1-8 equal arrays at once (the SpMV
    kernels read 3-5 arrays together), 1-8 blocks of 256 threads per CU,
    best of 3 means of 5 launches, and of a 1 GiB copy: the measured
    ceilings beside the 8 TB/s nominal peak."""
    nbytes = 2 * 1024 ** 3
    out = {"read_bytes": nbytes, "read_GBs": {}}
    for bpc in (1, 2, 4, 8):
        for k in (1, 2, 3, 4, 5, 8):
            ms = min(cel.stream_read_time(nbytes, streams=k, blocks_per_cu=bpc, reps=5)
                     for _ in range(3))
            out["read_GBs"][f"{k}_arrays_{bpc}_per_cu"] = nbytes / ms * 1e-6
            log(f"read {k} arrays, {bpc} blocks/CU: {nbytes / ms * 1e-6:.0f} GB/s")
    cb = 1024 ** 3
    ms = min(cel.stream_copy_time(cb, reps=5) for _ in range(3))
    out["copy_GBs"] = 2 * cb / ms * 1e-6
    log(f"copy: {out['copy_GBs']:.0f} GB/s (read + write)")
    best = max(out["read_GBs"].values())
    out["read_GBs_best"] = best
    out["read_frac_of_8TBs"] = best / 8000.0
    return out


def sort_cpu_baseline():
    """The oracle's restatement of the reference's COO sort
    (orc_coo_sort_cel: cel_sp_mat_format_sparse_31_sort, one thread, as the
    reference runs it) on a bounded sample: the 160^3 7-point matrix
    (28,518,400 entries) in the same shuffled form, about 15 s of CPU work."""
    import oracle as O

    A = O.gen_poisson7(160, 160, 160)
    rows = np.repeat(np.arange(1, A.n + 1, dtype=np.int32), np.diff(A.xadj))
    perm = np.random.default_rng(7).permutation(rows.size)
    r, c, v = rows[perm], A.cols[perm], A.vals[perm]
    t0 = time.perf_counter()
    O.coo_sort_cel(A.n, r, c, v)
    dt = time.perf_counter() - t0
    return {"value": rows.size / dt, "unit": "entries/s", "cores": 1, "kind": "port",
            "seconds": dt, "sample": f"160^3 poisson7 COO, {rows.size} entries, shuffled"}


def config3(cel, tol, steps, skip_cpu):
    """The precision workload: 384^3 7-point Poisson (n = 56,623,104, nnz =
    395,476,992), x_true ~ N(0,1) (numpy default_rng(2)), b = A x_true, x0 =
    0, to 1e-10: the mixed solve against the plain FP64 solve (CSR form), and
    both against the oracle's FP64 solve of the same system on the host."""
    N = 384
    M = cel.Matrix.poisson7(N, N, N)
    M.set_jacobi()
    xt = np.random.default_rng(2).normal(size=M.n)
    d_xt = cel.DeviceArray.from_host(xt)
    d_b = cel.DeviceArray(M.n)
    M.spmv(d_xt, d_b)
    b = d_b.to_host()
    bnorm = float(np.linalg.norm(b))
    kw = dict(tol=tol, tol_mode=cel.TOL_REL_B, x0_mode=cel.X0_ZERO, iter_max=20000,
              fmt=cel.FMT_CSR, allow_not_converged=False)
    out = {"workload": f"poisson7_{N}^3_x_true_N(0,1)_seed2_tol{tol:g}", "n": M.n, "nnz": M.nnz,
           "format": "csr"}
    runs = {}
    for name, prec in (("mixed", cel.PREC_MIXED), ("fp64", cel.PREC_FP64)):
        M.pcg(d_b, precision=prec, **kw)  # warm-up (FP32 values, work vectors)
        rs = [M.pcg(d_b, precision=prec, **kw) for _ in range(max(1, steps))]
        runs[name] = rs[-1]
        out[name + "_solve_ms"] = [r["solve_ms"] for r in rs]
        log(f"config3 {name}: " + ", ".join(f"{r['solve_ms']:.1f}" for r in rs) + " ms, "
            f"{rs[-1]['iters']} iterations")
    xm, xf = runs["mixed"]["x"].to_host(), runs["fp64"]["x"].to_host()
    res = cel.DeviceArray(M.n)
    M.spmv(runs["mixed"]["x"], res)
    out["mixed_inner_iters"] = runs["mixed"]["iters"]
    out["mixed_fp64_replacements"] = runs["mixed"]["outer_iters"]
    out["fp64_iters"] = runs["fp64"]["iters"]
    out["mixed_true_rel_residual"] = float(np.linalg.norm(b - res.to_host()) / bnorm)
    out["speedup_mixed_vs_fp64"] = float(np.median(out["fp64_solve_ms"]) /
                                         np.median(out["mixed_solve_ms"]))
    scale = float(np.abs(xf).max())
    out["max_rel_diff_mixed_vs_gpu_fp64"] = float(np.abs(xm - xf).max()) / scale
    out["max_abs_err_vs_x_true"] = {"mixed": float(np.abs(xm - xt).max()),
                                    "fp64": float(np.abs(xf - xt).max())}
    M.destroy()
    del d_xt, d_b, res
    if not skip_cpu:
        import oracle as O

        threads = _cpu_threads()
        t0 = time.time()
        A = O.gen_poisson7(N, N, N)
        jac = O.set_jacobi(O.csr_diagonal(A))
        ro = O.pcg_jacobi(A, jac, b, x0_mode=O.X0_ZERO, eps=tol * bnorm, iter_max=20000,
                          workers=threads - 1)
        xo = ro["x"]
        so = float(np.abs(xo).max())
        out["oracle"] = {"iters": ro["iters"], "seconds": time.time() - t0,
                         "max_abs_err_vs_x_true": float(np.abs(xo - xt).max())}
        for name, x in (("mixed", xm), ("fp64", xf)):
            out[f"{name}_max_rel_diff_vs_fp64_oracle"] = float(np.abs(x - xo).max()) / so
        log(f"config3 oracle: {ro['iters']} iterations in {time.time() - t0:.1f} s; "
            f"mixed vs oracle {out['mixed_max_rel_diff_vs_fp64_oracle']:.3e}, "
            f"fp64 vs oracle {out['fp64_max_rel_diff_vs_fp64_oracle']:.3e}")
        del A, jac
    return out


_THREADS = []


def _cpu_threads():
    """Host threads for the oracle: the cores this process may run on (at
    most OMP_NUM_THREADS), counted once, before an OpenMP runtime binds the
    main thread to one core."""
    if not _THREADS:
        threads = len(os.sched_getaffinity(0))
        _THREADS.append(max(2, min(threads, int(os.environ.get("OMP_NUM_THREADS", "16") or 16))))
    return _THREADS[0]


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--format", default="csr", choices=["csr", "jad", "dia"],
                    help="form of the block for `value` (the others are timed beside it)")
    ap.add_argument("--side-steps", type=int, default=5,
                    help="timed solves of each of the other two forms (at most --steps)")
    ap.add_argument("--skip-cpu", action="store_true", help="do not time the CPU baseline")
    ap.add_argument("--skip-config3", action="store_true", help="do not run the 384^3 workload")
    ap.add_argument("--only", default="all",
                    choices=["all", "fp64", "mixed", "spmv", "profile", "batch", "batch16", "sort",
                             "gmres", "config3", "bandwidth"],
                    help="run one part only; 'profile' = one streaming copy of known size "
                         "(counter calibration) and FP64, FP64-with-halo and mixed solves "
                         "capped at --iters iterations in every form, for hardware-counter "
                         "runs; 'batch' = the batched workload (64 variable-coefficient systems "
                         "+ 128^3 with 16 right-hand sides) against one-by-one solves; 'sort' = "
                         "COO ordering of the shuffled --size^3 matrix in HBM; 'gmres' = "
                         "Jacobi-GMRES(3) with the element solver's settings on the 128^3 block "
                         "(--iters > 0 caps its outer iterations); 'config3' = the 384^3 "
                         "precision workload; 'batch16' = the batch's 128^3 x 16 system alone "
                         "(the multi-rhs SpMV); 'bandwidth' = HBM read rate of 1-8 arrays read "
                         "together and of a copy (the measured ceiling beside the 8 TB/s peak)")
    ap.add_argument("--iters", type=int, default=None,
                    help="iteration cap of --only profile (default 20) and --only gmres (default none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the library runs one subdomain on one GPU; sharding across GPUs is not implemented)",
              file=sys.stderr)
        sys.exit(2)

    # bind the oracle's OpenMP workers to cores before its runtime starts
    _cpu_threads()
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import cel_v3_amd as cel

    if args.only == "batch":
        cel.init(0)
        print(json.dumps(batch_bench(cel, args.steps)))
        return
    if args.only == "gmres":
        cel.init(0)
        print(json.dumps(gmres_bench(cel, 128, args.steps, args.iters or 0)))
        return
    if args.only == "batch16":
        cel.init(0)
        print(json.dumps(batch16_bench(cel, args.steps)))
        return
    if args.only == "bandwidth":
        cel.init(0)
        print(json.dumps(bandwidth_bench(cel)))
        return
    if args.only == "sort":
        cel.init(0)
        print(json.dumps(sort_bench(cel, args.size, args.steps, args.warmup, args.skip_cpu)))
        return
    if args.only == "config3":
        cel.init(0)
        print(json.dumps(config3(cel, args.tol, args.steps, args.skip_cpu)))
        return

    N = args.size
    cel.init(0)
    dev, cus = cel.device_info()
    log(f"device {dev}, {cus} CUs; building {N}^3 Poisson on the device")
    M = cel.Matrix.poisson7(N, N, N)
    M.set_jacobi()
    FORMS = {"csr": cel.FMT_CSR, "jad": cel.FMT_JAD, "dia": cel.FMT_DIA}
    t0 = time.perf_counter()
    offs = M.to_dia()
    log(f"DIA form: {len(offs)} diagonals {offs}, built in {(time.perf_counter() - t0) * 1e3:.1f} ms")
    t0 = time.perf_counter()
    M.to_jad()
    jad_convert_ms = (time.perf_counter() - t0) * 1e3
    log(f"JAD form: converted on the device in {jad_convert_ms:.1f} ms")
    head = args.format
    fmt = FORMS[head]
    sides = [f for f in ("csr", "jad", "dia") if f != head]
    xt = np.random.default_rng(1).random(M.n)
    d_xt = cel.DeviceArray.from_host(xt)
    d_b = cel.DeviceArray(M.n)
    M.spmv(d_xt, d_b)
    bnorm = float(np.linalg.norm(d_b.to_host()))
    log(f"n={M.n} nnz={M.nnz} |b|={bnorm:.6e}")

    def solve(precision, form, profile=0, halo=0, mat=None):
        r = (mat or M).pcg(d_b, x0=None, tol=args.tol, tol_mode=cel.TOL_REL_B, x0_mode=cel.X0_ZERO,
                           iter_max=20000, precision=precision, fmt=FORMS[form], profile=profile,
                           use_halo=halo, allow_not_converged=False)
        return r

    solutions = {}

    def check(r, name):
        x = r["x"]
        res = cel.DeviceArray(M.n)
        M.spmv(x, res)
        xh = x.to_host()
        solutions[name] = xh
        rel = float(np.linalg.norm(d_b.to_host() - res.to_host()) / bnorm)
        err = float(np.abs(xh - xt).max())
        return rel, err

    out = {}
    if args.only == "profile":
        copy_bytes = 256 * 1024 ** 2
        ms = cel.stream_copy_time(copy_bytes, reps=2)
        prof = {"copy_bytes_read": copy_bytes, "copy_bytes_written": copy_bytes, "copy_ms": ms}
        H = None
        for form in ("csr", "jad", "dia"):
            runs = [("fp64_" + form, cel.PREC_FP64, 0), ("mixed_" + form, cel.PREC_MIXED, 0)]
            if form != "jad":
                runs.append(("fp64_halo_" + form, cel.PREC_FP64, 1))
            for name, prec, halo in runs:
                mat = M
                if halo:
                    if H is None:
                        H = cel.Matrix.poisson7(N, N, N)
                        H.set_jacobi()
                        for proc, rows, cols in face_blocks(N):
                            H.add_halo_block(proc, proc * N ** 3 + 1, N ** 3, rows, cols,
                                             -np.ones(rows.size))
                            H.set_send_list(proc, rows)
                        H.to_dia()
                    mat = H
                r = mat.pcg(d_b, x0=None, tol=args.tol, tol_mode=cel.TOL_REL_B, x0_mode=cel.X0_ZERO,
                            iter_max=args.iters or 20, precision=prec, fmt=FORMS[form], use_halo=halo,
                            profile=1)
                prof[name] = {"iters": r["iters"], "spmv_bytes": r["spmv_bytes"],
                              "mean_spmv_ms": r["spmv_ms"] / max(r["spmv_launches"], 1)}
                log(f"profile {name}: {r['iters']} iterations, K1 "
                    f"{prof[name]['mean_spmv_ms'] * 1e3:.1f} us")
        print(json.dumps({"profile": prof, "n": M.n, "nnz": M.nnz}))
        return
    # ------------------------------------------------------------ SpMV kernels alone
    if args.only in ("all", "spmv"):
        log("SpMV kernels")
        x64 = cel.DeviceArray.from_host(xt)
        y64 = cel.DeviceArray(M.n)
        x32 = cel.DeviceArray.from_host(xt.astype(np.float32))
        y32 = cel.DeviceArray(M.n, np.float32)
        spmv = {"jad_convert_ms": jad_convert_ms}
        for form in ("csr", "jad", "dia"):
            for pname, prec, x, y in (("fp64", cel.PREC_FP64, x64, y64),
                                      ("fp32", cel.PREC_MIXED, x32, y32)):
                ms, nb = M.spmv_time(x, y, fmt=FORMS[form], precision=prec, reps=20)
                spmv[f"{form}_{pname}"] = {"ms": ms, "bytes": nb, "GBs": nb / ms * 1e-6,
                                           "frac_hbm": nb / ms * 1e-6 / HBM_PEAK_GBS}
        cb = 2 * 1024 ** 3
        ms = cel.stream_copy_time(cb, reps=10)
        spmv["stream_copy_GBs"] = 2 * cb / ms * 1e-6
        for k, v in spmv.items():
            log(f"  {k}: " + (f"{v['GBs']:.1f} GB/s" if isinstance(v, dict) else f"{v:.1f}"))
        out["spmv"] = spmv
        del x64, y64, x32, y32

    if args.only == "spmv":
        print(json.dumps(out))
        return

    # ------------------------------------------------------------ warm-up
    for w in range(args.warmup):
        log(f"warm-up {w + 1}/{args.warmup}")
        if args.only in ("all", "fp64"):
            solve(cel.PREC_FP64, head)
        if args.only in ("all", "mixed"):
            solve(cel.PREC_MIXED, head)

    # ------------------------------------------------------------ timed region
    # both precisions solve the same local 256^3 block without halo blocks,
    # so the mixed/FP64 ratio compares like with like; neither is profiled
    fp64_runs, mixed_runs = [], []
    cel.lib().cel_gpu_synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        if args.only in ("all", "mixed"):
            r = solve(cel.PREC_MIXED, head)
            mixed_runs.append(r)
            log(f"step {k + 1}: mixed {head} {r['solve_ms']:.1f} ms, {r['iters']} inner iterations, "
                f"{r['outer_iters']} FP64 replacements")
    cel.lib().cel_gpu_synchronize()
    t_mixed = time.perf_counter() - t_start
    t_start = time.perf_counter()
    for k in range(args.steps):
        if args.only in ("all", "fp64"):
            r = solve(cel.PREC_FP64, head)
            fp64_runs.append(r)
            log(f"step {k + 1}: fp64 {head} {r['solve_ms']:.1f} ms, {r['iters']} iterations")
    cel.lib().cel_gpu_synchronize()
    t_fp64 = time.perf_counter() - t_start

    # the same solves in the other two forms
    side = {}
    if args.only == "all":
        ns = max(1, min(args.side_steps, args.steps))
        for form in sides:
            side[form] = {}
            for name, prec in (("mixed", cel.PREC_MIXED), ("fp64", cel.PREC_FP64)):
                solve(prec, form)  # warm-up
                rs = [solve(prec, form) for _ in range(ns)]
                side[form][name] = rs
                log(f"{form} {name}: " + ", ".join(f"{r['solve_ms']:.1f}" for r in rs) + " ms")

    # ------------------------------------------------------------ per-kernel timing (untimed)
    # one more solve of each form with HIP events around every K1 launch on
    # the library stream, and the one-subdomain halo model of the FP64 form
    prof = {}
    halo_timed = {}
    if args.only == "all":
        for form in ("csr", "jad", "dia"):
            prof["mixed_" + form] = solve(cel.PREC_MIXED, form, profile=1)
            prof["fp64_" + form] = solve(cel.PREC_FP64, form, profile=1)
        H = cel.Matrix.poisson7(N, N, N)
        H.set_jacobi()
        for proc, rows, cols in face_blocks(N):
            H.add_halo_block(proc, proc * N ** 3 + 1, N ** 3, rows, cols, -np.ones(rows.size))
            H.set_send_list(proc, rows)
        H.to_dia()
        # the local solve of the same handle (its launch shapes) to compare with
        halo_local = solve(cel.PREC_FP64, "csr", mat=H)["x"].to_host()
        for form in ("csr", "dia"):
            solve(cel.PREC_FP64, form, halo=1, mat=H)  # warm-up
            halo_timed[form] = solve(cel.PREC_FP64, form, halo=1, mat=H)
            # K1 of the same handle without and with its halo blocks, back to back
            prof["fp64_hlocal_" + form] = solve(cel.PREC_FP64, form, profile=1, mat=H)
            prof["fp64_halo_" + form] = solve(cel.PREC_FP64, form, profile=1, halo=1, mat=H)
        log("fp64 with halo faces: " + ", ".join(f"{k} {v['solve_ms']:.1f} ms" for k, v in halo_timed.items()))
        halo_x = {k: v["x"].to_host() for k, v in halo_timed.items()}
        H.destroy()

    def k1(r):
        ms = r["spmv_ms"] / max(r["spmv_launches"], 1)
        return ms, r["spmv_bytes"] / (ms * 1e-3) / 1e9

    res = dict(metric=METRIC, n_gpus=1, steps=args.steps, warmup=args.warmup,
               higher_is_better=False, scaling="weak", vs_baseline=None, data="synthetic",
               unit="s")
    cfg = {"workload": f"poisson7_{N}^3_jacobi_pcg_tol{args.tol:g}", "n": M.n, "nnz": M.nnz,
           "format": head, "x_true": "U[0,1) numpy default_rng(1)", "x0": 0,
           "stop": "||b-Ax|| <= tol*||b||", "parallelism": "1 GPU, one subdomain"}
    if mixed_runs:
        ms = [r["solve_ms"] for r in mixed_runs]
        rel, err = check(mixed_runs[-1], "mixed")
        res["value"] = float(np.median(ms)) / 1e3
        res["ms_per_step"] = t_mixed / args.steps * 1e3
        res["dtype"] = "f32+f64"
        res["mixed"] = {"format": head, "solve_ms": ms, "inner_iters": mixed_runs[-1]["iters"],
                        "fp64_replacements": mixed_runs[-1]["outer_iters"],
                        "true_rel_residual": rel, "max_abs_err_vs_x_true": err,
                        "wall_ms_per_step": t_mixed / args.steps * 1e3,
                        "iter_bytes": mixed_runs[-1]["iter_bytes"]}
        log(f"mixed {head}: {res['value']:.4f} s, true rel residual {rel:.3e}, max err {err:.3e}")
    if fp64_runs:
        ms = [r["solve_ms"] for r in fp64_runs]
        rel, err = check(fp64_runs[-1], "fp64")
        last = fp64_runs[-1]
        res["fp64"] = {"format": head, "solve_ms": ms, "iters": last["iters"], "true_rel_residual": rel,
                       "max_abs_err_vs_x_true": err, "wall_ms_per_step": t_fp64 / args.steps * 1e3,
                       "iter_bytes": last["iter_bytes"],
                       "iter_GBs": last["iter_bytes"] * last["iters"] / (np.median(ms) * 1e-3) / 1e9}
        log(f"fp64 {head}: {np.median(ms) / 1e3:.4f} s, {last['iters']} iterations")
        if "value" not in res:
            res["value"] = float(np.median(ms)) / 1e3
            res["ms_per_step"] = t_fp64 / args.steps * 1e3
            res["dtype"] = "f64"
    if mixed_runs and fp64_runs:
        res["speedup_mixed_vs_fp64"] = (np.median([r["solve_ms"] for r in fp64_runs]) /
                                        np.median([r["solve_ms"] for r in mixed_runs]))
    for form, rs in side.items():
        entry = {"mixed_solve_ms": [r["solve_ms"] for r in rs["mixed"]],
                 "fp64_solve_ms": [r["solve_ms"] for r in rs["fp64"]],
                 "mixed_inner_iters": rs["mixed"][-1]["iters"], "fp64_iters": rs["fp64"][-1]["iters"]}
        entry["speedup_mixed_vs_fp64"] = float(np.median(entry["fp64_solve_ms"]) /
                                               np.median(entry["mixed_solve_ms"]))
        for name in ("mixed", "fp64"):
            solutions[f"{form}_{name}"] = rs[name][-1]["x"].to_host()
        if "csr" in (head, form) and "dia" in (head, form):
            # the DIA form's products are the CSR ones bit for bit
            entry["solutions_bitwise_equal_to_csr" if form == "dia" else "solutions_bitwise_equal_to_dia"] = {
                name: bool(np.array_equal(solutions[f"{form}_{name}"], solutions[name]))
                for name in ("mixed", "fp64") if name in solutions}
        res[form] = entry
    fname = {"csr": "CSR", "jad": "JAD", "dia": "DIA"}
    roof_keys = [("roofline", "mixed_" + head, "float", head),
                 ("roofline_fp64", "fp64_" + head, "double", head)]
    for form in sides:
        roof_keys += [(f"roofline_{form}", "mixed_" + form, "float", form),
                      (f"roofline_{form}_fp64", "fp64_" + form, "double", form)]
    roof_keys += [("roofline_fp64_halo", "fp64_halo_" + "csr", "double", "csr"),
                  ("roofline_dia_fp64_halo", "fp64_halo_" + "dia", "double", "dia")]
    for key, pname, vt, form in roof_keys:
        if pname not in prof:
            continue
        r = prof[pname]
        halo = "halo" in pname
        name = (f"k_pcg_spmv<{vt},{fname[form]}" + (",halo>" if halo else ">") +
                (" (+6 halo faces, one-subdomain model)" if halo else " (v = A p fused with v.p)"))
        k1_ms, achieved = k1(r)
        res[key] = {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "traffic": read_traffic(name.split(" ")[0]),
                    "algorithmic_bytes": r["spmv_bytes"], "mean_launch_ms": k1_ms,
                    "launches": r["spmv_launches"], "solve_ms_profiled": r["solve_ms"]}
        log(f"{key}: {name.split(' ')[0]} {k1_ms * 1e3:.1f} us = {achieved:.0f} GB/s "
            f"({achieved / HBM_PEAK_GBS:.1%} of 8 TB/s)")
    if halo_timed:
        res["fp64_halo"] = {"csr_solve_ms": halo_timed["csr"]["solve_ms"],
                            "dia_solve_ms": halo_timed["dia"]["solve_ms"],
                            "iters": halo_timed["csr"]["iters"], "halo_faces": 6,
                            "face_entries": 6 * N * N,
                            "note": "one-subdomain model: 6 halo blocks and face packing every "
                                    "SpMV, received values zero",
                            "solutions_bitwise_equal_to_local_csr_solve": {
                                k: bool(np.array_equal(v, halo_local)) for k, v in halo_x.items()},
                            # mean K1 launch (us) of the same handle without and with its
                            # halo blocks, back to back
                            "k1_us_local_vs_halo": {
                                f: [k1(prof["fp64_hlocal_" + f])[0] * 1e3,
                                    k1(prof["fp64_halo_" + f])[0] * 1e3] for f in ("csr", "dia")}}
        for f, (a, b) in res["fp64_halo"]["k1_us_local_vs_halo"].items():
            log(f"halo K1 {f}: {b:.1f} us against {a:.1f} us without halo ({b / a:.3f}x)")
    res["config"] = cfg
    res.update(out)
    if not args.skip_config3 and args.only == "all":
        res["config3"] = config3(cel, args.tol, 1, args.skip_cpu)
    if not args.skip_cpu and fp64_runs and args.only == "all":
        base, x_oracle = cpu_baseline(N, args.tol)
        res["cpu_baseline"] = base
        # error of each GPU solution against the FP64 oracle's solution
        scale = float(np.abs(x_oracle).max())
        eo = float(np.abs(x_oracle - xt).max())
        res["fp64_oracle"] = {"iters": base["iters"], "max_abs_err_vs_x_true": eo}
        for key, sol in (("mixed", "mixed"), ("fp64", "fp64")):
            if sol in solutions and key in res:
                d = float(np.abs(solutions[sol] - x_oracle).max())
                res[key]["max_abs_diff_vs_fp64_oracle"] = d
                res[key]["max_rel_diff_vs_fp64_oracle"] = d / scale
                log(f"{key}: max |x - x_oracle| = {d:.3e} ({d / scale:.3e} of max |x|)")
        for form in side:
            for name in ("mixed", "fp64"):
                d = float(np.abs(solutions[f"{form}_{name}"] - x_oracle).max())
                res[form][f"{name}_max_rel_diff_vs_fp64_oracle"] = d / scale
        res["fp64"]["oracle_iters"] = base["iters"]
    print(json.dumps(res))


if __name__ == "__main__":
    main()
