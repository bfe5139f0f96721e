#!/usr/bin/env python3
"""Benchmark of the MI355X backend on HONEI's sparse-solver path.

Headline (one "step" = one ELL SpMV y = A x over the whole matrix):
  ELL SpMV, Matrix A = Q1 9-point stencil on 4097 x 4097 nodes
  (16,785,409 rows, 150,937,609 nonzeros, U[0.5,1.5]-scaled coefficients,
  strictly diagonally dominant), FP64, inputs resident in HBM.
  value = algorithmic bytes per SpMV / time per SpMV, with the byte model
  z(s+4) + n(2s+4) (values + 32-bit column indices + x + y + 4-byte row
  lengths).

Also reported under "extra": defect and FP32 forms, Matrix B (random ELL,
2^24 rows, 5..13 per row), Q1 banded SpMV, Poisson solves/s of
mixed-precision MG-IR (FP32 V-cycles, FP64 refinement) and plain FP64 CG,
the served stream of 4096 mixed-size MG-IR requests (solves/s, p50/p99),
and 64 right-hand sides of the variable-coefficient diffusion problem on
2049^2 (batched Jacobi-PCG and MG-IR).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Poisson solves/s at rel res 1e-8 (MG-IR vs FP64 CG), p50/p99; ELL SpMV GB/s"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_ell_spmv_matrix_a_4097_f64.txt")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def timed(fn, steps, warmup, stream=None):
    """Average ms per call of fn() over `steps` calls, HIP events on `stream`."""
    from honei_amd import device as hip
    for _ in range(warmup):
        fn()
    hip.synchronize(stream)
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    ms = e0.elapsed_ms(e1)
    hip.synchronize(stream)
    return ms / steps


def read_traffic():
    """HBM bytes per launch from the committed PMC summary, if present."""
    try:
        vals = {}
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if ":" in line and not line.startswith("#"):
                    k, v = line.split(":", 1)
                    vals[k.strip()] = v.strip()
        return float(vals["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def spmv_case(hip, P, E, steps, warmup, defect=False, keep_y=None, variant=None):
    """keep_y: a list that receives the host copy of the last y; variant: an
    explicit ELL kernel variant (hsolve_ell_set_variant), else automatic."""
    n = E["rows"]
    dtype = E["Ax"].dtype
    s = dtype.itemsize
    z = P.ell_nnz(E)
    M = hip.EllMatrix.from_dict(E)
    if variant is not None:
        M.set_variant(variant)
    x = hip.DeviceVector.from_numpy(
        P.rng_uniform(P.SEED, P.S_X, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype))
    y = hip.DeviceVector.zeros(n, dtype)
    rhs = None
    if defect:
        rhs = hip.DeviceVector.from_numpy(
            P.rng_uniform(P.SEED, P.S_RHS, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype))
    ms = timed(lambda: M.spmv(x, y, rhs=rhs), steps, warmup)
    nbytes = P.ell_defect_bytes(n, z, s) if defect else P.ell_spmv_bytes(n, z, s)
    out = dict(rows=n, nnz=z, ms=ms, bytes=nbytes, GBps=nbytes / (ms * 1e-3) / 1e9)
    out["frac_hbm"] = out["GBps"] / HBM_PEAK_GBS
    if keep_y is not None:
        keep_y.append(y.get())
    del M, x, y, rhs
    return out


def shard_case(hip, P, E, steps, warmup, parts=8):
    """One rank's share of a row-partitioned product on one GPU: y = A x on
    rows [k n / parts, (k + 1) n / parts) of the ELL handle
    (hsolve_ell_spmv_rows), each of the `parts` row blocks timed on its
    own; bytes by the same model on the block's rows and entries."""
    n = E["rows"]
    dtype = E["Ax"].dtype
    s = dtype.itemsize
    lens = np.asarray(E["Arl"], np.int64) * int(E.get("threads", 1))
    M = hip.EllMatrix.from_dict(E)
    x = hip.DeviceVector.from_numpy(
        P.rng_uniform(P.SEED, P.S_X, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype))
    y = hip.DeviceVector.zeros(n, dtype)
    blocks = []
    for k in range(parts):
        r0, r1 = k * n // parts, (k + 1) * n // parts
        ms = timed(lambda: M.spmv_rows(x, y, r0, r1), steps, warmup)
        nb = P.ell_spmv_bytes(r1 - r0, int(lens[r0:r1].sum()), s)
        blocks.append(dict(rows=[r0, r1], ms=ms, bytes=nb, GBps=nb / (ms * 1e-3) / 1e9))
    del M, x, y
    worst = min(b["GBps"] for b in blocks)
    return dict(parts=parts, blocks=blocks, ms_max=max(b["ms"] for b in blocks), GBps_min=worst,
                frac_hbm_min=worst / HBM_PEAK_GBS)


def csr_case(hip, P, E, steps, warmup, defect=False, variant=0):
    """Product / Defect of the same matrix stored as SparseMatrixCSR
    (blocksize 1) through a CSR handle: variant 0 lets the handle use its
    ELL form where it holds one, variant 1 times the CSR kernel itself."""
    n = E["rows"]
    M = P.ell_to_csr(E)
    dtype = M["Ax"].dtype
    s = dtype.itemsize
    z = M["nblocks"]
    H = hip.CsrMatrix.from_dict(M)
    H.set_variant(variant)
    del M
    x = hip.DeviceVector.from_numpy(
        P.rng_uniform(P.SEED, P.S_X, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype))
    y = hip.DeviceVector.zeros(n, dtype)
    rhs = None
    if defect:
        rhs = hip.DeviceVector.from_numpy(
            P.rng_uniform(P.SEED, P.S_RHS, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype))
    ms = timed(lambda: H.spmv(x, y, rhs=rhs), steps, warmup)
    nbytes = P.csr_spmv_bytes(n, z, s) + (n * s if defect else 0)
    out = dict(rows=n, nnz=z, ms=ms, bytes=nbytes, GBps=nbytes / (ms * 1e-3) / 1e9,
               device_bytes=H.device_bytes, kernel="ell form" if (H.ell_form and variant == 0) else "csr")
    out["frac_hbm"] = out["GBps"] / HBM_PEAK_GBS
    del H, x, y, rhs
    return out


def q1_case(hip, P, side, dtype, steps, warmup):
    n = side * side
    A = hip.Q1Matrix(P.matrix_a_bands(side, dtype=dtype))
    x = hip.DeviceVector.from_numpy(
        P.rng_uniform(P.SEED, P.S_X, np.arange(n, dtype=np.uint64), -1, 1).astype(dtype))
    y = hip.DeviceVector.zeros(n, dtype)
    ms = timed(lambda: hip.q1_product(A, x, y), steps, warmup)
    nbytes = 11 * n * np.dtype(dtype).itemsize
    return dict(rows=n, ms=ms, bytes=nbytes, GBps=nbytes / (ms * 1e-3) / 1e9,
                frac_hbm=nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS)


def mg_solve_bytes(level, min_level, inner, outer):
    """Algorithmic HBM bytes of one MG-IR solve on Q1 bands: per FP32 V-cycle
    (2 + 2 sweeps) the sum over its levels l of 5 M_l + 18 n_l 4 + 2 n_(l-1) 4
    with M_l = 36 n_l (four sweeps and one defect over the nine bands,
    the vectors they stream, restriction, prolongation with correction);
    per FP64 outer step 128 n (fused defect + norm 72n, scale + cast 24n,
    cast + update 12n + 20n).  5.69 GB per V-cycle and 2.15 GB per outer
    step at 4097^2."""
    n = lambda l: ((1 << l) + 1) ** 2
    cycle = sum(5 * 36 * n(l) + 18 * 4 * n(l) + (2 * 4 * n(l - 1) if l > min_level else 0)
                for l in range(min_level, level + 1))
    return inner * cycle + outer * 128 * n(level)


def mg_case(hip, P, level, min_level, reps, tol=1e-8, value_codes=True):
    """value_codes False: the same solve reading the plain bands (hsolve_mg_set_value_codes(0))."""
    n = P.side_of_level(level) ** 2
    mg = hip.MGSolver(hip.mg_params(min_level, level, n_max_iter=2, tolerance=tol))
    mg.set_value_codes(value_codes)
    b = hip.DeviceVector.from_numpy(P.rhs_uniform(n))
    x = hip.DeviceVector.zeros(n)
    lat = []
    info = None
    for r in range(reps + 1):
        x.fill(0.0)
        hip.synchronize()
        t0 = time.perf_counter()
        info = mg.solve(b, x)
        t1 = time.perf_counter()
        if r > 0:
            lat.append((t1 - t0) * 1e3)
    lat = np.asarray(lat)
    p50 = float(np.percentile(lat, 50))
    nbytes = mg_solve_bytes(level, min_level, info["inner"], info["outer"])
    gbps = nbytes / (p50 * 1e-3) / 1e9
    return dict(n=n, levels=f"{min_level}..{level}", tol=tol, outer=info["outer"], inner=info["inner"],
                rel_defect=info["final_defect"] / info["initial_defect"],
                ms_p50=p50, ms_p99=float(np.percentile(lat, 99)),
                solves_per_s=float(1e3 / lat.mean()), device_bytes=mg.device_bytes,
                roofline=dict(bound="hbm", bytes_per_solve=nbytes, achieved=round(gbps, 1), peak=HBM_PEAK_GBS,
                              unit="GB/s", frac=round(gbps / HBM_PEAK_GBS, 4),
                              model="per-operation bytes of the reference's V-cycle and outer step (36 B per node "
                                    "per band pass); the value-coded one-pass kernels read less, so frac can pass 1")), x


def serve_trace(count, rate, seed):
    """Served stream: side = nearest 2^k+1 to a log-uniform draw on [65, 2049]
    (k = 6..11), Poisson arrivals at `rate` per second (None: all at t = 0),
    one RHS seed per request."""
    rng = np.random.default_rng(seed)
    s = np.exp(rng.uniform(np.log(65.0), np.log(2049.0), count))
    levels = np.clip(np.rint(np.log2(s - 1.0)), 6, 11).astype(np.uint64)
    if rate is None:
        arrivals = np.zeros(count)
    else:
        arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
        arrivals -= arrivals[0]
    seeds = rng.integers(1, 1 << 62, count, dtype=np.uint64)
    return arrivals, levels, seeds


def serve_summary(res, arrivals, levels):
    lat = np.array([r["t_done"] - r["t_submit"] for r in res]) * 1e3
    wait = np.array([r["t_admit"] - r["t_submit"] for r in res]) * 1e3
    span = max(r["t_done"] for r in res) - float(np.min(arrivals))
    rel = np.array([r["final_defect"] / r["initial_defect"] for r in res])
    per_level = {}
    for L in sorted(set(int(v) for v in levels)):
        m = levels == L
        per_level[f"{(1 << L) + 1}^2"] = dict(
            requests=int(m.sum()), p50_ms=round(float(np.percentile(lat[m], 50)), 3),
            p99_ms=round(float(np.percentile(lat[m], 99)), 3),
            outer_mean=round(float(np.mean([res[i]["outer"] for i in np.nonzero(m)[0]])), 2),
            inner_mean=round(float(np.mean([res[i]["inner"] for i in np.nonzero(m)[0]])), 2))
    return dict(requests=len(res), solves_per_s=round(len(res) / span, 1), makespan_s=round(span, 4),
                p50_ms=round(float(np.percentile(lat, 50)), 3), p99_ms=round(float(np.percentile(lat, 99)), 3),
                admit_wait_p50_ms=round(float(np.percentile(wait, 50)), 3),
                admit_wait_p99_ms=round(float(np.percentile(wait, 99)), 3),
                max_rel_defect=float(rel.max()), all_converged=bool(all(r["status"] == 0 for r in res)),
                per_level=per_level)


def serve_case(hip, count, budget_gib=32, max_slots=64, load=0.7, seed=11):
    """Config: a served stream of `count` Poisson Q1 MG-IR solves (sides 65..2049,
    FP32 V-cycles + FP64 refinement to rel. residual 1e-8, device-generated
    seeded U[-1,1] RHS).  First all requests at t = 0 (saturation
    throughput), then Poisson arrivals at `load` x that rate (latency)."""
    srv = hip.Server(hip.serve_params(min_level=2, min_served_level=6, max_served_level=11, n_max_iter=2,
                                      tolerance=1e-8, memory_budget=budget_gib << 30, max_slots=max_slots,
                                      use_graphs=True))
    a, l, s = serve_trace(min(count, 256), None, seed + 1)
    srv.run(a, l, s)                                           # warm-up: graph captures, slot caches
    a, l, s = serve_trace(count, None, seed)
    burst = serve_summary(srv.run(a, l, s), a, l)
    rate = load * burst["solves_per_s"]
    a, l, s = serve_trace(count, rate, seed + 2)
    steady = serve_summary(srv.run(a, l, s), a, l)
    st = srv.stats()
    return dict(burst=burst, poisson_arrivals=dict(rate_per_s=round(rate, 1), **steady),
                memory_budget_bytes=st["budget"], shared_bytes=st["bytes_shared"],
                peak_running=st["peak_running"], max_slots_per_level=max_slots,
                graph_captures=st["graph_captures"], ticks=st["ticks"])


def diffusion_case(hip, P, level=11, min_level=2, nrhs=64, tol=1e-8, pcg_max=200000, pcg=True):
    """Variable-coefficient diffusion -div(k grad u) = f on the (2^level+1)^2
    Q1 grid (k = exp(N(0,1)) on 64 x 64 tiles; Dirichlet left/right, Neumann
    bottom/top) as SparseMatrixELL; nrhs seeded U[-1,1] right-hand sides
    solved in one batch to rel. residual tol by
      * Jacobi-PCG, CGSolver<VAR> in FP64 on a column-blocked multivector
        (one pass over the matrix per iteration for all columns), and
      * MG-IR, FP32 V-cycles (levels min_level..level, re-discretised with
        cell-averaged k) inside FP64 refinement, one engine slot per RHS.
    The true FP64 residual ||b - A x|| / ||b|| of every column is checked."""
    side = P.side_of_level(level)
    n = side * side
    t0 = time.perf_counter()
    bands = P.diffusion_q1_bands(level)
    A64 = P.q1_to_ell(bands)
    a32 = {L: P.diffusion_ell(L, dtype=np.float32) for L in range(min_level, level + 1)}
    log(f"diffusion {side}^2 hierarchy built in {time.perf_counter() - t0:.1f}s")
    out = dict(n=n, nrhs=nrhs, levels=f"{min_level}..{level}", tol=tol, nnz=P.ell_nnz(A64))
    A = hip.EllMatrix.from_dict(A64)

    def true_rel(X, B):
        Rm = hip.DeviceVector.zeros(n * nrhs)
        A.spmm(nrhs, X, Rm, rhs=B)
        r = Rm.get().reshape(n, nrhs)
        del Rm
        b = B.get().reshape(n, nrhs)
        return np.linalg.norm(r, axis=0) / np.linalg.norm(b, axis=0)

    Xp = None
    if pcg:
        Pd = hip.DeviceVector.from_numpy(1.0 / bands[4])
        B = hip.DeviceVector.zeros(n * nrhs)
        hip.rhs_uniform_multi(B, side, nrhs, nrhs, 1, mask=P.DIFFUSION_MASK)
        Xp = hip.DeviceVector.zeros(n * nrhs)
        hip.synchronize()
        t1 = time.perf_counter()
        used, rel = hip.cg_ell_multi(A, Pd, B, Xp, nrhs, pcg_max, tol, raise_on_cap=False)
        t2 = time.perf_counter()
        tr = true_rel(Xp, B)
        out["jacobi_pcg_fp64"] = dict(ms=(t2 - t1) * 1e3, solves_per_s=nrhs / (t2 - t1),
                                      iters_min=int(used.min()), iters_max=int(used.max()),
                                      max_rel_defect=float(rel.max()), max_true_rel_residual=float(tr.max()),
                                      ms_per_iteration=(t2 - t1) * 1e3 / float(used.max()))
        log("diffusion pcg", out["jacobi_pcg_fp64"])
        Xp = Xp.get().reshape(n, nrhs)
        del B, Pd
    mg = hip.MGSolver.from_ell(hip.mg_params(min_level, level, n_max_iter=2, tolerance=tol, mask=P.DIFFUSION_MASK),
                               A64, a32, None, slots=nrhs)
    del a32
    Bs = hip.DeviceVector.zeros(n * nrhs)
    hip.rhs_uniform_multi(Bs, side, nrhs, 1, n, mask=P.DIFFUSION_MASK)
    Xs = hip.DeviceVector.zeros(n * nrhs)
    bs = [Bs.view(j * n, n) for j in range(nrhs)]
    xs = [Xs.view(j * n, n) for j in range(nrhs)]
    mg.solve_batch(bs[:1], xs[:1])                 # warm-up (one slot)
    Xs.fill(0.0)
    hip.synchronize()
    t1 = time.perf_counter()
    infos = mg.solve_batch(bs, xs, raise_on_cap=False)
    t2 = time.perf_counter()
    tmp = hip.DeviceVector.zeros(n)
    tr = []
    for j in range(nrhs):
        A.spmv(xs[j], tmp, rhs=bs[j])
        tr.append(np.linalg.norm(tmp.get()) / np.linalg.norm(bs[j].get()))
    outer = [i["outer"] for i in infos]
    inner = [i["inner"] for i in infos]
    out["mg_ir_fp32_fp64"] = dict(ms=(t2 - t1) * 1e3, solves_per_s=nrhs / (t2 - t1), outer_min=min(outer),
                                  outer_max=max(outer), inner_max=max(inner),
                                  max_rel_defect=max(i["final_defect"] / i["initial_defect"] for i in infos),
                                  max_true_rel_residual=float(max(tr)), device_bytes=mg.device_bytes)
    if Xp is not None:
        xm = Xs.get().reshape(nrhs, n).T
        out["max_rel_diff_mg_vs_pcg"] = float(np.max(np.linalg.norm(xm - Xp, axis=0) / np.linalg.norm(Xp, axis=0)))
        out["mg_ir_speedup_vs_pcg"] = out["jacobi_pcg_fp64"]["ms"] / out["mg_ir_fp32_fp64"]["ms"]
    log("diffusion mg", out["mg_ir_fp32_fp64"])
    return out


def cg_case(hip, P, side, max_iters, tol=1e-8):
    n = side * side
    A = hip.Q1Matrix(P.poisson_q1_bands(n))
    b = hip.DeviceVector.from_numpy(P.rhs_uniform(n))
    x = hip.DeviceVector.zeros(n)
    hip.synchronize()
    t0 = time.perf_counter()
    used, rel = hip.cg_q1(A, b, x, max_iters, tol, raise_on_cap=False)
    t1 = time.perf_counter()
    return dict(n=n, iters=used, rel_defect=rel, ms=(t1 - t0) * 1e3, solves_per_s=1.0 / (t1 - t0))


def cg257_case(hip, P, reps=5, tol=1e-8, max_iters=10000, cpu=True):
    """Plain FP64 CG on the Q1 Poisson problem with 257 x 257 nodes, as
    BandedMatrixQx<Q1> and as SparseMatrixELL (x0 = 0, seeded U[-1, 1] rhs
    with zero boundary entries): GPU time per solve (median of reps), and the
    oracle's CG (the tags::CPU restatement, one thread) on the same inputs:
    its time, its iteration count and the relative difference of the
    solutions.  The working set (8 MB) stays in cache, so this is a latency
    figure, not a roofline one."""
    from oracle.pyoracle import Oracle
    n = 257 * 257
    bands = P.poisson_q1_bands(n)
    E = P.q1_to_ell(bands)
    b = P.rhs_uniform(n)
    out = {}
    for form in ("q1", "ell"):
        A = hip.Q1Matrix(bands) if form == "q1" else hip.EllMatrix.from_dict(E)
        bd, xd = hip.DeviceVector.from_numpy(b), hip.DeviceVector.zeros(n)
        lat = []
        for r in range(reps + 1):
            xd.fill(0.0)
            hip.synchronize()
            t0 = time.perf_counter()
            if form == "q1":
                used, rel = hip.cg_q1(A, bd, xd, max_iters, tol, raise_on_cap=False)
            else:
                used, rel = hip.cg_ell(A, bd, xd, max_iters, tol, raise_on_cap=False)
            t1 = time.perf_counter()
            if r > 0:
                lat.append((t1 - t0) * 1e3)
        res = dict(iters=used, rel_defect=rel, ms_p50=float(np.median(lat)), solves_per_s=float(1e3 / np.median(lat)))
        if cpu:
            O = Oracle()
            t0 = time.perf_counter()
            if form == "q1":
                xr, it, _ = O.cg_none_q1(bands, b, np.zeros(n), max_iters, tol)
            else:
                xr, it, _ = O.cg_none_ell(E, b, np.zeros(n), max_iters, tol)
            cpu_ms = (time.perf_counter() - t0) * 1e3
            x = xd.get()
            res.update(oracle_iters=it, oracle_ms=cpu_ms, speedup_vs_oracle=cpu_ms / res["ms_p50"],
                       rel_diff_vs_oracle=float(np.linalg.norm(x - xr) / np.linalg.norm(xr)))
        out[form] = res
        del A, bd, xd
    return out


def cpu_threads():
    """Host threads for the multi-core baseline: this process's CPU set,
    at most 16 (one GPU's share of a shared host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline_spmv(E, budget_s=10.0):
    """CPU ELL SpMV over the same Matrix A and x, repeated for about budget_s
    seconds: on one thread (Product<tags::CPU::SSE>), and with the rows split
    into one equal range per thread as Product<tags::CPU::MultiCore::SSE>
    does (product.hh:2087-2106).  The kernel is the reference's own SSE
    product_smell_dv (honei/backends/sse/product.cc), compiled from its sources
    into oracle/_ref, when that library is present ("reference"); else the
    oracle's C restatement of the same loop ("port").  Returns (multi-thread
    result, single-thread result, the product y), y checked against the
    oracle's."""
    import ctypes as C
    from concurrent.futures import ThreadPoolExecutor
    from honei_amd import problems as P
    from oracle.pyoracle import REF_SO, Oracle, have_reference
    n = E["rows"]
    x = P.rng_uniform(P.SEED, P.S_X, np.arange(n, dtype=np.uint64), -1, 1)
    y = np.zeros(n)
    Aj, Ax, Arl = E["Aj"], E["Ax"], np.ascontiguousarray(E["Arl"], np.uint64)
    stride, threads = E["stride"], E.get("threads", 1)
    O = Oracle()
    kind = "port"
    if have_reference():
        try:
            fn = C.CDLL(REF_SO).ref_sse_product_smell_dv_f64   # ctypes releases the GIL during the call
            kind = "reference"
        except (OSError, AttributeError):
            kind = "port"
    vp_ = C.c_void_p
    if kind == "reference":
        ptr = [a.ctypes.data_as(vp_) for a in (y, Aj, Ax, Arl, x)]

        def rows(r0, r1):
            fn(*ptr, C.c_uint64(stride), C.c_uint64(n), C.c_uint64(E["ncpr"]), C.c_uint64(r0), C.c_uint64(r1),
               C.c_uint64(threads))
        what = "the reference's SSE product_smell_dv (Product<tags::CPU::SSE>)"
    else:
        ofn = O.lib.or_ell_product_f64
        ptr = [a.ctypes.data_as(vp_) for a in (y, Aj, Ax, Arl, x)]

        def rows(r0, r1):
            ofn(*ptr, C.c_uint64(r0), C.c_uint64(r1), C.c_uint64(stride), C.c_uint64(threads))
        what = "the oracle's Product<tags::CPU> ELL"

    nbytes = P.ell_spmv_bytes(n, P.ell_nnz(E), 8)
    out = []
    for T in (1, cpu_threads()):
        bounds = [(n * t // T, n * (t + 1) // T) for t in range(T)]
        with ThreadPoolExecutor(max_workers=T) as ex:
            t0 = time.perf_counter()
            reps = 0
            while True:
                list(ex.map(lambda b: rows(*b), bounds))
                reps += 1
                if time.perf_counter() - t0 > budget_s:
                    break
            dt = (time.perf_counter() - t0) / reps
        out.append(dict(value=nbytes / dt / 1e9, unit="GB/s", cores=T, kind=kind,
                        sample=f"{reps} full SpMVs of Matrix A FP64 ({n} rows) by {what}, "
                               f"{dt * 1e3:.1f} ms each, {T} thread(s) over equal row ranges, on {cpu_model()}"))
    if kind == "reference":
        y_or = np.zeros(n)
        O.lib.or_ell_product_f64(y_or.ctypes.data_as(vp_), Aj.ctypes.data_as(vp_), Ax.ctypes.data_as(vp_),
                                 Arl.ctypes.data_as(vp_), x.ctypes.data_as(vp_), C.c_uint64(0), C.c_uint64(n),
                                 C.c_uint64(stride), C.c_uint64(threads))
        if not np.array_equal(y, y_or):
            raise AssertionError("reference SSE product differs from the oracle's")
    return out[1], out[0], y


def mg_vs_oracle(P, level, min_level, tol, x_gpu, info_gpu):
    """The oracle's MG-IR (C restatement of Multigrid<tags::CPU, ..., MIXED>,
    one thread) on the same problem: iteration counts, relative difference of
    the solutions, and the CPU solve time."""
    from oracle.pyoracle import Oracle
    O = Oracle()
    n = P.side_of_level(level) ** 2
    a32, d32 = {}, {}
    for L in range(min_level, level + 1):
        bands = [b.astype(np.float32) for b in P.poisson_q1_bands(P.side_of_level(L) ** 2)]
        a32[L] = bands
        # ElementInverse, then Scale by 0.7 in double (poisson_mg_mixed_banded_TEST.cc:411-419)
        inv = bands[4].copy()
        nz = inv != 0
        inv[nz] = np.float32(1.0) / inv[nz]
        d32[L] = (inv.astype(np.float64) * 0.7).astype(np.float32)
    A64 = P.poisson_q1_bands(n)
    b = P.rhs_uniform(n)
    t0 = time.perf_counter()
    r = O.mg_mixed_q1(A64, b, np.zeros(n), a32, d32, min_level=min_level, max_level=level, tolerance=tol)
    cpu_s = time.perf_counter() - t0
    rel = float(np.linalg.norm(x_gpu - r["x"]) / np.linalg.norm(r["x"]))
    return dict(oracle_outer=r["outer"], oracle_inner=r["inner"], gpu_outer=info_gpu["outer"],
                gpu_inner=info_gpu["inner"], rel_diff_vs_oracle=rel,
                oracle_rel_defect=r["final_defect"] / r["initial_defect"],
                cpu_solve_s=cpu_s, cpu_threads=1)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--quick", action="store_true", help="headline only (no extras, no CPU baseline)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--side", type=int, default=4097)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 2 GiB device copy kernel (for byte-counter calibration)")
    ap.add_argument("--mg-only", type=int, default=0, metavar="REPS",
                    help="only time REPS MG-IR solves on the 4097^2 Poisson problem (for profiling)")
    ap.add_argument("--serve-only", type=int, default=0, metavar="REQUESTS",
                    help="only run the served mixed-size stream with REQUESTS requests")
    ap.add_argument("--serve-budget-gib", type=int, default=32,
                    help="with --serve-only: device-memory budget of the request server (GiB)")
    ap.add_argument("--serve-max-slots", type=int, default=64,
                    help="with --serve-only: slots per grid level of the request server")
    ap.add_argument("--diffusion-only", type=int, default=0, metavar="LEVEL",
                    help="only run the 64-RHS variable-coefficient diffusion case on the 2^LEVEL+1 grid")
    ap.add_argument("--no-pcg", action="store_true", help="with --diffusion-only: MG-IR leg only")
    ap.add_argument("--cg257-only", action="store_true",
                    help="plain FP64 CG on the 257x257 Q1 Poisson problem (banded and ELL) with the oracle beside it")
    ap.add_argument("--shard-only", action="store_true",
                    help="the row-block legs only: Matrix A and Matrix B products on each eighth of the rows")
    ap.add_argument("--matb-only", action="store_true",
                    help="ELL product / defect on Matrix B (FP64, FP32) only")
    ap.add_argument("--csr-only", action="store_true",
                    help="SparseMatrixCSR product / defect on Matrix A (FP64, FP32) and Matrix B (FP64) only")
    ap.add_argument("--pcg-max", type=int, default=200000,
                    help="with --diffusion-only: iteration cap of the PCG leg (a short cap for profiling)")
    args = ap.parse_args()
    if args.cg257_only:
        from honei_amd import device as hip
        from honei_amd import problems as P
        print(json.dumps({"cg_fp64_poisson_257": cg257_case(hip, P)}))
        return
    if args.shard_only:
        from honei_amd import device as hip
        from honei_amd import problems as P
        E = P.q1_to_ell(P.matrix_a_bands(args.side, dtype=np.float64))
        out = {"ell_spmv_matrix_a_f64": spmv_case(hip, P, E, args.steps, args.warmup),
               "ell_spmv_matrix_a_f64_row_blocks_of_8": shard_case(hip, P, E, args.steps, args.warmup)}
        del E
        EB = P.matrix_b_ell(1 << 24)
        out["ell_spmv_matrix_b_f64"] = spmv_case(hip, P, EB, args.steps, args.warmup)
        out["ell_spmv_matrix_b_f64_row_blocks_of_8"] = shard_case(hip, P, EB, args.steps, args.warmup)
        print(json.dumps(out))
        return
    if args.matb_only:
        from honei_amd import device as hip
        from honei_amd import problems as P
        EB = P.matrix_b_ell(1 << 24)
        out = {}
        if args.copy_calibration:
            nb = 2 << 30
            src, dst = hip.DeviceBuffer(nb), hip.DeviceBuffer(nb)
            ms = timed(lambda: hip.copy_kernel(src, dst, nb), max(args.steps // 5, 3), 1)
            out["copy_kernel_2GiB"] = dict(bytes=2 * nb, ms=ms, GBps=2 * nb / (ms * 1e-3) / 1e9)
            del src, dst
        out.update({"ell_spmv_matrix_b_f64": spmv_case(hip, P, EB, args.steps, args.warmup),
               "ell_defect_matrix_b_f64": spmv_case(hip, P, EB, args.steps, args.warmup, defect=True),
               "ell_spmv_matrix_b_f64_no_lds": spmv_case(hip, P, EB, args.steps, args.warmup, variant=4)})
        EB["Ax"] = EB["Ax"].astype(np.float32)
        out["ell_spmv_matrix_b_f32"] = spmv_case(hip, P, EB, args.steps, args.warmup)
        print(json.dumps(out))
        return
    if args.csr_only:
        from honei_amd import device as hip
        from honei_amd import problems as P
        E = P.q1_to_ell(P.matrix_a_bands(args.side, dtype=np.float64))
        out = {}
        if args.copy_calibration:
            nb = 2 << 30
            src, dst = hip.DeviceBuffer(nb), hip.DeviceBuffer(nb)
            ms = timed(lambda: hip.copy_kernel(src, dst, nb), max(args.steps // 5, 3), 1)
            out["copy_kernel_2GiB"] = dict(bytes=2 * nb, ms=ms, GBps=2 * nb / (ms * 1e-3) / 1e9)
            del src, dst
        out.update({"csr_spmv_matrix_a_f64": csr_case(hip, P, E, args.steps, args.warmup),
                    "csr_spmv_matrix_a_f64_csr_kernel": csr_case(hip, P, E, args.steps, args.warmup, variant=1),
                    "csr_defect_matrix_a_f64_csr_kernel": csr_case(hip, P, E, args.steps, args.warmup, defect=True,
                                                                   variant=1)})
        E["Ax"] = E["Ax"].astype(np.float32)
        out["csr_spmv_matrix_a_f32_csr_kernel"] = csr_case(hip, P, E, args.steps, args.warmup, variant=1)
        del E
        EB = P.matrix_b_ell(1 << 24)
        out["csr_spmv_matrix_b_f64"] = csr_case(hip, P, EB, args.steps, args.warmup)
        out["csr_spmv_matrix_b_f64_csr_kernel"] = csr_case(hip, P, EB, args.steps, args.warmup, variant=1)
        print(json.dumps(out))
        return
    if args.diffusion_only:
        from honei_amd import device as hip
        from honei_amd import problems as P
        print(json.dumps({f"diffusion_{P.side_of_level(args.diffusion_only)}_64rhs":
                          diffusion_case(hip, P, level=args.diffusion_only, pcg=not args.no_pcg,
                                         pcg_max=args.pcg_max)}))
        return
    if args.serve_only:
        from honei_amd import device as hip
        print(json.dumps({"serve_mixed_poisson": serve_case(hip, args.serve_only, budget_gib=args.serve_budget_gib,
                                                            max_slots=args.serve_max_slots)}))
        return
    if args.mg_only:
        from honei_amd import device as hip
        from honei_amd import problems as P
        mg, _ = mg_case(hip, P, 12, 2, reps=args.mg_only, tol=1e-8)
        print(json.dumps({"mg_ir_poisson_4097_tol1e-8": mg}))
        return

    if args.gpus != 1:
        print(f"bench.py: this build targets one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by grid row blocks: one rank per GPU, halo exchange and RCCL reductions)",
              file=sys.stderr)
        sys.exit(2)

    from honei_amd import device as hip
    from honei_amd import problems as P

    side = args.side
    t0 = time.perf_counter()
    E64 = P.q1_to_ell(P.matrix_a_bands(side, dtype=np.float64))
    log(f"Matrix A {side}^2 built in {time.perf_counter() - t0:.1f}s: rows={E64['rows']} nnz={P.ell_nnz(E64)}")

    extra = {}
    if args.copy_calibration:
        nb = 2 << 30
        src, dst = hip.DeviceBuffer(nb), hip.DeviceBuffer(nb)
        ms = timed(lambda: hip.copy_kernel(src, dst, nb), max(args.steps // 5, 3), 1)
        extra["copy_kernel_2GiB"] = dict(bytes=2 * nb, ms=ms, GBps=2 * nb / (ms * 1e-3) / 1e9)
        log("copy", extra["copy_kernel_2GiB"])
        del src, dst

    y_head = []
    head = spmv_case(hip, P, E64, args.steps, args.warmup, keep_y=y_head)
    log("headline", head)
    traffic = read_traffic()
    if not args.quick:
        extra["ell_defect_matrix_a_f64"] = spmv_case(hip, P, E64, args.steps, args.warmup, defect=True)
        E32 = dict(E64)
        E32["Ax"] = E64["Ax"].astype(np.float32)
        extra["ell_spmv_matrix_a_f32"] = spmv_case(hip, P, E32, args.steps, args.warmup)
        extra["ell_defect_matrix_a_f32"] = spmv_case(hip, P, E32, args.steps, args.warmup, defect=True)
        del E32
        extra["ell_spmv_matrix_a_f64_row_blocks_of_8"] = shard_case(hip, P, E64, args.steps, args.warmup)
        extra["csr_spmv_matrix_a_f64"] = csr_case(hip, P, E64, args.steps, args.warmup)
        extra["csr_spmv_matrix_a_f64_csr_kernel"] = csr_case(hip, P, E64, args.steps, args.warmup, variant=1)
        log("matrix A extras done")
        t1 = time.perf_counter()
        EB = P.matrix_b_ell(1 << 24)
        log(f"Matrix B built in {time.perf_counter() - t1:.1f}s")
        extra["ell_spmv_matrix_b_f64"] = spmv_case(hip, P, EB, args.steps, args.warmup)
        extra["ell_spmv_matrix_b_f64_row_blocks_of_8"] = shard_case(hip, P, EB, args.steps, args.warmup)
        extra["csr_spmv_matrix_b_f64"] = csr_case(hip, P, EB, args.steps, args.warmup)
        extra["csr_spmv_matrix_b_f64_csr_kernel"] = csr_case(hip, P, EB, args.steps, args.warmup, variant=1)
        EB["Ax"] = EB["Ax"].astype(np.float32)
        extra["ell_spmv_matrix_b_f32"] = spmv_case(hip, P, EB, args.steps, args.warmup)
        del EB
        extra["q1_spmv_matrix_a_f64"] = q1_case(hip, P, side, np.float64, args.steps, args.warmup)
        log("spmv extras done")
        mg, x_mg = mg_case(hip, P, 12, 2, reps=5, tol=1e-8)
        extra["mg_ir_poisson_4097_tol1e-8"] = mg
        if not args.no_cpu:
            mg["vs_oracle"] = mg_vs_oracle(P, 12, 2, 1e-8, x_mg.get(), mg)
            log("mg vs oracle", mg["vs_oracle"])
        del x_mg
        mg10, _ = mg_case(hip, P, 12, 2, reps=3, tol=1e-10)
        extra["mg_ir_poisson_4097_tol1e-10"] = mg10
        mgp, _ = mg_case(hip, P, 12, 2, reps=3, tol=1e-8, value_codes=False)
        extra["mg_ir_poisson_4097_tol1e-8_plain_bands"] = mgp
        log("mg", mg, mg10, mgp)
        extra["cg_fp64_poisson_257"] = cg257_case(hip, P, cpu=not args.no_cpu)
        log("cg 257", extra["cg_fp64_poisson_257"])
        cg = cg_case(hip, P, 4097, 100000)
        extra["cg_fp64_poisson_4097_tol1e-8"] = cg
        extra["mg_ir_speedup_vs_fp64_cg_4097"] = cg["ms"] / mg["ms_p50"]
        log("cg", cg)
        extra["serve_mixed_poisson"] = serve_case(hip, 4096)
        log("serve", {k: v for k, v in extra["serve_mixed_poisson"].items() if k != "burst"})
        extra["diffusion_2049_64rhs"] = diffusion_case(hip, P, level=11)

    cpu = None
    if not args.quick and not args.no_cpu:
        cpu, cpu1, y_oracle = cpu_baseline_spmv(E64)
        extra["cpu_baseline_1thread"] = cpu1
        diff = np.abs(y_head[0] - y_oracle)
        extra["headline_vs_oracle"] = dict(rows=int(y_oracle.size), rows_differing=int(np.count_nonzero(diff)),
                                           max_abs_diff=float(diff.max()))
        log("cpu", cpu, cpu1, extra["headline_vs_oracle"])

    res = {
        "metric": METRIC,
        "value": round(head["GBps"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"ELL SpMV y=Ax, Matrix A: Q1 9-point stencil {side}x{side} "
                        f"({head['rows']} rows, {head['nnz']} nnz), FP64, inputs resident in HBM",
            "bytes_model": "z*(s+4) + n*(2s+4), s=8",
            "bytes_per_step": head["bytes"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(head["GBps"], 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(head["GBps"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            # measured HBM bytes per launch over the launch time: the physical
            # rate (the stencil-form kernel reads no column indices of full rows)
            "frac_traffic": (round(traffic / (head["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None),
            # the general-ELL case (Matrix B, 2^24 rows) on the same byte model
            "frac_matrix_b": (round(extra["ell_spmv_matrix_b_f64"]["frac_hbm"], 4)
                              if "ell_spmv_matrix_b_f64" in extra else None),
        },
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
