#!/usr/bin/env python3
"""Benchmark of the factor-and-solve path on one MI355X.

Workload (default): 64^3 7-point convection-diffusion operator, n = 262,144,
nnz = 1,810,432, diagonal 6.1, off-diagonals -1 + 0.25 (u - 0.5) (seed 0),
b = A x_true with x_true ~ U[-1,1) (seed 1), FP64, 128 nested-dissection
leaves of ~1.5-2k rows (the library's graph ND, seed 0).  One step = one numeric
factorization from A's values resident in HBM plus one solve with b resident
in HBM (the reference's "Factsolve", src/solver.cu:68-76).  Construct
(ordering, symbolic analysis, upload) is outside the timed region.

Prints one JSON line.  Besides the contract fields it reports the factor and
solve times, solves/s with resident factors (1 and 16 right-hand sides), the
backward error and the relative difference to the CPU oracle, the roofline
of the dominant kernel (the grouped FP64 MFMA GEMM, measured with HIP events
on its stream) and the CPU baseline (the oracle's MKL host path, timed on
this host's cores on the same workload).

Three further legs (skippable) report the other forms of the path; every
FP32 factorization refines to the metric's 1e-12 backward error:
  "precision": 96^3 27-point anisotropic operator (n = 884,736, 512 leaves),
     FP32 factors + FP64 refinement to <= 1e-12 backward error against the
     plain FP64 path (factor and solve times, refinement steps, errors);
  "batch": 512 2D 9-point matrices (sides log-uniform in [16, 192]) factorized
     and solved as one batch, FP64 and FP32 + refinement, against the same
     matrices run one by one;
  "serving": 16 resident FP32 + refinement factorizations (8 3D 7-point grids,
     sides in 32..80; 8 2D 5-point grids, sides in 256..1024), a Poisson
     stream of solve requests with 2^k columns (k uniform in 0..6), 5% of
     requests new factorizations admitted under a 200 GB budget; p50 / p99
     latency and solves/s at the stated rate; then the same 16 matrices
     under a budget of half their device bytes ("under_pressure"): the
     server parks idle least-recently-used matrices and admits a parked one
     again when a request for it arrives; end-to-end latency from arrival
     (re-admission included), re-admissions and evictions.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "factor time & solve/s at 1e-12 backward error, FP64 vs FP32+IR; p99 latency"
# FP32 refinement target of every leg (precision, batch, serving): the
# metric's backward error
IR_TOL = 1e-12
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_BYTES = 8.0e12


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c2", "c3"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-precision", action="store_true", help="skip the 96^3 FP32+IR leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-matrix batch leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also launch a plain 1 GiB copy kernel (snu::k_copy), the known-size "
                         "reference for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="requests per second (0: half the capacity measured by a closed-loop burst)")
    ap.add_argument("--serve-requests", type=int, default=400)
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this library runs on one GPU only (asked for {max(args.gpus, world)}); "
              "the separator tree is not sharded across GPUs, so there is no multi-GPU figure.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import snusolver_amd as S

    if args.calibrate_copy:
        import ctypes as C
        ms = C.c_double()
        S._check(S.lib().snu_bench_copy(C.c_size_t(1 << 30), 3, C.byref(ms)))
        print(f"copy kernel: 1 GiB in {ms.value:.3f} ms ({2 * (1 << 30) / ms.value / 1e6:.0f} GB/s read+write)",
              file=sys.stderr, flush=True)
    A, grid, np_ = S.config_matrix(args.config)
    n = A.n
    _, b1 = S.gen_rhs(A, 1, 1)
    _, b16 = S.gen_rhs(A, 16, 1)
    t0 = time.perf_counter()
    order, sizes, order_name = S.config_order(args.config, A, grid, np_)
    order_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    M = S.SnuMat(A, np_, order=order, sizes=sizes, device=0)
    construct_s = time.perf_counter() - t0
    counts = M.counts()

    d_b = S.DeviceBuffer(8 * n * 16)
    d_x = S.DeviceBuffer(8 * n * 16)
    d_b.upload(np.asfortranarray(b1))

    def step():
        M.factorize()
        M.solve_device_ptr(d_b.ptr.value, d_x.ptr.value, 1)

    for _ in range(args.warmup):
        step()
    S.device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    S.device_sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    tim = M.timings()
    x = np.zeros(n)
    d_x.download(x)
    berr = S.backward_error(A, x, b1[:, 0])

    # solves per second with resident factors
    def solve_rate(nrhs, reps):
        if nrhs > 1:
            d_b.upload(np.asfortranarray(b16[:, :nrhs]))
        M.solve_device_ptr(d_b.ptr.value, d_x.ptr.value, nrhs)
        S.device_sync()
        t = time.perf_counter()
        for _ in range(reps):
            M.solve_device_ptr(d_b.ptr.value, d_x.ptr.value, nrhs)
        S.device_sync()
        return reps / (time.perf_counter() - t)

    solves1 = solve_rate(1, 10)
    solves16 = solve_rate(16, 5)
    X16 = np.zeros((n, 16), order="F")
    d_x.download(X16)
    berr16 = S.backward_error(A, X16, b16)

    # dominant kernel: the grouped FP64 MFMA GEMM (C -= sum A B) instantiation
    # (64 x 64 or 128 x 128 tiles) with the larger share of kernel time,
    # timed per launch with HIP events on the library's stream (every launch
    # of it in a factorization: the leaf fronts' and the dense levels')
    M.set_profiling(True)
    M.factorize()
    M.set_profiling(False)
    pt = M.timings()
    sp = M.gemm_split()
    tiles = {64: (sp["t64_ms"], sp["t64_launches"], sp["t64_gflop"]),
             128: (sp["t128_ms"], sp["t128_launches"], sp["t128_gflop"])}
    dom = 128 if sp["t128_ms"] > sp["t64_ms"] else 64
    gemm_ms, gemm_launches, gemm_gflop = tiles[dom]
    oth_ms, oth_launches, oth_gflop = tiles[192 - dom]
    achieved = gemm_gflop / gemm_ms if gemm_ms > 0 else 0.0  # GFLOP/ms = TFLOP/s
    blocks = 4 if dom == 128 else 2
    roofline = {
        "bound": "mfma",
        "achieved": round(achieved, 3),
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
        "traffic": profiled_traffic(dom),
        "kernel": f"snu::k_gemm<double, 0, {blocks}, {blocks}, 16, 1> "
                  f"(grouped FP64 MFMA 16x16x4, {dom} x {dom} tiles, C -= sum A B)",
        "launches": int(gemm_launches),
        "avg_launch_ms": round(gemm_ms / max(gemm_launches, 1), 4),
        "gflop_per_launch": round(gemm_gflop / max(gemm_launches, 1), 4),
        "other_tile": {"tile": 192 - dom, "ms": round(oth_ms, 3), "launches": int(oth_launches),
                       "tflops": round(oth_gflop / oth_ms, 3) if oth_ms > 0 else None},
        "assign_variant": {"ms": round(pt["gemm_assign_ms"], 3),
                           "launches": int(pt["gemm_assign_launches"]),
                           "gflop": round(pt["gemm_assign_gflop"], 3)},
    }
    work = M.work()
    M.close()
    d_b.free()
    d_x.free()

    line = {
        "metric": METRIC,
        "value": round(1e3 / ms_per_step, 4),
        "unit": "factorize+solve/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "64^3 7-point convection-diffusion, FP64, 128 ND leaves, 1 rhs, factorize+solve"
            if args.config == "c2" else "96^3 27-point anisotropic, FP64, 512 ND leaves, 1 rhs",
            "n": n, "nnz": A.nnz, "leaves": np_, "nrhs": 1, "ordering": order_name,
        },
        "factor_ms": round(tim["factor_ms"], 3),
        "solve_ms": round(tim["solve_ms"], 3),
        "phases_ms": {k: round(tim[k], 3) for k in
                      ["scatter_ms", "leaf_fronts_ms", "leaf_schur_ms", "assemble_ms", "dense_ms"]},
        # achieved work: the flops the dense launches execute (zero tiles
        # skipped) over the dense phase's time, and the bytes one solve sweep
        # reads (factors actually touched + vectors) over the solve time
        "dense_executed_flops": work["dense_flops"],
        "dense_tflops": round(work["dense_flops"] / (tim["dense_ms"] * 1e9), 3),
        "dense_frac_of_fp64_peak": round(work["dense_flops"] / (tim["dense_ms"] * 1e9) / PEAK_FP64_TFLOPS, 4),
        "dense_launches": int(work["dense_launches"]),
        "diag_lu_launches": int(work["diag_lu_launches"]),
        "solve_executed_bytes_per_rhs": work["solve_bytes"],
        "solve_hbm_gbs": round(work["solve_bytes"] / (tim["solve_ms"] * 1e6), 1),
        "solve_frac_of_hbm": round(work["solve_bytes"] / (tim["solve_ms"] * 1e-3) / PEAK_HBM_BYTES, 4),
        # nominal counts (zero tiles included, as the reference's dense
        # blocks would compute them): not achieved work
        "nominal_flops": {"dense": counts["f_dense"], "leaf": counts["f_leaf"]},
        "nominal_solve_bytes_per_rhs": counts["b_solve"],
        "solves_per_s_nrhs1": round(solves1, 2),
        "solves_per_s_nrhs16": round(solves16, 2),
        "backward_error": berr,
        "backward_error_nrhs16": berr16,
        "order_s": round(order_s, 3),
        "construct_s": round(construct_s, 3),
        "roofline": roofline,
    }
    if not args.no_cpu:
        line["cpu_baseline"], xo = cpu_baseline(A, order, sizes, np_, b1[:, 0], args.cpu_threads)
        line["rel_diff_vs_oracle"] = float(np.max(np.abs(x - xo)) / np.max(np.abs(xo)))
    legs = ((args.no_precision, "precision", precision_leg), (args.no_batch, "batch", batch_leg),
            (args.no_serve, "serving", lambda S_, np_: serve_leg(S_, np_, args.serve_rate, args.serve_requests)))
    for flag, key, leg in legs:
        if flag:
            continue
        try:
            line[key] = leg(S, np)
        except Exception as e:  # report the failure of a side leg, keep the main line
            line[key] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(line), flush=True)


def profiled_traffic(tile):
    """HBM bytes per launch of the dominant GEMM instantiation (tile 64 or
    128) on this workload, from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (scaled by the counters' reading for a copy kernel of
    known size); None if not recorded."""
    import re
    path = os.path.join(ROOT, "profiles", "k_gemm_64cube_7pt_fp64.md")
    try:
        m = re.search(r"traffic_bytes_per_launch_tile%d:\s*([0-9.eE+]+)" % tile, open(path).read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def precision_leg(S, np, reps=3):
    """96^3 27-point anisotropic operator: plain FP64 vs FP32 factors with
    FP64 refinement (factor and solve times are the best of reps).  The FP32
    factorization runs its GEMM tiles as three-piece bf16 splits
    (the default); the same FP32 path on the FP32 MFMA is timed beside it."""
    A, grid, np_ = S.config_matrix("c3")
    _, b = S.gen_rhs(A, 1, 1)
    b = b[:, 0]
    out = {"workload": "96^3 27-point anisotropic 1:10:100, 512 ND leaves, 1 rhs",
           "n": A.n, "nnz": A.nnz}
    xs = {}
    for key, prec, form in (("fp64", "fp64", "split"), ("fp32", "fp32", "split"),
                            ("fp32_f32_mfma", "fp32", "f32")):
        print(f"precision leg: {key}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        # refinement stops at the metric's backward-error target (1e-12)
        M = S.SnuMat(A, np_, grid=grid, device=0, precision=prec, ir_tol=IR_TOL, fp32_gemm=form)
        construct = time.perf_counter() - t0
        d_b = S.DeviceBuffer(8 * A.n)
        d_x = S.DeviceBuffer(8 * A.n)
        d_b.upload(np.asfortranarray(b.reshape(-1, 1)))
        fms, sms, steps, ibe, phases = [], [], 0, 0.0, None
        for _ in range(reps):
            M.factorize()
            ft = M.timings()
            if not fms or ft["factor_ms"] < min(fms):
                phases = {k: round(ft[k], 3) for k in
                          ["scatter_ms", "leaf_fronts_ms", "leaf_schur_ms", "assemble_ms", "dense_ms"]}
            fms.append(ft["factor_ms"])
            M.solve_device_ptr(d_b.ptr.value, d_x.ptr.value, 1)
            t = M.timings()
            sms.append(t["solve_ms"])
            steps, ibe = t["ir_steps"], t["ir_berr"]
        # Schur-update GEMM throughput of this precision (one profiled factorization)
        M.set_profiling(True)
        M.factorize()
        M.set_profiling(False)
        pt = M.timings()
        gs = M.gemm_split()
        x = np.zeros(A.n)
        d_x.download(x)
        xs[key] = x
        out[key] = {"factor_ms": round(min(fms), 3), "solve_ms": round(min(sms), 3), "phases_ms": phases,
                     "factorize_solve_per_s": round(1e3 / (min(fms) + min(sms)), 4),
                     "backward_error": S.backward_error(A, x, b), "construct_s": round(construct, 2),
                     "device_gb": round(M.info()["device_bytes"] / 1e9, 2),
                     "schur_gemm": {"ms": round(pt["gemm_sub_ms"], 3), "launches": int(pt["gemm_sub_launches"]),
                                    "tflops": round(pt["gemm_sub_gflop"] / max(pt["gemm_sub_ms"], 1e-9), 3)}}
        for t in (64, 128):
            tt = {"ms": round(gs[f"t{t}_ms"], 3), "launches": int(gs[f"t{t}_launches"]),
                  "tflops": round(gs[f"t{t}_gflop"] / max(gs[f"t{t}_ms"], 1e-9), 3)}
            if prec == "fp32" and form == "split":
                # six bf16 products per FP32 multiply-add: the MFMA work executed
                tt["form"] = "FP32 operands as three bf16 pieces, six products on v_mfma_f32_16x16x32_bf16"
                tt["bf16_mfma_tflops"] = round(6 * tt["tflops"], 3)
                tt["frac_of_bf16_peak"] = round(6 * tt["tflops"] / 2500.0, 4)
                tt["fp32_equiv_frac_of_fp32_peak"] = round(tt["tflops"] / 157.3, 4)
            out[key]["schur_gemm"][f"tile{t}"] = tt
        if prec == "fp32":
            out[key]["gemm_form"] = ("split: FP32 operands as three bf16 pieces, six products, FP32 accumulation"
                                     if form == "split" else "f32: FP32 MFMA")
            out[key]["ir_steps"] = int(steps)
            out[key]["ir_backward_error"] = ibe
        M.close()
        d_b.free()
        d_x.free()
    out["rel_diff_fp32ir_vs_fp64"] = float(np.max(np.abs(xs["fp32"] - xs["fp64"])) / np.max(np.abs(xs["fp64"])))
    out["speedup_fp32ir_vs_fp64"] = round(out["fp32"]["factorize_solve_per_s"] / out["fp64"]["factorize_solve_per_s"], 3)
    out["speedup_fp32ir_f32_mfma_vs_fp64"] = round(
        out["fp32_f32_mfma"]["factorize_solve_per_s"] / out["fp64"]["factorize_solve_per_s"], 3)
    out["rel_diff_split_vs_f32_mfma"] = float(np.max(np.abs(xs["fp32"] - xs["fp32_f32_mfma"]))
                                              / np.max(np.abs(xs["fp64"])))
    return out


def serve_matrix(S, rng, kind):
    """One matrix of the serving mix: kind 0 a 3D 7-point grid (side 32..80),
    kind 1 a 2D 5-point grid (side 256..1024), convection-diffusion recipe."""
    if kind == 0:
        d = int(rng.integers(32, 81))
        A = S.gen_grid(d, d, d, 7, 0, 6.1, int(rng.integers(1 << 30)))
        return A, (d, d, d), S.batch_np(A.n)
    d = int(rng.integers(256, 1025))
    A = S.gen_grid(d, d, 1, 5, 0, 4.1, int(rng.integers(1 << 30)))
    return A, (d, d, 1), S.batch_np(A.n)


def serve_leg(S, np, rate, nreq, budget=200e9):
    """Serving: the offered rate is stated in the output; by default half the
    capacity measured by a closed-loop burst, so latency reflects service
    rather than an unbounded queue."""
    import threading
    rng = np.random.default_rng(4)
    mix = [serve_matrix(S, rng, kind) for kind in (0, 1) for _ in range(8)]
    out = {"workload": "16 resident FP32+IR factorizations (8 3D 7-pt sides 32..80, 8 2D 5-pt sides 256..1024), "
                       "Poisson arrivals (rate below), nrhs 2^U(0..6), 5% new factorizations, 200 GB budget",
           "requests": nreq}
    print("serving leg: admitting 16 matrices", file=sys.stderr, flush=True)
    srv = S.Server(budget_bytes=budget, slots=2, max_cols=256)
    t0 = time.perf_counter()
    ids = [srv.add(A, np_, grid=g, precision="fp32", ir_tol=IR_TOL) for A, g, np_ in mix]
    out["admit_s"] = round(time.perf_counter() - t0, 2)
    mats = {i: A for i, (A, _, _) in zip(ids, mix)}
    arr = np.random.default_rng(5)
    # capacity: a closed-loop burst of requests drawn like the stream's
    burst = []
    for _ in range(64):
        mid = ids[int(arr.integers(len(ids)))]
        burst.append((mid, arr.standard_normal((mats[mid].n, 1 << int(arr.integers(0, 7))))))
    t0 = time.perf_counter()
    for t in [srv.submit(mid, b) for mid, b in burst]:
        srv.wait(t)
    cap = len(burst) / (time.perf_counter() - t0)
    srv.stats(reset=True)
    if rate <= 0:
        rate = 0.5 * cap
    nreq = max(40, min(nreq, int(60 * rate)))  # about a minute of arrivals at most
    out["capacity_req_per_s"] = round(cap, 2)
    out["rate_per_s"] = round(rate, 2)
    out["requests"] = nreq
    gaps = arr.exponential(1.0 / rate, nreq)
    tickets, adds, lock = [], [], threading.Lock()

    def admit(A, g, np_):
        mid = srv.add(A, np_, grid=g, precision="fp32", ir_tol=IR_TOL)
        with lock:
            mats[mid] = A

    print("serving leg: request stream", file=sys.stderr, flush=True)
    start = time.perf_counter()
    due = start
    for q in range(nreq):
        due += gaps[q]
        dt = due - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        if arr.random() < 0.05:
            A, g, np_ = serve_matrix(S, arr, int(arr.integers(2)))
            th = threading.Thread(target=admit, args=(A, g, np_))
            th.start()
            adds.append(th)
            continue
        with lock:
            live = list(mats.keys())
        mid = live[int(arr.integers(len(live)))]
        k = 1 << int(arr.integers(0, 7))
        b = arr.standard_normal((mats[mid].n, k))
        try:
            tickets.append((mid, b, srv.submit(mid, b)))
        except S.SnuError:
            pass  # evicted between the pick and the submit
    worst = 0.0
    cols = 0
    for mid, b, t in tickets:
        x, _ = srv.wait(t)
        cols += b.shape[1]
        if len(tickets) < 50 or t % 10 == 0:
            worst = max(worst, S.backward_error(mats[mid], x, b))
    for th in adds:
        th.join()
    wall = time.perf_counter() - start
    st = srv.stats()
    srv.close()
    out.update({"answered": int(st["requests"]), "columns": cols, "packed_solves": int(st["packed_solves"]),
                "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3), "mean_ms": round(st["mean_ms"], 3),
                "solves_per_s": round(st["requests"] / wall, 2), "columns_per_s": round(cols / wall, 2),
                "new_factorizations": len(adds), "evictions": int(st["evictions"]),
                "worst_backward_error_sampled": worst, "wall_s": round(wall, 2)})
    out["under_pressure"] = serve_pressure(S, mix, 150)
    return out


def serve_pressure(S, mix, nreq):
    """The serving mix under a budget of half its device bytes, requests
    uniform over all 16 matrices: a request for a parked matrix waits for its
    admission (upload + factorization) on the submitting thread.  Offered at
    half the capacity a closed-loop burst measures under the same budget.
    Latency is end to end: from the request's arrival to its answer."""
    import ctypes
    import numpy as np
    need = []
    for A, g, np_ in mix:
        with S.SnuMat(A, np_, grid=g, device=-1, precision="fp32") as M:
            nb = ctypes.c_size_t()
            S._check(S.lib().snumat_bytes_needed(M._h, ctypes.byref(nb)))
            need.append(nb.value)
    budget = int(0.5 * sum(need))
    print("serving leg: under a binding budget", file=sys.stderr, flush=True)
    srv = S.Server(budget_bytes=budget, slots=2, max_cols=256)
    ids = [srv.add(A, np_, grid=g, precision="fp32", ir_tol=IR_TOL) for A, g, np_ in mix]
    arr = np.random.default_rng(6)

    def draw():
        key = int(arr.integers(len(mix)))
        return key, arr.standard_normal((mix[key][0].n, 1 << int(arr.integers(0, 7))))

    t0 = time.perf_counter()
    burst = [draw() for _ in range(32)]
    for t in [srv.submit(ids[key], b) for key, b in burst]:
        srv.wait(t)
    cap = len(burst) / (time.perf_counter() - t0)
    rate = 0.5 * cap
    nreq = max(20, min(nreq, int(40 * rate)))
    srv.stats(reset=True)
    adm0 = srv.stats()
    gaps = arr.exponential(1.0 / rate, nreq)
    tickets = []
    start = time.perf_counter()
    due = start
    for q in range(nreq):
        due += gaps[q]
        dt = due - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        key, b = draw()
        t = srv.submit(ids[key], b)
        tickets.append((key, b, t, time.perf_counter() - due))  # arrival -> queued
    lat, worst, cols = [], 0.0, 0
    for key, b, t, queued in tickets:
        x, ms = srv.wait(t)
        lat.append(1e3 * queued + ms)
        cols += b.shape[1]
        if t % 10 == 0:
            worst = max(worst, S.backward_error(mix[key][0], x, b))
    wall = time.perf_counter() - start
    st = srv.stats()
    srv.close()
    lat = np.sort(np.array(lat))
    return {"budget_gb": round(budget / 1e9, 2), "resident_set_gb": round(sum(need) / 1e9, 2),
            "capacity_req_per_s": round(cap, 2), "rate_per_s": round(rate, 2), "requests": nreq, "columns": cols,
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_ms": round(float(lat.mean()), 3), "solves_per_s": round(nreq / wall, 2),
            "readmissions": int(st["readmissions"] - adm0["readmissions"]),
            "evictions": int(st["evictions"] - adm0["evictions"]),
            "worst_backward_error_sampled": worst, "wall_s": round(wall, 2)}


def batch_leg(S, np, count=512):
    """512 2D 9-point matrices factorized + solved as one batch vs one by one."""
    mem = S.config_batch(count)
    bs = [S.gen_rhs(A, 1, 3)[1][:, 0] for A, _, _ in mem]
    out = {"workload": f"{count} 2D 9-point convection-diffusion, sides log-uniform [16,192], 1 rhs each",
           "total_n": int(sum(A.n for A, _, _ in mem))}
    for prec in ("fp64", "fp32"):
        print(f"batch leg: {prec}", file=sys.stderr, flush=True)
        mats = [S.SnuMat(A, np_, grid=(side, side, 1), device=0, precision=prec, ir_tol=IR_TOL)
                for A, side, np_ in mem]
        with S.Batch(mats) as B:
            B.factorize()
            B.solve(bs)
            S.device_sync()
            t0 = time.perf_counter()
            reps = 3
            for _ in range(reps):
                B.factorize()
                xs = B.solve(bs)
            S.device_sync()
            batched = (time.perf_counter() - t0) / reps
            bt = B.timings()
            fl, sl = B.launches()
            # the same batch with the factorization replayed from a HIP graph
            # (captured on the first factorize after switching it on)
            B.use_graph(True)
            B.factorize()
            B.solve(bs)
            S.device_sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                B.factorize()
                xg = B.solve(bs)
            S.device_sync()
            graphed = (time.perf_counter() - t0) / reps
            gt = B.timings()
            B.use_graph(False)
        same = all(np.array_equal(x, y) for x, y in zip(xs, xg))
        worst = max(S.backward_error(A, x, b) for (A, _, _), x, b in zip(mem, xs, bs))
        t0 = time.perf_counter()
        for M, b in zip(mats, bs):
            M.factorize()
            M.solve(b)
        single = time.perf_counter() - t0
        for M in mats:
            M.close()
        out[prec] = {"batch_factorize_solve_s": round(batched, 4), "one_by_one_s": round(single, 4),
                     "matrices_per_s": round(count / batched, 2), "speedup_vs_one_by_one": round(single / batched, 2),
                     "batch_factor_ms": round(bt["factor_ms"], 3), "batch_solve_ms": round(bt["solve_ms"], 3),
                     "dense_factor_launches": fl, "dense_solve_launches": sl,
                     "graph_factorize_solve_s": round(graphed, 4), "graph_factor_ms": round(gt["factor_ms"], 3),
                     "graph_results_equal_eager": bool(same),
                     "worst_backward_error": worst}
    return out


def host_cpu():
    """The host cores this job may use: the CPUs of its affinity mask,
    limited by its cgroup CPU quota and by OMP_NUM_THREADS when the
    environment sets one (the job's share of a shared host); plus what they
    are (all figures are reported)."""
    info = {"os_cpu_count": os.cpu_count()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["affinity_cpus"] = os.cpu_count() or 1
    cores = info["affinity_cpus"]
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            info["cgroup_quota_cpus"] = round(int(quota) / int(period), 2)
            cores = min(cores, max(1, int(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        info["omp_num_threads"] = int(omp)
        cores = min(cores, int(omp))
    info["cpu_model"] = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, info


def cpu_baseline(A, order, sizes, np_, b, threads):
    """The oracle's MKL host path (reference offlvl = -1 semantics), one full
    factor+solve of the same workload, on every host core this job may use
    (leaves one per thread, dense levels through MKL's threads)."""
    usable, info = host_cpu()
    cores = threads or usable
    os.environ["MKL_NUM_THREADS"] = str(cores)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, dense_backend

    O = Oracle(A, np_, order, sizes)
    t0 = time.perf_counter()
    x = O.factor_solve(b, nthreads=cores)
    dt = time.perf_counter() - t0
    O.close()
    return ({
        "value": round(1.0 / dt, 5),
        "unit": "factorize+solve/s",
        "cores": cores,
        "host": info,
        "kind": "port",
        "sample": "one full factorize+solve of the same 64^3 workload (oracle C restatement, "
                  f"leaves on {cores} threads, dense levels through "
                  f"{'MKL' if dense_backend() else 'fallback loops'} with {cores} threads), "
                  f"{dt:.1f} s",
        "seconds": round(dt, 3),
    }, x)


if __name__ == "__main__":
    main()
