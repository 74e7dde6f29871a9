#!/usr/bin/env python3
"""EBE-PCG benchmark on one MI355X.

Workload: the p121 hex20 linear-elastic cube of 100 x 100 x 100 = 1,000,000
20-node bricks (12,180,600 equations), unit edges, base fixed, uniform unit
pressure on the top face, per-element Young's modulus log-normal
(mu = ln 100, sigma = 0.5, SplitMix64/Box-Muller seed 1), nu = 0.3, element
matrices assembled on the device.

Headline (value): iterations/s of the mixed-precision Jacobi EBE-PCG --
element matrices stored as FP32 hi/lo in a packed symmetric layout, element
products from hi in FP32,
vectors, dots and recurrences in FP64, reliable updates with the FP64
(hi+lo) residual.  A step is one PCG iteration (gather -> element products
-> scatter -> dots -> recurrences -> stopping test) with every input
resident in HBM; the stopping tolerance is 0 during the timed steps so every
step does full work.  The same measurement of the plain FP64 path, both
times-to-solution at relative residual 1e-10 (the solves/s half of the
metric) and the CPU oracle's rate are reported beside it.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "EBE-PCG iter/s + % HBM peak, hex20 elastic; mixed-prec solves/s, p99 latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ebe_bytes(ntot, store, precision="mixed"):
    """Algorithmic HBM bytes of one element product: the element matrix once
    (full FP64, or the FP32 hi part packed symmetric: the ntot(ntot+1)/2
    entries of one triangle padded to a multiple of 4, hi and lo for the
    split precision) and the int32 steering once."""
    if store == "fp64":
        return ntot * ntot * 8 + ntot * 4
    parts = 2 if precision == "split" else 1
    packed = (ntot * (ntot + 1) // 2 + 3) // 4 * 4
    return parts * packed * 4 + ntot * 4


def iteration_bytes(nels, ntot, neq, store, precision="mixed"):
    """Per-iteration algorithmic bytes: element matrices + steering + 15 FP64
    vector streams of length neq."""
    return nels * ebe_bytes(ntot, store, precision) + 15 * neq * 8


def ebe_kernel_name(ntot, precision, epl):
    """The iteration's element-product kernel as rocprofv3 names it."""
    kind, out = {"fp64": (0, "double"), "mixed": (1, "float"), "split": (2, "double"),
                 "mixed64": (3, "double")}[precision]
    return f"k_ebe<{ntot}, {kind}, true, {out}, false, {epl}>"


def time_iterations(ctx, pr, precision, steps, warmup):
    """Iterations/s with tol = 0 (no early exit), then a per-kernel pass with
    HIP events on the context's stream."""
    limit = warmup + steps + 200
    ctx.pcg_begin(pr.r0, limit, 0.0, stop="relres", precision=precision)
    ctx.pcg_run(warmup)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.pcg_run(steps)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    res = ctx.pcg_poll()
    if res.status == 6 or res.iters != warmup + steps:
        raise RuntimeError(f"{precision}: unexpected state {res}")
    nprof = min(steps, 40)
    ctx.timing_enable(True)
    ctx.pcg_run(nprof)
    tk = ctx.timing_get()
    ctx.timing_enable(False)
    return elapsed, tk, nprof


def solve(ctx, pr, precision, store, tol):
    from parafem_amd import problems as P, tune as T
    t0 = time.perf_counter()
    r = P.setup(ctx, pr, store=store)
    ctx.sync()
    t_setup = time.perf_counter() - t0
    T.tune_ebe(ctx, r, precision)          # the shape tuned for this path (cached)
    t0 = time.perf_counter()
    x, res = ctx.pcg_solve(r, 100000, tol, stop="relres", precision=precision, check_every=64)
    t_solve = time.perf_counter() - t0
    y = ctx.apply(x)
    true_rel = float(np.linalg.norm(y - pr.r0) / np.linalg.norm(pr.r0))
    return x, {"iters": res.iters, "refinements": res.refinements, "converged": res.converged,
               "time_s": round(t_solve, 3), "setup_assembly_s": round(t_setup, 3),
               "solves_per_s": round(1.0 / t_solve, 5), "true_relres": true_rel}


def profiled_traffic(workload, kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of this workload (calibrated against a
    plain copy), or None when there is none."""
    path = os.path.join(HERE, "profiles", "hbm_traffic.json")
    try:
        with open(path) as fh:
            entry = json.load(fh)[workload]
        return entry["bytes_per_dispatch"].get(kernel), entry["source"]
    except (OSError, KeyError, ValueError):
        return None, None


def build_problem(workload, n):
    from parafem_amd import problems as P
    if workload == "hex20":
        return P.elastic_cube(n, lognormal_e=True)
    return P.laplace_cube(n, lognormal_k=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload: str, n_sample: int, nels_full: int, budget_s: float = 12.0):
    """The CPU oracle's FP64 EBE-PCG on an n^3 sample of the same workload
    with OMP_NUM_THREADS threads (the host cores granted to this GPU's
    process: 16 per GPU on the MI355X hosts here, whatever nproc shows), rate
    scaled by elements to the full problem.  nproc and the CPU model are
    reported beside it."""
    import oracle as O
    nproc = os.cpu_count() or 1
    threads = min(nproc, int(os.environ.get("OMP_NUM_THREADS", str(nproc))))
    pr = build_problem(workload, n_sample)
    if pr.kind == "elastic":
        K = O.assemble_elastic(pr.g_coord(), pr.e, pr.v, pr.e_elem, nthreads=threads)
    else:
        K = O.assemble_laplace(pr.g_coord(), pr.kx, pr.ky, pr.kz, pr.k_elem, nthreads=threads)
    dg = O.diag(pr.g_g, K, pr.neq)

    def run(n):
        t = time.perf_counter()
        O.pcg(pr.g_g, K, pr.r0, dg, 0.0, n, stop=O.STOP_RELRES, nthreads=threads,
              dot_order=O.DOT_THREADS)
        return time.perf_counter() - t

    run(2)                                   # thread pool and caches warm
    k = 4
    while True:                              # a pilot run of at least 0.5 s
        t = run(k)
        if t >= 0.5 or k >= 4096:
            break
        k *= 2
    n = max(4, min(int(budget_s * k / t), 100000))
    log(f"cpu baseline: {k} iterations in {t:.2f} s, timing {n}")
    t = run(n)
    rate = n / t
    return {
        "value": rate * pr.nels / nels_full,
        "unit": "iter/s",
        "cores": threads,
        "nproc": nproc,
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": f"{n_sample}^3 {workload} cube ({pr.nels} elements, {pr.neq} equations), {n} FP64 "
                  f"PCG iterations in {t:.1f} s = {rate:.2f} iter/s, scaled by elements to "
                  f"{nels_full:,}",
    }


def block_bench(args):
    """64 load cases at once: the matrix-core block solve (FP32 MFMA block
    PCG, FP64 residual replacement) to relres 1e-10 per case, against plain
    FP64 single-RHS solves of the same cases one after another (a sample of
    them timed, the time scaled to all 64)."""
    from parafem_amd import Context, problems as P, tune as T
    n = args.n or 60
    k = 64
    t0 = time.perf_counter()
    pr, F = P.block_cube(n, k)
    log(f"block problem: nels={pr.nels} neq={pr.neq} k={k} ({time.perf_counter() - t0:.1f} s)")
    ctx = Context(0)
    P.setup(ctx, pr, store="fp64")
    ctx.sync()
    shape = T.tune_block(ctx)
    log(f"block: autotuned {shape['blk_waves']} elements per workgroup "
        f"({'cache' if shape['cached'] else 'timed ' + str(shape['ms'])})")
    ctx.block_solve(F[:, :1], 1e-4, outer_limit=1)          # warm: packing, allocation, code paths
    t0 = time.perf_counter()
    X, res = ctx.block_solve(F, 1e-10)
    t_block = time.perf_counter() - t0
    log(f"block solve: {t_block:.3f} s, {res.inner_iters} iterations, {res.outer} residual replacements, "
        f"converged {res.converged}/{k}, max relres {res.max_relres:.3e}")
    # the kernels' own durations inside the iteration: HIP events around
    # every launch of the same solve run again (the operands' bits move the
    # matrix cores' clock, so a short or synthetic run would not represent it)
    ctx.timing_enable(True)
    ctx.block_solve(F, 1e-10)
    tk = ctx.timing_get()
    ctx.timing_enable(False)
    ebe_ms = tk["ebe"][0] / max(tk["ebe"][1], 1)
    log("block iteration kernels (ms per launch): " +
        ", ".join(f"{kk}={v[0] / max(v[1], 1):.4f}" for kk, v in tk.items() if v[1]))
    btraffic, bsource = profiled_traffic("block", "k_blk_ebe_sym<60>") if n == 60 else (None, None)
    flops = 2.0 * pr.ntot * pr.ntot * k * pr.nels
    tflops = flops / (ebe_ms * 1e-3) / 1e12
    kt_bytes = pr.nels * 2560 * 4       # the 10 stored 16x16 tiles of the upper tile triangle
    ut_bytes = pr.nels * pr.ntot * k * 4
    log(f"FP32 MFMA element product: {ebe_ms:.3f} ms, {tflops:.1f} TFLOP/s useful, "
        f"{(kt_bytes + ut_bytes) / (ebe_ms * 1e-3) / 1e9:.0f} GB/s (K tiles + U_e)")
    t_one = []
    iters_one = []
    for c in range(min(args.fp64_sample, k)):
        t1 = time.perf_counter()
        x, r = ctx.pcg_solve(F[:, c], 100000, 1e-10, stop="relres", check_every=64)
        t_one.append(time.perf_counter() - t1)
        iters_one.append(r.iters)
        err = float(np.abs(X[:, c] - x).max() / np.abs(x).max())
        log(f"FP64 single solve {c}: {t_one[-1]:.3f} s, {r.iters} iterations, block vs single {err:.2e}")
    t_seq = float(np.mean(t_one)) * k if t_one else None
    out = {
        "metric": METRIC,
        "value": round(k / t_block, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": k,
        "warmup": 1,
        "ms_per_step": round(t_block / k * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"multi-RHS: hex20 elastic cube {n}^3 ({pr.nels} elements, {pr.neq} equations), "
                        f"E=100, nu=0.3, {k} load cases (top-face traction components U[-1,1] per node, "
                        "seed 3), FP32 MFMA block PCG with FP64 residual replacement to relres 1e-10 per case",
            "block_solve_s": t_block, "replacements": res.outer, "iterations": res.inner_iters,
            "converged": res.converged, "max_relres": res.max_relres,
            "fp64_sequential_s_estimated": t_seq, "fp64_sample_solves": len(t_one),
            "fp64_iterations": iters_one,
            "speedup_vs_fp64_sequential": round(t_seq / t_block, 3) if t_seq else None,
            "tile_shape": {"blk_waves": shape["blk_waves"], "source": "autotuner"},
        },
        "roofline": {
            "bound": "mfma", "kernel": "k_blk_ebe_sym<60> (v_mfma_f32_16x16x4f32)",
            "achieved": round(tflops, 2), "peak": 157.3, "unit": "TFLOP/s", "frac": round(tflops / 157.3, 4),
            "traffic": btraffic, "traffic_source": f"profiles/{bsource}" if bsource else None,
            "ebe_ms": ebe_ms,
            "hbm_gbs": round((kt_bytes + ut_bytes) / (ebe_ms * 1e-3) / 1e9, 1),
        },
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def serve_sample(requests: int, nmin: int, nmax: int, precision: str, slots: int = 32):
    """A reduced serving stream for the default line: closed-loop capacity of
    the batched server, then the same requests at Poisson arrivals of 80%
    of it (p50/p99 latency), and the FP64 one-by-one capacity on them."""
    from parafem_amd import serve as S
    cache = S.CubeCache(keep=64)
    closed = S.make_stream(requests, 0.0, seed=4, nmin=nmin, nmax=nmax)
    for r in closed:
        cache.get(r.n)
    srv = S.Server(slots=slots, cache=cache, precision=precision)
    try:
        srv.serve(S.make_stream(4, 0.0, seed=5, nmin=nmin, nmax=nmin + 2))    # warm-up
        t0 = time.perf_counter()
        cap = S.summarize(srv.serve(closed), time.perf_counter() - t0)
        rate = 0.8 * cap["solves_per_s"]
        stream = S.make_stream(requests, rate, seed=4, nmin=nmin, nmax=nmax)
        t0 = time.perf_counter()
        mix = S.summarize(srv.serve(stream), time.perf_counter() - t0)
    finally:
        srv.close()
    t0 = time.perf_counter()
    seq = S.summarize(S.serve_sequential(closed, precision="fp64", cache=cache), time.perf_counter() - t0)
    return {
        "stream": f"{requests} hex20 cube requests, n log-uniform in [{nmin},{nmax}], E U[1,200], nu U[0.2,0.4], "
                  f"pressure U[0.1,10] (seed 4), relres 1e-10, {precision} server with {slots} slots",
        "closed_loop_solves_per_s": round(cap["solves_per_s"], 3),
        "rate_per_s": round(rate, 3),
        "solves_per_s": round(mix["solves_per_s"], 3),
        "p50_latency_s": round(mix["p50_latency_s"], 4),
        "p99_latency_s": round(mix["p99_latency_s"], 4),
        "fp64_one_by_one_solves_per_s": round(seq["solves_per_s"], 3),
        "capacity_vs_fp64_one_by_one": round(cap["solves_per_s"] / seq["solves_per_s"], 3),
        "all_converged": cap["all_converged"] and mix["all_converged"] and seq["all_converged"],
    }


def serve_bench(args):
    """The serving loop on a request stream: closed-loop capacity, then the
    stream at a fixed Poisson rate through the mixed-precision server, then
    the same stream one request at a time on the plain FP64 path (each phase
    can be left out; the FP64 phase can take the first N requests only)."""
    from parafem_amd import serve as S
    phases = set(args.phases.split(","))
    if not phases <= {"closed", "stream", "fp64"} or not phases:
        raise SystemExit("--phases: a comma-separated subset of closed,stream,fp64")
    if "stream" in phases and "closed" not in phases and args.rate <= 0:
        raise SystemExit("the stream phase without the closed loop needs --rate")
    cache = S.CubeCache(keep=80)
    closed = S.make_stream(args.requests, 0.0, seed=4, nmin=args.nmin, nmax=args.nmax)
    for r in closed:
        cache.get(r.n)
    log(f"serve: {args.requests} requests, n in [{args.nmin},{args.nmax}], sizes cached, phases {sorted(phases)}")
    done = [0]

    def progress(msg):
        # one line per 25 finished requests, so a long stream shows progress
        done[0] += 1
        if done[0] % 25 == 0:
            log(f"{done[0]} finished; last: {msg}")

    cap = mix = seq = None
    rate = args.rate
    if phases & {"closed", "stream"}:
        srv = S.Server(slots=args.slots, cache=cache, precision=args.serve_precision, chunk=args.chunk,
                       persistent_elems=args.persistent_elems, solo_elems=args.solo_elems, lanes=args.lanes,
                       workers=args.workers, refine_delta=args.refine_delta)
        try:
            # warm-up: compile/capture paths once
            srv.serve(S.make_stream(args.slots, 0.0, seed=5, nmin=args.nmin, nmax=args.nmin + 2))
            if "closed" in phases:
                done[0] = 0
                t0 = time.perf_counter()
                res_closed = srv.serve(closed, log=progress)
                cap = S.summarize(res_closed, time.perf_counter() - t0)
                log(f"closed loop: {cap}")
                log(f"server host time: {srv.stats}")
                if rate <= 0:
                    rate = 0.8 * cap["solves_per_s"]
            if "stream" in phases:
                stream = S.make_stream(args.requests, rate, seed=4, nmin=args.nmin, nmax=args.nmax)
                done[0] = 0
                t0 = time.perf_counter()
                res_mix = srv.serve(stream, log=progress)
                mix = S.summarize(res_mix, time.perf_counter() - t0)
                log(f"stream ({args.serve_precision} server, {args.slots} slots, {rate:.3f}/s): {mix}")
        finally:
            srv.close()
    if "fp64" in phases:
        n64 = args.fp64_requests if args.fp64_requests > 0 else args.requests
        # the same requests (a seed gives the same ones at every rate), at
        # the stream's arrival times when there is a rate
        part = S.make_stream(args.requests, rate if rate > 0 else 0.0, seed=4, nmin=args.nmin,
                             nmax=args.nmax)[:n64]
        done[0] = 0
        t0 = time.perf_counter()
        res_seq = S.serve_sequential(part, precision="fp64", cache=cache, log=progress)
        seq = S.summarize(res_seq, time.perf_counter() - t0)
        log(f"FP64 one by one ({len(part)} requests): {seq}")
    head = mix or cap or seq
    out = {
        "metric": METRIC,
        "value": round(head["solves_per_s"], 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.requests,
        "warmup": args.slots,
        "ms_per_step": round(1e3 / head["solves_per_s"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"serving stream: {args.requests} hex20 elastic cube requests, n log-uniform in "
                        f"[{args.nmin},{args.nmax}], E U[1,200], nu U[0.2,0.4], pressure U[0.1,10] (seed 4), "
                        + (f"Poisson arrivals at {rate:.2f}/s, " if rate > 0 else "closed loop, ")
                        + f"{args.serve_precision} PCG (FP32-stored element matrices, FP64 vectors, reliable "
                        f"updates) to relres {1e-10}, {args.slots} slots",
            "phases": sorted(phases),
            "value_from": "stream" if mix else ("closed" if cap else "fp64"),
            "rate_per_s": rate if rate > 0 else None,
            "closed_loop": cap,
            "server": mix,
            "fp64_one_by_one": seq,
            "capacity_vs_fp64_one_by_one": round(cap["solves_per_s"] / seq["solves_per_s"], 3)
            if cap and seq else None,
            "p99_speedup_vs_fp64": round(seq["p99_latency_s"] / mix["p99_latency_s"], 3)
            if mix and seq and seq["requests"] == mix["requests"] else None,
        },
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="hex20", choices=["hex20", "hex8", "serve", "block"],
                    help="hex20: p121 elastic cube (default, 100^3); hex8: p123 Laplace cube (200^3); "
                         "serve: a stream of hex20 cube requests through the serving loop; "
                         "block: 64 load cases on a 60^3 hex20 cube through the matrix-core block solve")
    ap.add_argument("--fp64-sample", type=int, default=4,
                    help="block: single-RHS FP64 solves timed for the sequential baseline")
    ap.add_argument("--requests", type=int, default=200, help="serve: stream length")
    ap.add_argument("--nmin", type=int, default=10, help="serve: smallest cube edge")
    ap.add_argument("--nmax", type=int, default=80, help="serve: largest cube edge")
    ap.add_argument("--rate", type=float, default=0.0,
                    help="serve: Poisson arrival rate (0 = 80%% of the measured closed-loop capacity)")
    ap.add_argument("--slots", type=int, default=32, help="serve: concurrent solves (one batch)")
    ap.add_argument("--chunk", type=int, default=16, help="serve: iterations per batch run")
    ap.add_argument("--lanes", type=int, default=1,
                    help="serve: independent batches on their own streams")
    ap.add_argument("--refine-delta", type=float, default=0.1,
                    help="serve: reliable-update threshold of the mixed server (0: the library's 1e-2)")
    ap.add_argument("--workers", type=int, default=1,
                    help="serve: host threads setting up requests and reading back results")
    ap.add_argument("--solo-elems", type=int, default=0,
                    help="serve: solves of at least this many elements run their own launches inside the batch "
                         "(0: none)")
    ap.add_argument("--persistent-elems", type=int, default=0,
                    help="serve: batches of at most this many elements run as persistent kernels")
    ap.add_argument("--phases", default="closed,stream,fp64",
                    help="serve: which of closed (closed-loop capacity), stream (Poisson arrivals) and fp64 "
                         "(the same stream one by one on the FP64 path) to run")
    ap.add_argument("--fp64-requests", type=int, default=0,
                    help="serve: FP64 one-by-one phase on the first N requests only (0 = all)")
    ap.add_argument("--serve-sample", type=int, default=48,
                    help="hex20: requests of the reduced serving stream reported beside the iteration rate "
                         "(0 = skip)")
    ap.add_argument("--serve-precision", default="mixed64", choices=["fp64", "mixed", "mixed64", "split"],
                    help="serve: arithmetic of the element products in the server")
    ap.add_argument("--n", type=int, default=0, help="elements per cube edge (0 = the workload's size)")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"],
                    help="path whose iteration rate is the headline value")
    ap.add_argument("--solve-tol", type=float, default=1e-10,
                    help="relative residual of the time-to-solution runs (0 = skip them)")
    ap.add_argument("--cpu-sample", type=int, default=24, help="edge of the CPU-baseline cube (0 = skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the streaming copy kernel over MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "parafem_amd runs on one GPU: element sharding across GPUs "
                          "is a design only; use --gpus 1"}))
        sys.exit(2)

    from parafem_amd import Context, DeviceArray, problems as P, tune as T

    if args.copy_calibration:
        ctx = Context(0)
        n = args.copy_calibration * 2**20 // 8
        a, b = DeviceArray(ctx, n), DeviceArray(ctx, n)
        for _ in range(5):
            ctx.dev_copy(b, a)
        ctx.sync()
        reps = 50
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.dev_copy(b, a)
        ctx.sync()
        t = (time.perf_counter() - t0) / reps
        # A read-only stream: DOT_PRODUCT_P of the two arrays (each call
        # returns its sum to the host, so the wall clock includes that wait).
        ctx.vec_fill(a, 1.0)
        ctx.vec_fill(b, 0.5)
        ctx.dot_product_p(a, b)
        t1 = time.perf_counter()
        for _ in range(reps):
            ctx.dot_product_p(a, b)
        td = (time.perf_counter() - t1) / reps
        # copy: read n*8 bytes + write n*8 bytes per launch; dot: read 2*n*8
        print(json.dumps({"copy_bytes": n * 8, "launches": 5 + reps, "ms_per_copy": round(t * 1e3, 4),
                          "copy_gbs": round(2 * n * 8 / t / 1e9, 1), "ms_per_dot": round(td * 1e3, 4),
                          "dot_read_gbs": round(2 * n * 8 / td / 1e9, 1)}), flush=True)
        return

    if args.workload == "serve":
        serve_bench(args)
        return
    if args.workload == "block":
        block_bench(args)
        return
    if args.n == 0:
        args.n = 100 if args.workload == "hex20" else 200
    t0 = time.perf_counter()
    pr = build_problem(args.workload, args.n)
    log(f"problem built: nels={pr.nels} neq={pr.neq} ({time.perf_counter() - t0:.1f} s)")
    ctx = Context(0)
    paths = {}
    order = ["fp64", "split", "mixed64", "mixed", "mixed_ranges"] if args.precision == "mixed" \
        else ["mixed_ranges", "mixed", "mixed64", "split", "fp64"]
    for name in order:
        # mixed_ranges: the mixed solver with range-fused products and sums
        prec = "mixed" if name == "mixed_ranges" else name
        store = "fp64" if prec == "fp64" else "fp32"
        ctx.scatter_mode = "ranges" if name == "mixed_ranges" else "slots"
        t0 = time.perf_counter()
        P.setup(ctx, pr, store=store)
        ctx.sync()
        t_setup = time.perf_counter() - t0
        shape = T.tune_ebe(ctx, pr.r0, prec)
        log(f"{name}: autotuned elements per lane group {shape['ebe_epl']} "
            f"({'cache' if shape['cached'] else 'timed ' + str(shape['ms'])})")
        elapsed, tk, nprof = time_iterations(ctx, pr, prec, args.steps, args.warmup)
        ebe_ms = tk["ebe"][0] / max(tk["ebe"][1], 1)
        per_elem = ebe_bytes(pr.ntot, store, prec)
        achieved = pr.nels * per_elem / (ebe_ms * 1e-3) / 1e9
        iter_ms = sum(v[0] for v in tk.values()) / nprof
        paths[name] = {
            "iter_per_s": round(args.steps / elapsed, 3),
            "ms_per_step": round(elapsed / args.steps * 1e3, 4),
            "elapsed_s": elapsed,
            "store": store,
            "kernel_ms_per_iter": {k: round(v[0] / nprof, 4) for k, v in tk.items()},
            "ebe_ms": ebe_ms,
            "ebe_gbs": round(achieved, 1),
            "bytes_per_element": per_elem,
            "iteration_gbs": round(iteration_bytes(pr.nels, pr.ntot, pr.neq, store, prec) / (iter_ms * 1e-3) / 1e9, 1),
            "setup_assembly_s": round(t_setup, 3),
            "device_gib": round(ctx.mem_used() / 2**30, 2),
            "tile_shape": {"ebe_epl": shape["ebe_epl"], "source": "autotuner"},
            "kernel": (f"k_ebe_range<{pr.ntot}> (products and their sums into u, one launch per colour)"
                       if name == "mixed_ranges" else ebe_kernel_name(pr.ntot, prec, shape["ebe_epl"])),
        }
        log(f"{name}: {paths[name]['iter_per_s']} iter/s, element products {ebe_ms:.3f} ms "
            f"({achieved:.0f} GB/s), kernels " +
            ", ".join(f"{k}={v}" for k, v in paths[name]["kernel_ms_per_iter"].items()))
    ctx.scatter_mode = "slots"

    solves = None
    if args.solve_tol > 0:
        solves = {}
        x64, solves["fp64"] = solve(ctx, pr, "fp64", "fp64", args.solve_tol)
        log(f"fp64 solve: {solves['fp64']}")
        for name in ("mixed", "mixed_ranges", "mixed64", "split"):
            prec = "mixed" if name == "mixed_ranges" else name
            ctx.scatter_mode = "ranges" if name == "mixed_ranges" else "slots"
            xm, solves[name] = solve(ctx, pr, prec, "fp32", args.solve_tol)
            solves[name]["rel_err_vs_fp64"] = float(np.abs(xm - x64).max() / np.abs(x64).max())
            solves[f"speedup_{name}_vs_fp64"] = round(solves["fp64"]["time_s"] / solves[name]["time_s"], 3)
            log(f"{name} solve: {solves[name]}")
        ctx.scatter_mode = "slots"

    serving = None
    if args.serve_sample > 0 and args.workload == "hex20":
        ctx.reset()
        ctx.trim()
        log("serving sample ...")
        serving = serve_sample(args.serve_sample, 10, 40, args.serve_precision)
        log(f"serving sample: {serving}")

    cpu = None
    if args.cpu_sample > 0:
        log("cpu baseline ...")
        cpu = cpu_baseline(args.workload, args.cpu_sample if args.workload == "hex20" else 2 * args.cpu_sample,
                           pr.nels)
        log(f"cpu baseline {cpu['value']:.3f} iter/s ({cpu['cores']} threads)")

    head = paths[args.precision]
    mixed = args.precision == "mixed"
    kname = head["kernel"]
    traffic, tsource = profiled_traffic(args.workload, kname) if args.n == (100 if args.workload == "hex20" else 200) \
        else (None, None)
    out = {
        "metric": METRIC,
        "value": head["iter_per_s"],
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"p121 hex20 elastic cube {args.n}^3 ({pr.nels} elements, {pr.neq} equations), "
                         "log-normal E (seed 1)" if args.workload == "hex20" else
                         f"p123 hex8 Laplace cube {args.n}^3 ({pr.nels} elements, {pr.neq} equations), "
                         "log-normal k (seed 2), T=0 on x=0 by penalty, unit flux on x=n")
                        + ", Jacobi EBE-PCG iteration, "
                        + ("packed symmetric FP32-stored element matrices with FP64 vectors and reliable updates"
                           if args.precision == "mixed" else "FP64 element matrices"),
            "nels": pr.nels, "neq": pr.neq, "precision": args.precision,
            "paths": {k: {kk: vv for kk, vv in v.items() if kk != "elapsed_s"} for k, v in paths.items()},
            "solve_relres": args.solve_tol,
            "solves": solves,
            "serving": serving,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname + " element products (" + ("FP32 hi, packed symmetric" if mixed else "FP64") + ")",
            "achieved": head["ebe_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(head["ebe_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": f"profiles/{tsource}" if tsource else None,
            "algorithmic_bytes": pr.nels * head["bytes_per_element"],
            "bytes_per_element": head["bytes_per_element"],
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
