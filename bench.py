#!/usr/bin/env python3
"""EBE-PCG benchmark on one MI355X.

Workload: the p121 hex20 linear-elastic cube of 100 x 100 x 100 = 1,000,000
20-node bricks (12,180,600 equations), unit edges, base fixed, uniform unit
pressure on the top face, per-element Young's modulus log-normal
(mu = ln 100, sigma = 0.5, SplitMix64/Box-Muller seed 1), nu = 0.3, element
matrices assembled on the device and stored in FP64, Jacobi-preconditioned
PCG in FP64.  A step is one PCG iteration (gather -> element products ->
scatter -> dots -> recurrences -> stopping test) with every input resident
in HBM.  The stopping tolerance is 0 so every timed step does full work.

Prints one JSON line (see the "roofline" and "cpu_baseline" keys for the
dominant kernel's achieved HBM bandwidth and the CPU oracle's rate).
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


def algorithmic_bytes(nels, ntot, neq, s_k=8):
    """Per-iteration algorithmic HBM bytes: every element matrix once, the
    int32 steering once, and 15 FP64 vector streams of length neq."""
    per_elem = ntot * ntot * s_k + ntot * 4
    return nels * per_elem, nels * per_elem + 15 * neq * 8, per_elem


def cpu_baseline(n_sample: int, budget_s: float = 12.0):
    """The CPU oracle's FP64 EBE-PCG on an n^3 sample of the same workload,
    all host threads (capped at 16), rate scaled to the 10^6-element cube."""
    import oracle as O
    from parafem_amd import problems as P
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")), 16)
    pr = P.elastic_cube(n_sample, lognormal_e=True)
    K = O.assemble_elastic(pr.g_coord(), pr.e, pr.v, pr.e_elem, nthreads=threads)
    dg = O.diag(pr.g_g, K, pr.neq)

    def run(n):
        t = time.perf_counter()
        O.pcg(pr.g_g, K, pr.r0, dg, 0.0, n, stop=O.STOP_RELRES, nthreads=threads,
              dot_order=O.DOT_THREADS)
        return time.perf_counter() - t

    t1 = run(2)
    per = max(run(4) - t1, 1e-6) / 2
    n = max(4, int(budget_s / per))
    t = run(n + 2) - t1
    rate = n / t                                   # iterations/s of the sample
    return {
        "value": rate * pr.nels / 1_000_000,
        "unit": "iter/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n_sample}^3 hex20 cube ({pr.nels} elements, {pr.neq} equations), {n} FP64 "
                  f"PCG iterations in {t:.1f} s = {rate:.2f} iter/s, scaled by elements to "
                  f"1,000,000",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=100, help="elements per cube edge")
    ap.add_argument("--store", default="fp64", choices=["fp64"])
    ap.add_argument("--cpu-sample", type=int, default=24, help="edge of the CPU-baseline cube (0 = skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the streaming copy kernel over MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU: the multi-GPU element "
                          "sharding is a design only (see DESIGN.md); use --gpus 1"}))
        sys.exit(2)

    from parafem_amd import Context, DeviceArray, problems as P

    if args.copy_calibration:
        ctx = Context(0)
        n = args.copy_calibration * 2**20 // 8
        a, b = DeviceArray(ctx, n), DeviceArray(ctx, n)
        for _ in range(5):
            ctx.dev_copy(b, a)
        ctx.sync()
        print(json.dumps({"copy_bytes": n * 8, "launches": 5}), flush=True)
        return

    t0 = time.perf_counter()
    pr = P.elastic_cube(args.n, lognormal_e=True)
    log(f"problem built: nels={pr.nels} neq={pr.neq} ({time.perf_counter() - t0:.1f} s)")
    ctx = Context(0)
    t0 = time.perf_counter()
    ctx.mesh_set(pr.g_coord())
    ctx.steer_set(pr.g_g, pr.neq)
    ctx.sync()
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    ctx.assemble_elastic(pr.e, pr.v, pr.e_elem, store=args.store)
    ctx.sync()
    t_asm = time.perf_counter() - t0
    log(f"set-up {t_setup:.2f} s, assembly {t_asm * 1e3:.1f} ms, device memory {ctx.mem_used() / 2**30:.1f} GiB")

    limit = args.warmup + 2 * args.steps + 100
    ctx.pcg_begin(pr.r0, limit, 0.0, stop="relres")
    ctx.pcg_run(args.warmup)
    ctx.sync()
    log("warm-up done")
    t0 = time.perf_counter()
    ctx.pcg_run(args.steps)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    res = ctx.pcg_poll()
    assert res.iters == args.warmup + args.steps, res
    assert res.status != 6, "breakdown"
    log(f"timed {args.steps} iterations: {elapsed * 1e3:.1f} ms")

    # Per-kernel HIP-event timing on the context's stream (separate pass).
    nprof = min(args.steps, 50)
    ctx.timing_enable(True)
    ctx.pcg_run(nprof)
    tk = ctx.timing_get()
    ctx.timing_enable(False)
    ebe_ms = tk["ebe"][0] / max(tk["ebe"][1], 1)
    iter_ms = sum(v[0] for v in tk.values()) / nprof
    b_ebe, b_iter, per_elem = algorithmic_bytes(pr.nels, pr.ntot, pr.neq)
    achieved = b_ebe / (ebe_ms * 1e-3) / 1e9
    log("kernel ms/launch: " + ", ".join(f"{k}={v[0] / max(v[1], 1):.3f}" for k, v in tk.items()))

    cpu = None
    if args.cpu_sample > 0:
        log("cpu baseline ...")
        cpu = cpu_baseline(args.cpu_sample)
        log(f"cpu baseline {cpu['value']:.3f} iter/s ({cpu['cores']} threads)")

    value = args.steps / elapsed
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"p121 hex20 elastic cube {args.n}^3 ({pr.nels} elements, {pr.neq} equations), "
                        "log-normal E (seed 1), FP64 storkm, Jacobi EBE-PCG iteration",
            "nels": pr.nels, "neq": pr.neq, "store": args.store, "precision": "fp64",
            "iteration_gbs": round(b_iter / (iter_ms * 1e-3) / 1e9, 1),
            "iteration_bytes": b_iter,
            "kernel_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in tk.items()},
            "assembly_ms": round(t_asm * 1e3, 1),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_ebe_f64<60> (gather + element products)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_element": per_elem,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
