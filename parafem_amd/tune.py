"""Autotuner for the launch shapes of the element-product kernels.

The shapes (C ABI ``pf_tune``):

* single-RHS product ``k_ebe``: ``ebe_epl``, the number of elements a lane
  group loads before it multiplies (more bytes in flight per wavefront
  against more registers and fewer resident wavefronts);
* block products ``k_blk_ebe_sym`` (FP32) and ``k_blk_ebe`` (FP64):
  ``blk_waves``, elements (one wavefront each) per workgroup, tuned on the
  FP32 product; ``blk_rows``, 16-row output tiles per pass of the FP64
  product (more accumulators against more re-reads of the gathered block;
  the FP32 product forms all row tiles in one pass), left at its default.

No shape changes a result: every lane forms its row in the same order
whatever the shape, so the tuner only measures time.  Each candidate is
timed with HIP events around its own launches on the problem loaded in the
context, the fastest is set on the context, and the choice is cached in a
JSON file keyed by device (architecture name and compute units), kernel,
ntot, precision, k and the element count's power of two.  The cache file is
``$PARAFEM_AMD_TUNE_CACHE`` or ``~/.cache/parafem_amd/tune.json``; an
unwritable location keeps the results in memory only.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

EPL_CANDIDATES = {60: (1, 2), 24: (1, 2, 4), 8: (1, 2, 4, 8)}
BLOCK_WAVES = (1, 2, 4)
# every block shape the library accepts (waves x FP64 rows per pass)
BLOCK_SHAPES = tuple((w, r) for w in BLOCK_WAVES for r in (1, 2, 4))

_memory: dict[str, dict] = {}


def cache_path() -> str:
    return os.environ.get("PARAFEM_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "parafem_amd", "tune.json"))


def _load() -> dict:
    try:
        with open(cache_path()) as fh:
            data = json.load(fh)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(key: str, entry: dict):
    _memory[key] = entry
    data = _load()
    data[key] = entry
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def lookup(key: str) -> dict | None:
    if key in _memory:
        return _memory[key]
    return _load().get(key)


def cache_key(arch: str, cus: int, kernel: str, ntot: int, precision: str, k: int, nels: int) -> str:
    size = int(round(math.log2(max(nels, 1))))
    return f"{arch}|cu{cus}|{kernel}|ntot{ntot}|{precision}|k{k}|nels~2^{size}"


def _key(ctx, kernel: str, precision: str, k: int) -> str:
    arch, cus = ctx.device_info()
    return cache_key(arch, cus, kernel, ctx.ntot, precision, k, ctx.nels)


def pick(times: dict) -> object:
    """The fastest candidate; ties go to the earlier (simpler) one."""
    best = None
    for cand, t in times.items():
        if best is None or t < times[best] * (1 - 1e-3):
            best = cand
    return best


def tune_ebe(ctx, r0, precision: str, reps: int = 20, warm: int = 3, use_cache: bool = True) -> dict:
    """Pick ``ebe_epl`` for the PCG iteration's element product of
    `precision` on the problem loaded in `ctx` (right-hand side r0), set it
    on the context and return {"ebe_epl", "ms", "cached"}.  The PCG state
    is left at a tol-0 run: call pcg_begin/pcg_solve afterwards."""
    cur = ctx.tune_get()
    if precision == "mixed" and ctx.scatter_mode == "ranges":
        # the range-fused kernel has one shape (an element per lane group)
        return {"ebe_epl": cur["ebe_epl"], "ms": {}, "cached": True}
    key = _key(ctx, "k_ebe", precision, 1)
    hit = lookup(key) if use_cache else None
    if hit is not None:
        ctx.tune_set(hit["ebe_epl"], cur["blk_waves"], cur["blk_rows"])
        return {**hit, "cached": True}
    cands = EPL_CANDIDATES[ctx.ntot]
    ctx.pcg_begin(r0, len(cands) * (warm + reps) + 10, 0.0, stop="relres", precision=precision)
    times = {}
    for e in cands:
        ctx.tune_set(e, cur["blk_waves"], cur["blk_rows"])
        ctx.pcg_run(warm)
        ctx.timing_enable(True)
        ctx.pcg_run(reps)
        ms, n = ctx.timing_get()["ebe"]
        ctx.timing_enable(False)
        times[e] = ms / max(n, 1)
    best = pick(times)
    ctx.tune_set(best, cur["blk_waves"], cur["blk_rows"])
    entry = {"ebe_epl": int(best), "ms": {str(e): round(t, 5) for e, t in times.items()}}
    _store(key, entry)
    return {**entry, "cached": False}


def tune_block(ctx, reps: int = 10, use_cache: bool = True) -> dict:
    """Pick ``blk_waves`` for the FP32 block element product on the problem
    loaded in `ctx` (FP64 element matrices), set it (``blk_rows`` keeps the
    context's value) and return {"blk_waves", "ms", "cached"}."""
    key = _key(ctx, "k_blk_ebe_sym", "fp32", 64)
    cur = ctx.tune_get()
    hit = lookup(key) if use_cache else None
    if hit is not None:
        ctx.tune_set(cur["ebe_epl"], hit["blk_waves"], cur["blk_rows"])
        return {**hit, "cached": True}
    # block buffers, FP32 tiles and a block of operands with full mantissas
    # (the matrix cores draw less power, and clock higher, on zeros)
    ctx.block_apply(np.random.default_rng(0).uniform(-1.0, 1.0, (ctx.neq, 64)), "mixed")
    times = {}
    for w in BLOCK_WAVES:
        ctx.tune_set(cur["ebe_epl"], w, cur["blk_rows"])
        times[w] = ctx.block_time_ebe(reps)
    w = pick(times)
    ctx.tune_set(cur["ebe_epl"], w, cur["blk_rows"])
    entry = {"blk_waves": int(w), "ms": {str(a): round(t, 5) for a, t in times.items()}}
    _store(key, entry)
    return {**entry, "cached": False}
