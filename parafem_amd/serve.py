"""A request loop that serves many EBE-PCG solves of uneven size on one GPU.

Each request is a p121-style hex20 elastic cube (n x n x n elements, unit
edges, base fixed, uniform pressure on the top face) with its own E, nu and
pressure.  The loop admits arrived requests in arrival order while a context
("slot") and the HBM budget allow (counting what every context holds,
cached blocks included, and handing cached blocks back when they are in the
way), builds the problem on that slot (mesh upload, steering and maps built
on the device, element matrices assembled there) and begins its PCG.  All
admitted solves then run as one batch (pf_batch): every launch of the
iteration's kernels -- element products, scatter, folds, updates --
advances every slot, each slot's blocks doing what its own launch would.
The host drives the batch from one thread in chunks of iterations (a
replayed HIP graph), keeping one chunk queued behind a non-blocking state
snapshot; when the snapshot shows a solve finished, its displacements are
read back, its device buffers freed, the slot takes the next request and
the next chunk runs the new set.

The arithmetic of a solve does not depend on what else runs: every kernel is
deterministic, a solve's blocks do the work of its own launch and its state
is its own, so a result is bit-identical alone and in the mix.

Stream generator (``make_stream``): numpy ``default_rng(seed)``; arrival gaps
exponential with the given rate (unit-rate draws scaled by 1/rate; all zero
for rate 0, the closed loop, which therefore holds the same requests; the
first request arrives at 0); n =
round(exp(U(ln nmin, ln nmax))); E ~ U[1,200]; nu ~ U[0.2,0.4]; pressure ~
U[0.1,10].  The load is pressure times the unit-pressure consistent nodal
load of the top face.
"""
from __future__ import annotations

import queue as queue_mod
import threading
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field

import numpy as np

from . import mesh as M
from .api import BATCH_MAX, Batch, Context


@dataclass
class Request:
    rid: int
    arrival: float      # seconds after the start of the stream
    n: int
    e: float
    v: float
    pressure: float


@dataclass
class Result:
    rid: int
    n: int
    neq: int
    iters: int
    converged: bool
    refinements: int
    t_admit: float
    t_done: float
    latency: float
    x: np.ndarray | None = None
    checksum: float = 0.0


def make_stream(count: int, rate: float, seed: int = 4, nmin: int = 10, nmax: int = 80) -> list[Request]:
    rng = np.random.default_rng(seed)
    # unit-rate gaps are drawn whatever the rate, so one seed gives the same
    # requests closed-loop (rate 0) and at any arrival rate
    unit = rng.exponential(1.0, count)
    gaps = unit / rate if rate > 0 else np.zeros(count)
    arrival = np.concatenate([[0.0], np.cumsum(gaps[1:])])
    n = np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), count))).astype(int)
    e = rng.uniform(1.0, 200.0, count)
    v = rng.uniform(0.2, 0.4, count)
    pr = rng.uniform(0.1, 10.0, count)
    return [Request(i, float(arrival[i]), int(n[i]), float(e[i]), float(v[i]), float(pr[i])) for i in range(count)]


class CubeCache:
    """Host-side mesh (coordinates, g_num, rest) and unit load of the n^3
    cube, kept for the most recently used sizes; the steering itself is built
    on the device at admission."""

    def __init__(self, keep: int = 24):
        self.keep = keep
        self._d: OrderedDict[int, dict] = OrderedDict()
        self._lock = threading.Lock()     # set-up threads share the cache

    def get(self, n: int) -> dict:
        with self._lock:
            if n in self._d:
                self._d.move_to_end(n)
                return self._d[n]
        # built outside the lock (two threads missing the same size both build it)
        m = M.box_hex20(n, n, n, 1.0, 1.0, 1.0)
        rest = M.base_fixed_rest(m)
        nf, neq = M.steering(m.nn, 3, rest)
        nodes, vals = M.top_pressure_loads(m, -1.0)
        r_unit = M.nodal_loads_to_rhs(nf, nodes, vals, neq)
        entry = {"g_coord": m.g_coord(), "g_num": m.g_num, "rest": rest, "nn": m.nn, "neq": neq,
                 "nels": m.nels, "r_unit": r_unit}
        with self._lock:
            self._d[n] = entry
            while len(self._d) > self.keep:
                self._d.popitem(last=False)
        return entry


def device_bytes(n: int, precision: str) -> int:
    """Peak device bytes of one request: coordinates (480 B per element),
    steering (240), transposed steering (240 + 4 per equation), FP64
    element products (480) and the PCG vectors (12 per equation), plus the
    element matrices as stored -- FP64 (28,800 B per element) or the packed
    symmetric FP32 hi and lo (2 x 7,328) with FP32 products (240); FP32
    storage is assembled straight into hi/lo, with no FP64 copy."""
    nels = n ** 3
    neq = 3 * ((2 * n + 1) * (n + 1) * (n + 1) + (n + 1) * (n + 1) * n)   # upper bound
    per_el = 480 + 240 + 240 + 480
    per_el += 28800 if precision == "fp64" else 2 * 7328 + 240
    return nels * per_el + neq * 8 * 12 + neq * 4 + (16 << 20)


def solve_alone(ctx: Context, cache: CubeCache, req: Request, precision: str, tol: float,
                limit: int = 100000, refine_delta: float = 0.0):
    """One request by itself on `ctx` (blocking): (x, PcgResult)."""
    c = cache.get(req.n)
    ctx.mesh_set(c["g_coord"])
    ctx.steer_build(c["g_num"], c["rest"], c["nn"])
    ctx.assemble_elastic(req.e, req.v, None, store="fp64" if precision == "fp64" else "fp32")
    x, res = ctx.pcg_solve(c["r_unit"] * req.pressure, limit, tol, stop="relres", precision=precision,
                           check_every=64, refine_delta=refine_delta)
    ctx.reset()
    return x, res


@dataclass
class _Slot:
    ctx: Context
    req: Request | None = None
    t_admit: float = 0.0
    bytes: int = 0


@dataclass
class Server:
    """Continuous batching: every admitted solve is a slot of one pf_batch,
    so each launch of the iteration's kernels advances all of them; solves
    join and leave the batch between chunks of iterations.  Set-up of
    arriving requests and read-back of finished ones run on other host
    threads (the library's calls release the interpreter lock), so the
    thread driving the batch never waits for them."""
    slots: int = 32
    hbm_budget: float = 240e9
    chunk: int = 16
    precision: str = "mixed64"
    tol: float = 1e-10
    limit: int = 100000
    keep_x: bool = False
    device: int = 0
    cache: CubeCache = field(default_factory=CubeCache)
    # batches of at most this many elements in total run each chunk as one
    # persistent cooperative kernel (pf_batch_run_persistent) instead of a
    # replayed graph of launches: small solves are launch-bound
    persistent_elems: int = 0
    # solves of at least this many elements run their own launches inside
    # the batch's runs (pf_batch_solo_set; 0: every solve shares the batched
    # launches): a large solve fills the device alone, and its own kernels
    # skip the per-workgroup slot lookup.  Off by default: on the [10, 80]
    # stream it changed the capacity by less than the run-to-run spread.
    solo_elems: int = 0
    # independent batches, each on its own stream, the admitted solves dealt
    # to the one with the fewest elements (one batch measured fastest: two
    # or three lost 2-12% of the closed-loop capacity on [10, 80] and
    # [10, 40] streams)
    lanes: int = 1
    # host threads setting up arriving requests and reading back finished
    # ones, each on its own slot's context and stream
    workers: int = 1
    # mixed solvers: replace the recurrence residual by the true one each
    # time it has fallen by this factor (the library's default is 1e-2; on
    # uniform-material cubes 0.1 took 5-11% fewer iterations at n = 40-60,
    # and 1e-4 up to 24 times more)
    refine_delta: float = 0.1

    def __post_init__(self):
        if not 1 <= self.slots <= BATCH_MAX:
            raise ValueError(f"1..{BATCH_MAX} slots")
        self._slots = [_Slot(Context(self.device)) for _ in range(self.slots)]
        if self.lanes < 1:
            raise ValueError("at least one lane")
        if self.workers < 1:
            raise ValueError("at least one set-up thread")
        self._stats_lock = threading.Lock()
        self._batches = [Batch(self.device) for _ in range(self.lanes)]
        for b in self._batches:
            b.solo_set(self.solo_elems)
        # host-side time of the last serve() on the set-up threads (summed
        # over them): set-up of requests, read-backs of finished solves;
        # cache trims (each waits for the device) and batch cycles on the
        # driving thread
        self.stats = {"admit_s": 0.0, "readback_s": 0.0, "trims": 0, "cycles": 0}

    def close(self):
        for b in self._batches:
            b.close()
        for s in self._slots:
            s.ctx.close()

    def _device_bytes(self, with_worker: set) -> int:
        """Device memory the slots hold now: live buffers and each context's
        cache of freed blocks, read from the contexts this thread owns, and
        the estimate of what a context on the worker holds (it is not
        touched from two threads)."""
        return sum(s.bytes if id(s) in with_worker else s.ctx.mem_used() for s in self._slots)

    def _fits(self, slot: _Slot, need: int, busy: bool, idle: list[_Slot], with_worker: set) -> bool:
        """Whether a request needing `need` bytes can start on `slot`:
        within the HBM budget after handing back the cached blocks of the
        slot, then of the other idle slots, if needed (an idle context holds
        only cached blocks; nothing is handed back when that cannot make
        room).  With nothing running a request is always admitted."""
        used = self._device_bytes(with_worker)
        if used + need <= self.hbm_budget:
            return True
        order = [slot] + [t for t in idle if t is not slot]
        held = [s.ctx.mem_used() for s in order]
        if used - sum(held) + need <= self.hbm_budget:
            for s, h in zip(order, held):
                if not h:
                    continue
                s.ctx.trim()
                self.stats["trims"] += 1
                if self._device_bytes(with_worker) + need <= self.hbm_budget:
                    return True
        return not busy

    def _admit(self, s: _Slot, req: Request):
        c = self.cache.get(req.n)
        ctx = s.ctx
        ctx.mesh_set(c["g_coord"])
        ctx.steer_build(c["g_num"], c["rest"], c["nn"])      # steering and maps built on the device
        ctx.assemble_elastic(req.e, req.v, None, store="fp64" if self.precision == "fp64" else "fp32")
        ctx.pcg_begin(c["r_unit"] * req.pressure, self.limit, self.tol, stop="relres", precision=self.precision,
                      refine_delta=self.refine_delta)

    def _worker(self, jobs: queue_mod.Queue, results: queue_mod.Queue):
        """Admissions and retirements, one at a time, on their slot's
        context (no slot is in the batch while it is here)."""
        while True:
            job = jobs.get()
            if job is None:
                return
            kind, s, payload, t0 = job
            tb = time.perf_counter()
            try:
                if kind == "admit":
                    self._admit(s, payload)
                    s.t_admit = tb - t0
                    with self._stats_lock:
                        self.stats["admit_s"] += time.perf_counter() - tb
                    results.put(("admitted", s, None))
                else:
                    r = payload
                    x = s.ctx.pcg_x()
                    t = time.perf_counter() - t0
                    q = s.req
                    res = Result(q.rid, q.n, s.ctx.neq, r.iters, bool(r.converged), r.refinements, s.t_admit, t,
                                 t - q.arrival, x if self.keep_x else None, float(np.abs(x).sum()))
                    s.ctx.reset()
                    with self._stats_lock:
                        self.stats["readback_s"] += time.perf_counter() - tb
                    results.put(("retired", s, res))
            except Exception as exc:              # handed to the driving thread
                results.put(("error", s, exc))

    def serve(self, requests: list[Request], log=None) -> list[Result]:
        """Run the stream in real time (arrival times relative to the start).

        For each batch the driving thread enqueues a chunk of iterations for
        its current slots, takes a state snapshot behind it and enqueues one
        more chunk, so the GPU keeps working while it waits for the
        snapshot; when it arrives, finished solves leave the batch (the
        worker reads them back and frees their buffers), requests the worker
        has set up join the batch with the fewest elements, and the next
        cycle runs the new set."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting: deque[Request] = deque()
        free = list(self._slots)
        nb = len(self._batches)
        active: list[list[_Slot]] = [[] for _ in range(nb)]
        ready: list[_Slot] = []          # set up by the worker, not yet in a batch
        busy_worker = 0                  # jobs handed to the worker and not back yet
        with_worker: set = set()         # id() of the slots the worker holds
        out: list[Result] = []
        by_ctx = {s.ctx.h: s for s in self._slots}
        in_flight = [False] * nb
        self.stats = {"admit_s": 0.0, "readback_s": 0.0, "trims": 0, "cycles": 0}
        jobs: queue_mod.Queue = queue_mod.Queue()
        results: queue_mod.Queue = queue_mod.Queue()
        workers = [threading.Thread(target=self._worker, args=(jobs, results), daemon=True)
                   for _ in range(self.workers)]
        for w in workers:
            w.start()
        t0 = time.perf_counter()

        def elems(lane):
            return sum(s.req.n ** 3 for s in active[lane])

        try:
            while pending or waiting or any(active) or ready or busy_worker:
                now = time.perf_counter() - t0
                while pending and pending[0].arrival <= now:
                    waiting.append(pending.popleft())
                # hand arrived requests to the worker while slots and memory allow
                while waiting and free:
                    need = device_bytes(waiting[0].n, self.precision)
                    if not self._fits(free[-1], need, bool(any(active) or ready or busy_worker), free, with_worker):
                        break
                    s = free.pop()
                    s.req = waiting.popleft()
                    # while the worker builds the problem the context still
                    # holds its cache of freed blocks, which the new buffers
                    # may not reuse: count both
                    s.bytes = s.ctx.mem_used() + need
                    with_worker.add(id(s))
                    jobs.put(("admit", s, s.req, t0))
                    busy_worker += 1
                # what the worker finished
                while True:
                    try:
                        kind, s, res = results.get_nowait()
                    except queue_mod.Empty:
                        break
                    busy_worker -= 1
                    with_worker.discard(id(s))
                    if kind == "error":
                        raise res
                    if kind == "admitted":
                        ready.append(s)
                    else:
                        out.append(res)
                        if log:
                            log(f"request {res.rid} n={res.n} iters={res.iters} latency={res.latency:.3f}s")
                        s.req = None
                        free.append(s)
                # set-up solves join the idle batches, the smallest first
                idle = [k for k in range(nb) if not in_flight[k]]
                while ready and idle:
                    k = min(idle, key=elems)
                    active[k].append(ready.pop(0))
                progressed = False
                for k in range(nb):
                    if in_flight[k] or not active[k]:
                        continue
                    ctxs = [s.ctx for s in active[k]]
                    pers = elems(k) <= self.persistent_elems
                    b = self._batches[k]
                    b.run(ctxs, self.chunk, persistent=pers)
                    b.query()                        # snapshot behind this chunk
                    b.run(ctxs, self.chunk, persistent=pers)   # and one more chunk queued
                    in_flight[k] = True
                    self.stats["cycles"] += 1
                for k in range(nb):
                    got = self._batches[k].query() if in_flight[k] else None
                    if got is None:
                        continue
                    progressed = True
                    in_flight[k] = False
                    snap, res_list = got
                    for ctx, r in zip(snap, res_list):
                        s = by_ctx[ctx.h]
                        # converged, broken down (status 6) or at the limit:
                        # the solve is frozen on the device; anything else
                        # still runs
                        if r.converged or r.status == 6 or r.iters >= self.limit:
                            active[k].remove(s)
                            s.bytes = s.ctx.mem_used()       # freeing keeps the blocks in its cache
                            with_worker.add(id(s))
                            jobs.put(("retire", s, r, t0))
                            busy_worker += 1
                if not progressed:
                    if not any(in_flight) and not busy_worker and pending and not waiting:
                        time.sleep(max(0.0, min(0.01, pending[0].arrival - (time.perf_counter() - t0))))
                    else:
                        time.sleep(20e-6)
        finally:
            for _ in workers:
                jobs.put(None)
            for w in workers:
                w.join()
        return sorted(out, key=lambda r: r.rid)


def serve_sequential(requests: list[Request], precision: str = "fp64", tol: float = 1e-10,
                     device: int = 0, keep_x: bool = False, cache: CubeCache | None = None,
                     log=None, refine_delta: float = 0.0) -> list[Result]:
    """The same stream one request at a time (blocking solves, arrival order,
    real time)."""
    cache = cache or CubeCache()
    ctx = Context(device)
    out = []
    t0 = time.perf_counter()
    try:
        for r in sorted(requests, key=lambda q: q.arrival):
            now = time.perf_counter() - t0
            if r.arrival > now:
                time.sleep(r.arrival - now)
            ta = time.perf_counter() - t0
            x, res = solve_alone(ctx, cache, r, precision, tol, refine_delta=refine_delta)
            t = time.perf_counter() - t0
            out.append(Result(r.rid, r.n, x.size, res.iters, bool(res.converged), res.refinements, ta, t,
                              t - r.arrival, x if keep_x else None, float(np.abs(x).sum())))
            if log:
                log(f"request {r.rid} n={r.n} iters={res.iters} latency={t - r.arrival:.3f}s")
    finally:
        ctx.close()
    return out


def summarize(results: list[Result], wall: float) -> dict:
    lat = np.array([r.latency for r in results])
    return {
        "requests": len(results),
        "solves_per_s": len(results) / wall if wall > 0 else None,
        "wall_s": wall,
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()),
        "all_converged": all(r.converged for r in results),
        "iterations": int(sum(r.iters for r in results)),
    }
