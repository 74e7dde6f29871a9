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
exponential with the given rate (the first request arrives at 0); n =
round(exp(U(ln nmin, ln nmax))); E ~ U[1,200]; nu ~ U[0.2,0.4]; pressure ~
U[0.1,10].  The load is pressure times the unit-pressure consistent nodal
load of the top face.
"""
from __future__ import annotations

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
    gaps = rng.exponential(1.0 / rate, count) if rate > 0 else np.zeros(count)
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

    def get(self, n: int) -> dict:
        if n in self._d:
            self._d.move_to_end(n)
            return self._d[n]
        m = M.box_hex20(n, n, n, 1.0, 1.0, 1.0)
        rest = M.base_fixed_rest(m)
        nf, neq = M.steering(m.nn, 3, rest)
        nodes, vals = M.top_pressure_loads(m, -1.0)
        r_unit = M.nodal_loads_to_rhs(nf, nodes, vals, neq)
        entry = {"g_coord": m.g_coord(), "g_num": m.g_num, "rest": rest, "nn": m.nn, "neq": neq,
                 "nels": m.nels, "r_unit": r_unit}
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
                limit: int = 100000):
    """One request by itself on `ctx` (blocking): (x, PcgResult)."""
    c = cache.get(req.n)
    ctx.mesh_set(c["g_coord"])
    ctx.steer_build(c["g_num"], c["rest"], c["nn"])
    ctx.assemble_elastic(req.e, req.v, None, store="fp64" if precision == "fp64" else "fp32")
    x, res = ctx.pcg_solve(c["r_unit"] * req.pressure, limit, tol, stop="relres", precision=precision,
                           check_every=64)
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
    join and leave the batch between chunks of iterations."""
    slots: int = 32
    hbm_budget: float = 240e9
    chunk: int = 16
    precision: str = "mixed64"
    tol: float = 1e-10
    limit: int = 100000
    keep_x: bool = False
    device: int = 0
    cache: CubeCache = field(default_factory=CubeCache)

    def __post_init__(self):
        if not 1 <= self.slots <= BATCH_MAX:
            raise ValueError(f"1..{BATCH_MAX} slots")
        self._slots = [_Slot(Context(self.device)) for _ in range(self.slots)]
        self._batch = Batch(self.device)

    def close(self):
        self._batch.close()
        for s in self._slots:
            s.ctx.close()

    def _device_bytes(self) -> int:
        """Device memory the slots hold now: live buffers and each context's
        cache of freed blocks."""
        return sum(s.ctx.mem_used() for s in self._slots)

    def _fits(self, slot: _Slot, need: int, busy: bool) -> bool:
        """Whether a request needing `need` bytes can start on `slot`:
        within the HBM budget after handing back the slot's cached blocks,
        then those of idle slots, if needed.  With nothing running a request
        is always admitted."""
        if self._device_bytes() + need <= self.hbm_budget:
            return True
        slot.ctx.trim()
        if self._device_bytes() + need <= self.hbm_budget:
            return True
        for s in self._slots:
            if s.req is None:
                s.ctx.trim()
        return self._device_bytes() + need <= self.hbm_budget or not busy

    def _admit(self, s: _Slot, req: Request, now: float):
        c = self.cache.get(req.n)
        ctx = s.ctx
        ctx.mesh_set(c["g_coord"])
        ctx.steer_build(c["g_num"], c["rest"], c["nn"])      # steering and maps built on the device
        ctx.assemble_elastic(req.e, req.v, None, store="fp64" if self.precision == "fp64" else "fp32")
        ctx.pcg_begin(c["r_unit"] * req.pressure, self.limit, self.tol, stop="relres", precision=self.precision)
        s.req, s.t_admit, s.bytes = req, now, device_bytes(req.n, self.precision)

    def serve(self, requests: list[Request], log=None) -> list[Result]:
        """Run the stream in real time (arrival times relative to the start).

        Each cycle enqueues a chunk of iterations for the current slots,
        takes a state snapshot behind it and enqueues one more chunk, so the
        GPU keeps working while the host waits for the snapshot; when it
        arrives, finished solves are read back and retired, waiting requests
        admitted, and the next cycle runs the new set."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        queue: deque[Request] = deque()
        free = list(self._slots)
        active: list[_Slot] = []
        out: list[Result] = []
        by_ctx = {s.ctx.h: s for s in self._slots}
        done: list[tuple] = []           # (slot, result) finished, still to read back
        in_flight = False
        t0 = time.perf_counter()
        while pending or queue or active or done:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            if not in_flight:
                while queue and free:
                    need = device_bytes(queue[0].n, self.precision)
                    if not self._fits(free[-1], need, bool(active)):
                        break
                    s = free.pop()
                    self._admit(s, queue.popleft(), time.perf_counter() - t0)
                    active.append(s)
                if active:
                    ctxs = [s.ctx for s in active]
                    self._batch.run(ctxs, self.chunk)
                    self._batch.query()              # snapshot behind this chunk
                    self._batch.run(ctxs, self.chunk)   # and one more chunk queued
                    in_flight = True
            # read back the solves that finished in the previous snapshot
            # while the GPU works on the chunks just queued
            for s, r in done:
                x = s.ctx.pcg_x()
                t = time.perf_counter() - t0
                q = s.req
                out.append(Result(q.rid, q.n, s.ctx.neq, r.iters, bool(r.converged), r.refinements,
                                  s.t_admit, t, t - q.arrival, x if self.keep_x else None,
                                  float(np.abs(x).sum())))
                if log:
                    log(f"request {q.rid} n={q.n} iters={r.iters} latency={t - q.arrival:.3f}s")
                s.ctx.reset()
                s.req = None
                free.append(s)
            done = []
            got = self._batch.query() if in_flight else None
            if got is None:
                if not active and pending:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - (time.perf_counter() - t0))))
                elif in_flight:
                    time.sleep(20e-6)
                continue
            in_flight = False
            snap, res = got
            for ctx, r in zip(snap, res):
                s = by_ctx[ctx.h]
                # converged, broken down (status 6) or at the limit: the solve
                # is frozen on the device; anything else is still running
                if r.converged or r.status == 6 or r.iters >= self.limit:
                    active.remove(s)
                    done.append((s, r))
        return sorted(out, key=lambda r: r.rid)


def serve_sequential(requests: list[Request], precision: str = "fp64", tol: float = 1e-10,
                     device: int = 0, keep_x: bool = False, cache: CubeCache | None = None,
                     log=None) -> list[Result]:
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
            x, res = solve_alone(ctx, cache, r, precision, tol)
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
