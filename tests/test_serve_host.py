"""Host logic of the serving loop (no device): the request stream and the
admission estimate."""
import numpy as np
import pytest

from parafem_amd import serve as S


def test_stream_holds_the_same_requests_at_every_rate():
    closed = S.make_stream(200, 0.0, seed=4, nmin=10, nmax=80)
    timed = S.make_stream(200, 1.5, seed=4, nmin=10, nmax=80)
    assert [(r.n, r.e, r.v, r.pressure) for r in closed] == [(r.n, r.e, r.v, r.pressure) for r in timed]
    assert all(r.arrival == 0.0 for r in closed)
    t = np.array([r.arrival for r in timed])
    assert t[0] == 0.0 and np.all(np.diff(t) > 0)
    # mean gap 1/rate (200 draws: within 20%)
    assert abs(np.diff(t).mean() * 1.5 - 1.0) < 0.2


def test_stream_ranges():
    reqs = S.make_stream(500, 0.0, seed=7, nmin=10, nmax=80)
    n = np.array([r.n for r in reqs])
    assert n.min() >= 10 and n.max() <= 80
    # log-uniform: about as many requests below the geometric mean as above
    assert abs((n < np.sqrt(10 * 80)).mean() - 0.5) < 0.1
    assert all(1.0 <= r.e <= 200.0 and 0.2 <= r.v <= 0.4 and 0.1 <= r.pressure <= 10.0 for r in reqs)


def test_device_bytes_estimate():
    # FP32 storage holds hi + lo packed (2 x 7,328 B per element) instead of
    # the 28,800 B FP64 matrix, and grows with the element count
    for n in (10, 40, 80):
        assert S.device_bytes(n, "mixed64") < S.device_bytes(n, "fp64")
    assert S.device_bytes(80, "mixed64") > 8 * S.device_bytes(40, "mixed64") * 0.9
    nels = 80 ** 3
    assert S.device_bytes(80, "fp64") > nels * 28800


class _Res:
    def __init__(self, iters, converged):
        self.iters, self.converged, self.status, self.refinements = iters, converged, 0, 0


class _Device:
    """Device memory of the mock contexts: what each holds live and in its
    cache of freed blocks (never reused, the worst case), against a fixed
    capacity."""

    def __init__(self, capacity):
        import threading
        self.capacity, self.lock, self.peak, self.ctxs = capacity, threading.Lock(), 0, []

    def total(self):
        return sum(c.live + c.cache for c in self.ctxs)


def _mock_classes(dev):
    import itertools

    ids = itertools.count(1)

    class Ctx:
        def __init__(self, device=0):
            self.h, self.neq, self.it, self.target, self.live, self.cache, self.n = next(ids), 0, 0, None, 0, 0, 0
            dev.ctxs.append(self)

        def mesh_set(self, g):
            pass

        def steer_build(self, g_num, rest, nn):
            self.neq = nn
            return nn, None

        def assemble_elastic(self, e, v, ee, store="fp64"):
            with dev.lock:
                need = S.device_bytes(self.n, "mixed64")
                if dev.total() + need > dev.capacity:
                    raise MemoryError("device out of memory")
                self.live = need
                dev.peak = max(dev.peak, dev.total())

        def pcg_begin(self, r, limit, tol, **kw):
            self.it, self.target = 0, 20 + 3 * self.n

        def pcg_x(self):
            return np.ones(self.neq)

        def reset(self):
            with dev.lock:
                self.cache += self.live
                self.live, self.target = 0, None

        def trim(self):
            with dev.lock:
                self.cache = 0

        def mem_used(self):
            return self.live + self.cache

        def close(self):
            pass

    class Batch:
        def __init__(self, device=0):
            self._last, self._snap = [], None

        def close(self):
            pass

        def solo_set(self, min_elems):
            pass

        def run(self, ctxs, n, graph=True, persistent=False):
            for c in ctxs:
                assert c.target is not None, "a slot in the batch that is not set up"
                c.it = min(c.target, c.it + n)
            self._last = list(ctxs)

        def query(self):
            if self._snap is None:
                self._snap = (self._last, [_Res(c.it, c.it >= c.target) for c in self._last])
                return None
            snap, self._snap = self._snap, None
            return snap

    class Cache(S.CubeCache):
        def get(self, n):
            return {"g_coord": None, "g_num": None, "rest": None, "nn": n, "r_unit": np.ones(n)}

    return Ctx, Batch, Cache


@pytest.mark.parametrize("lanes,workers", [(1, 1), (2, 1), (1, 3)])
def test_admission_counts_cached_blocks_of_slots_being_set_up(monkeypatch, lanes, workers):
    """A slot handed to a set-up thread still holds its cache of freed
    blocks while the new problem is allocated: admission must count both, so
    the device never overflows, whatever order the sizes come in (mock
    contexts whose freed blocks are never reused); with one batch or two,
    one set-up thread or several, every request is served once."""
    budget = 3 * S.device_bytes(80, "mixed64")
    dev = _Device(budget)
    Ctx, Batch, Cache = _mock_classes(dev)
    monkeypatch.setattr(S, "Context", Ctx)
    monkeypatch.setattr(S, "Batch", Batch)
    reqs = S.make_stream(120, 0.0, seed=9, nmin=10, nmax=80)
    for r in reqs:
        r.n = int(r.n)
    srv = S.Server(slots=8, hbm_budget=budget, chunk=16, cache=Cache(), lanes=lanes, workers=workers)

    orig_admit = S.Server._admit

    def admit(self, s, req):
        s.ctx.n = req.n
        orig_admit(self, s, req)

    monkeypatch.setattr(S.Server, "_admit", admit)
    out = srv.serve(reqs)
    assert sorted(r.rid for r in out) == list(range(120))
    assert dev.peak <= budget
    assert srv.stats["trims"] > 0          # the budget was reached and caches handed back
    # ... but only caches that hold something, and only when that makes room
    assert srv.stats["trims"] <= len(reqs)


def test_server_and_one_by_one_use_the_same_replacement_threshold(monkeypatch):
    """The server's mixed solves and serve_sequential's pass the same
    residual-replacement threshold to the library (the serving tests compare
    the two bit for bit)."""
    seen = []

    class Ctx:
        neq = 4

        def __init__(self, device=0):
            pass

        def mesh_set(self, g):
            pass

        def steer_build(self, g_num, rest, nn):
            return nn, None

        def assemble_elastic(self, e, v, ee, store="fp64"):
            pass

        def pcg_begin(self, r, limit, tol, **kw):
            seen.append(("server", kw.get("refine_delta")))

        def pcg_solve(self, r, limit, tol, **kw):
            seen.append(("alone", kw.get("refine_delta")))
            return np.zeros(4), _Res(3, True)

        def reset(self):
            pass

        def close(self):
            pass

    class Cache(S.CubeCache):
        def get(self, n):
            return {"g_coord": None, "g_num": None, "rest": None, "nn": 4, "r_unit": np.ones(4)}

    monkeypatch.setattr(S, "Context", Ctx)
    reqs = S.make_stream(3, 0.0, seed=1, nmin=2, nmax=3)
    S.serve_sequential(reqs, precision="mixed64", cache=Cache(), refine_delta=S.Server.refine_delta)
    srv = S.Server.__new__(S.Server)                 # only _admit is exercised: no batch, no slots
    srv.cache, srv.precision, srv.limit, srv.tol, srv.refine_delta = Cache(), "mixed64", 100, 1e-10, 0.1
    srv._admit(S._Slot(Ctx()), reqs[0])
    assert {d for _, d in seen} == {0.1} and {k for k, _ in seen} == {"server", "alone"}
