"""Batched iterations (pf_batch_*): several solves of different sizes
advanced by one launch per kernel.  Bars: every solve's iteration count and
displacements bit-identical to the same solve run alone on its context
(pf_pcg_solve), in every precision, with checon and residual stopping; the
FP64 solves also bit-identical to the oracle in the device's reduction order
(p121.f90:89-104 / p123.f90:135-151 per solve)."""
import sys

import numpy as np
import pytest

import oracle as O
from parafem_amd import Batch, Context, problems as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _elastic_mix():
    out = []
    for n, e, v, q in ((3, 100.0, 0.3, 1.0), (5, 37.0, 0.22, 4.0), (2, 150.0, 0.35, 0.3), (4, 80.0, 0.28, 2.5)):
        out.append(P.elastic_cube(n, e=e, v=v, pressure=q))
    out.append(P.elastic_cube(4, lognormal_e=True))
    return out


def _run_batch(ctxs, rs, limit, tol, stop, precision, graph, chunk=7, solo=0):
    """graph: True (replayed graph), False (direct launches) or
    "persistent" (one cooperative kernel per chunk); solo: solves of at
    least that many elements run their own launches in the batch."""
    for c, r in zip(ctxs, rs):
        c.pcg_begin(r, limit, tol, stop=stop, precision=precision)
    b = Batch(0)
    try:
        b.solo_set(solo)
        live = list(range(len(ctxs)))
        done = {}
        while live:
            b.run([ctxs[k] for k in live], chunk, graph=graph is True, persistent=graph == "persistent")
            got = None
            while got is None:
                got = b.query()
            snap, res = got
            assert [c.h for c in snap] == [ctxs[k].h for k in live]
            keep = []
            for k, rr in zip(live, res):
                if rr.converged or rr.status == 6 or rr.iters >= limit:
                    done[k] = (rr.iters, bool(rr.converged), ctxs[k].pcg_x())
                else:
                    keep.append(k)
            live = keep       # the set shrinks as solves finish
        return [done[k] for k in range(len(ctxs))]
    finally:
        b.close()


@pytest.mark.parametrize("precision,stop,graph", [("fp64", "checon", True), ("fp64", "relres", False),
                                                  ("split", "relres", True), ("mixed64", "relres", True),
                                                  ("mixed", "relres", True), ("mixed", "relres", False),
                                                  ("fp64", "checon", "persistent"),
                                                  ("split", "relres", "persistent"),
                                                  ("mixed64", "relres", "persistent"),
                                                  ("mixed", "relres", "persistent")])
def test_batch_equals_alone_hex20(precision, stop, graph):
    probs = _elastic_mix()
    store = "fp64" if precision == "fp64" else "fp32"
    tol = 1e-7 if stop == "checon" else 1e-10
    ctxs = [Context(0) for _ in probs]
    try:
        rs = [P.setup(c, pr, store=store) for c, pr in zip(ctxs, probs)]
        say(f"batch hex20 {precision} {stop} graph={graph}: alone")
        alone = []
        for c, r in zip(ctxs, rs):
            x, res = c.pcg_solve(r, 3000, tol, stop=stop, precision=precision)
            alone.append((res.iters, res.converged, x))
        say(f"batch hex20 {precision} {stop} graph={graph}: batched")
        got = _run_batch(ctxs, rs, 3000, tol, stop, precision, graph)
        for (ia, ca, xa), (ib, cb, xb) in zip(alone, got):
            assert ca and cb and ia == ib
            assert np.array_equal(xa, xb)
        if precision == "fp64":
            pr = probs[1]
            Ko = O.assemble_elastic(pr.g_coord(), pr.e, pr.v, pr.e_elem)
            so = O.STOP_CHECON if stop == "checon" else O.STOP_RELRES
            xo, it, conv = O.pcg(pr.g_g, Ko, pr.r0, O.diag(pr.g_g, Ko, pr.neq), tol, 3000, stop=so,
                                 dot_order=O.DOT_DEVICE)
            assert it == got[1][0] and np.array_equal(xo, got[1][2])
    finally:
        for c in ctxs:
            c.close()


@pytest.mark.parametrize("precision,graph", [("fp64", True), ("mixed", True), ("fp64", "persistent"),
                                             ("mixed", "persistent")])
def test_batch_equals_alone_hex8_laplace_penalty(precision, graph):
    """hex8 Laplace cubes with penalty fixed freedoms (p123's fixed rows in
    the scatter) and, for the mixed path, the FP32 copy of p."""
    probs = [P.laplace_cube(n, lognormal_k=True, seed=s) for n, s in ((5, 2), (3, 7), (7, 9))]
    store = "fp64" if precision == "fp64" else "fp32"
    ctxs = [Context(0) for _ in probs]
    try:
        rs = [P.setup(c, pr, store=store) for c, pr in zip(ctxs, probs)]
        say(f"batch hex8 laplace {precision}: alone")
        alone = [c.pcg_solve(r, 3000, 1e-10, stop="relres", precision=precision) for c, r in zip(ctxs, rs)]
        say(f"batch hex8 laplace {precision} {graph}: batched")
        got = _run_batch(ctxs, rs, 3000, 1e-10, "relres", precision, graph, chunk=5)
        for (xa, ra), (ib, cb, xb) in zip(alone, got):
            assert ra.converged and cb and ra.iters == ib
            assert np.array_equal(xa, xb)
    finally:
        for c in ctxs:
            c.close()


@pytest.mark.parametrize("precision,stop,graph,solo", [("mixed64", "relres", True, 60), ("fp64", "checon", False, 60),
                                                       ("mixed", "relres", True, 1), ("split", "relres", True, 100)])
def test_batch_with_solo_solves_equals_alone(precision, stop, graph, solo):
    """Solves above the solo threshold run their own launches inside the
    batch's runs (all of them for threshold 1), the rest share the batched
    launches: every solve still bit-identical to itself alone."""
    probs = _elastic_mix()
    store = "fp64" if precision == "fp64" else "fp32"
    tol = 1e-7 if stop == "checon" else 1e-10
    ctxs = [Context(0) for _ in probs]
    try:
        rs = [P.setup(c, pr, store=store) for c, pr in zip(ctxs, probs)]
        say(f"batch solo>={solo} {precision} {stop} graph={graph}: alone")
        alone = []
        for c, r in zip(ctxs, rs):
            x, res = c.pcg_solve(r, 3000, tol, stop=stop, precision=precision)
            alone.append((res.iters, res.converged, x))
        say(f"batch solo>={solo} {precision} {stop} graph={graph}: batched")
        got = _run_batch(ctxs, rs, 3000, tol, stop, precision, graph, solo=solo)
        for (ia, ca, xa), (ib, cb, xb) in zip(alone, got):
            assert ca and cb and ia == ib
            assert np.array_equal(xa, xb)
    finally:
        for c in ctxs:
            c.close()


def test_batch_rejects_mixed_settings():
    """Solves of one batch share ntot, precision and stopping rule."""
    from parafem_amd import PFError
    a, b = Context(0), Context(0)
    try:
        ra = P.setup(a, P.elastic_cube(2))
        rb = P.setup(b, P.elastic_cube(3))
        a.pcg_begin(ra, 100, 1e-8, stop="relres")
        b.pcg_begin(rb, 100, 1e-8, stop="checon")
        with Batch(0) as bt:
            with pytest.raises(PFError):
                bt.run([a, b], 2)
            with pytest.raises(PFError):
                bt.run([a, a], 2)
    finally:
        a.close()
        b.close()
