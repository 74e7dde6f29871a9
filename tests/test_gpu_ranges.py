"""Range-fused products and sums of the mixed solver (PF_SCATTER_RANGES).

The element products are the slot-list path's FP32 values; only the order of
the FP64 sums into u and of p.u differs.  Bars: one iteration from the same
start gives x within 1e-12 (relative, max norm) of the slot-list path; solves
to relres 1e-10 are deterministic (bit-identical when repeated), converge,
have a true residual <= 1e-10 and lie within 1e-9 of the FP64 solution (the
oracle's arithmetic bit for bit); fixed rows (p123's penalty) included."""
import sys

import numpy as np
import pytest

from parafem_amd import Context, PFError, mesh as M, problems as P

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _hex8_elastic(n):
    m = M.box_hex8(n, n, n, 1.0, 1.0, 1.0)
    rest = M.base_fixed_rest(m)
    nf, neq = M.steering(m.nn, 3, rest)
    nodes, vals = M.top_pressure_loads(m, -1.0)
    r0 = M.nodal_loads_to_rhs(nf, nodes, vals, neq)
    return P.Problem("elastic", m, M.g_g_from(nf, m.g_num), neq, r0, nf,
                     e_elem=M.lognormal(3, m.nels, np.log(100.0), 0.5))


CASES = {
    "hex20": lambda: P.elastic_cube(6, lognormal_e=True),
    "hex20_ragged": lambda: P.elastic_cube(5, lognormal_e=True, nels_xyz=(5, 3, 7)),
    "hex8_elastic": lambda: _hex8_elastic(7),
    "hex8_laplace": lambda: P.laplace_cube(9, lognormal_k=True),
}


@pytest.mark.parametrize("case", list(CASES))
def test_ranges_mixed_solve(case):
    pr = CASES[case]()
    say(f"ranges {case} nels={pr.nels} neq={pr.neq}")
    with Context(0) as ctx:
        r = P.setup(ctx, pr, store="fp64")
        x64, r64 = ctx.pcg_solve(r, 5000, 1e-10, stop="relres")
        r = P.setup(ctx, pr, store="fp32")
        # one iteration from the same start, both orders of the sums
        ctx.pcg_begin(r, 10, 0.0, stop="relres", precision="mixed")
        ctx.pcg_run(1)
        xs = ctx.pcg_x()
        ctx.scatter_mode = "ranges"
        ctx.pcg_begin(r, 10, 0.0, stop="relres", precision="mixed")
        ctx.pcg_run(1)
        xr = ctx.pcg_x()
        assert np.abs(xr - xs).max() <= 1e-12 * np.abs(xs).max()
        # whole solves: deterministic, converged, true residual, vs FP64
        xa, ra = ctx.pcg_solve(r, 5000, 1e-10, stop="relres", precision="mixed")
        xb, rb = ctx.pcg_solve(r, 5000, 1e-10, stop="relres", precision="mixed")
        assert ra.converged and ra.iters == rb.iters and np.array_equal(xa, xb)
        f = P.start_r(ctx, pr)
        y = ctx.apply(xa)
        assert np.linalg.norm(y - f) <= 1e-10 * np.linalg.norm(f)
        err = np.abs(xa - x64).max() / np.abs(x64).max()
        say(f"ranges {case}: {ra.iters} iterations (FP64 {r64.iters}), vs FP64 {err:.2e}")
        assert err <= 1e-9
        ctx.scatter_mode = "slots"
        xc, rc = ctx.pcg_solve(r, 5000, 1e-10, stop="relres", precision="mixed")
        assert rc.converged and abs(rc.iters - ra.iters) <= max(3, ra.iters // 20)


def test_ranges_mode_rules():
    """The mode applies to the mixed solver; the reference dot order and
    the batch refuse it."""
    from parafem_amd import Batch
    pr = P.elastic_cube(3)
    with Context(0) as ctx:
        r = P.setup(ctx, pr, store="fp32")
        ctx.scatter_mode = "ranges"
        ctx.dot_order = "seq"
        with pytest.raises(PFError):
            ctx.pcg_begin(r, 10, 1e-8, stop="relres", precision="mixed")
        ctx.dot_order = "tree"
        ctx.pcg_begin(r, 10, 1e-8, stop="relres", precision="mixed")
        with Batch(0) as b:
            with pytest.raises(PFError):
                b.run([ctx], 2)
        # other precisions keep the slot-list sums (same bits as without the mode)
        xs, rs = ctx.pcg_solve(r, 500, 1e-9, stop="relres", precision="split")
        ctx.scatter_mode = "slots"
        xt, rt = ctx.pcg_solve(r, 500, 1e-9, stop="relres", precision="split")
        assert rs.iters == rt.iters and np.array_equal(xs, xt)


def test_ranges_refuse_an_equation_twice_in_one_element():
    """The device build of the ranges checks that no element lists an
    equation twice (its rows would be summed in one round); the solve is
    refused in range mode and runs with the slot-list sums."""
    pr = P.elastic_cube(3)
    g = pr.g_g.copy()
    # the last element: slot j's equation listed again in slot i, where j's
    # own equation is listed by another element too (every equation keeps a slot)
    e = pr.nels - 1
    uses = np.bincount(g.ravel(), minlength=pr.neq + 1)
    free = np.flatnonzero(g[e] > 0)
    i = int(free[-1])
    j = int(next(k for k in free if k != i and uses[g[e, k]] > 1))
    g[e, j] = g[e, i]
    bad = P.Problem(pr.kind, pr.mesh, g, pr.neq, pr.r0, pr.nf, e=pr.e, v=pr.v)
    with Context(0) as ctx:
        r = P.setup(ctx, bad, store="fp32")
        ctx.scatter_mode = "ranges"
        with pytest.raises(PFError):
            ctx.pcg_begin(r, 10, 1e-8, stop="relres", precision="mixed")
        ctx.scatter_mode = "slots"
        ctx.pcg_begin(r, 10, 1e-8, stop="relres", precision="mixed")
        ctx.pcg_run(2)
