"""The allocation-free entry point pf_apply_noalloc against pf_apply (same
operator, bit for bit) through the C ABI: two calls in a row on one GPU, the
second on a smaller problem (the fixed scratch keeps the first call's
contents), and the size limit refused with PF_ERR_TOO_LARGE."""
import sys

import numpy as np
import pytest

from parafem_amd import Context, PFError, problems as P
from parafem_amd.api import PF_ERR_TOO_LARGE

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _problem(ctx, pr, store):
    P.setup(ctx, pr, store=store)
    return np.random.default_rng(pr.neq).normal(size=pr.neq)


@pytest.mark.parametrize("store", ["fp64", "fp32"])
def test_apply_noalloc_equals_apply(store):
    with Context(0) as ctx:
        big = P.elastic_cube(6, lognormal_e=True)
        say(f"apply_noalloc {store} n=6 neq={big.neq}")
        x = _problem(ctx, big, store)
        y_ref = ctx.apply(x)
        y1 = ctx.apply_noalloc(x)
        assert np.array_equal(y1, y_ref)
        # a smaller problem next: stale entries past its end must not matter
        small = P.laplace_cube(4)
        say(f"apply_noalloc {store} laplace n=4 neq={small.neq}")
        x2 = _problem(ctx, small, store)
        y_ref2 = ctx.apply(x2)
        y2 = ctx.apply_noalloc(x2)
        assert np.array_equal(y2, y_ref2)
        assert np.any(y2 != 0.0)


def test_apply_noalloc_refuses_beyond_scratch():
    """50^3 hex20 needs 8*60*125000 + 2*8*(1545300+1) bytes (81 MiB) of
    scratch: refused on the host before any launch, while pf_apply still
    serves it."""
    with Context(0) as ctx:
        pr = P.elastic_cube(50)
        P.setup(ctx, pr, assemble=False)
        km = np.eye(pr.ntot)
        ctx.storkm_shared(km)
        say(f"apply_noalloc size refusal nels={pr.nels} neq={pr.neq}")
        x = np.ones(pr.neq)
        with pytest.raises(PFError) as ei:
            ctx.apply_noalloc(x)
        assert ei.value.status == PF_ERR_TOO_LARGE
        y = ctx.apply(x)                 # the allocating entry point still serves it
        # identity element matrices: y(q) counts the element slots of q
        counts = np.bincount(pr.g_g[pr.g_g > 0] - 1, minlength=pr.neq)
        assert np.array_equal(y, counts.astype(float))
