"""Host logic of the serving loop (no device): the request stream and the
admission estimate."""
import numpy as np

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
