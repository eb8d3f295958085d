"""Host logic of the request loop (pof/serve.py) that needs no GPU: the FIFO
plain baseline's queueing and completion hook, and the latency statistics."""
import numpy as np

from pof import serve


def test_fifo_plain_queues_in_arrival_order():
    arrivals = np.array([0.004, 0.0, 0.002, 0.002])
    served = []

    def solve_one(req):
        served.append(req)
        return np.full(2, float(req))

    seen = []
    recs = serve.fifo_plain(list(range(4)), arrivals, solve_one, on_done=lambda r, n: seen.append((r.rid, n)))
    assert served == [1, 2, 3, 0]  # stable order by arrival
    assert seen == [(1, 1), (2, 2), (3, 3), (0, 4)]
    for r in recs:
        assert r.done >= r.admitted >= r.arrival - 1e-9
        assert np.array_equal(r.means, np.full(2, float(r.rid)))


def test_latency_stats():
    recs = []
    for i, lat in enumerate([0.010, 0.020, 0.030, 0.040]):
        r = serve.Record(i, 1.0)
        r.done = 1.0 + lat
        recs.append(r)
    st = serve.latency_stats(recs)
    assert abs(st["p50_ms"] - 25.0) < 1e-9
    assert abs(st["mean_ms"] - 25.0) < 1e-9
    assert abs(st["p99_ms"] - np.percentile([10, 20, 30, 40], 99)) < 1e-9
    assert serve.latency_stats([serve.Record(0, 0.0)])["p50_ms"] is None


def test_latency_stats_skip_rejected():
    a = serve.Record(0, 0.0)
    a.done = 0.010
    b = serve.Record(1, 0.0)
    b.done = 5.0
    b.status = "rejected"
    assert abs(serve.latency_stats([a, b])["p50_ms"] - 10.0) < 1e-9
