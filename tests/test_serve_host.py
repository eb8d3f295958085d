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


class _FakeBatch:
    """Stand-in for pof.batch.Batch with the same admission interface: `entries` of workspace,
    `slots` IVPs; an admitted IVP finishes after one launch / finish round."""

    made = []

    def __init__(self, d, order, entries, slots):
        self.d, self.order, self.entries, self.slots = d, order, entries, slots
        self.used, self.live, self.launched, self.closed = 0, {}, False, False
        self.next = 1
        _FakeBatch.made.append(self)

    @property
    def free_entries(self):
        return self.entries - self.used

    @property
    def active(self):
        return len(self.live)

    def admit(self, f, y0, ts, init="prior", calibrate=True, maxiters=100):
        need = 2 * len(ts)
        if need > self.free_entries or len(self.live) >= self.slots:
            raise serve._batch.BatchFull("full")
        self.used += need
        t = self.next
        self.next += 1
        self.live[t] = (need, len(ts))
        return t

    def launch(self):
        self.launched = True

    def ready(self):
        return True

    def finish(self):
        import torch

        self.launched = False
        out = []
        for t, (need, n) in list(self.live.items()):
            self.used -= need
            del self.live[t]
            out.append((t, (serve._batch.MVNSqrt(torch.zeros(n, 1), None), {"iterations": 1.0, "status": "converged"})))
        return out

    def close(self):
        self.closed = True


def _fake_server(monkeypatch, free_bytes, **kw):
    import torch

    _FakeBatch.made = []
    monkeypatch.setattr(serve._batch, "Batch", _FakeBatch)
    monkeypatch.setattr(serve._batch, "bytes_for", lambda D, entries, slots: 1000 * entries)
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda *a: (free_bytes, 1 << 40))
    return serve.Server(**kw)


class _F:
    d = 1


def test_admission_grows_a_group_by_free_memory(monkeypatch):
    """Requests of one group beyond its first batch's workspace get further batches while the budget
    and the free memory allow; emptied extra batches are freed; all requests are served."""
    srv = _fake_server(monkeypatch, free_bytes=10 ** 12, budget_bytes=40_000_000, n_groups=4, slots=64,
                       reserve_bytes=0)
    # target batch size: budget / n_groups = 10 MB -> 10,000 entries of 1000 bytes
    reqs = [(_F(), np.zeros(1), np.zeros(3000), 2, "constant") for _ in range(6)]  # 6000 entries each
    recs = srv.run(reqs, np.zeros(6))
    assert all(r.status == "converged" for r in recs)
    assert srv.batches_created >= 2          # one 10,000-entry batch holds one request
    assert len(srv.groups[(1, 2)]) == 1      # the extra batches were freed once empty
    assert sum(b.closed for b in _FakeBatch.made) == srv.batches_created - 1
    assert srv.used == 10_000_000


def test_admission_rejects_what_can_never_fit(monkeypatch):
    """A request larger than any batch the budget allows, with no extra batch to free, is rejected;
    the others are served."""
    srv = _fake_server(monkeypatch, free_bytes=10 ** 12, budget_bytes=1, n_groups=1, slots=8, reserve_bytes=0)
    reqs = [(_F(), np.zeros(1), np.zeros(n), 2, "constant") for n in (100, 5000, 200)]
    recs = srv.run(reqs, np.zeros(3))
    assert [r.status for r in recs] == ["converged", "rejected", "converged"]


def test_admission_respects_free_memory(monkeypatch):
    """No batch is created beyond the device's free memory less the reserve: with room for none, the
    request is rejected at once."""
    srv = _fake_server(monkeypatch, free_bytes=5_000_000, budget_bytes=1 << 40, n_groups=1, slots=8,
                       reserve_bytes=2_000_000)
    recs = srv.run([(_F(), np.zeros(1), np.zeros(3000), 2, "constant")], np.zeros(1))
    assert recs[0].status == "rejected" and srv.batches_created == 0
