"""Continuous-batching request loop for IEKS solves on one GPU.

Requests (IVPs of any ``(d, order)``) arrive over time.  Each ``(d, order)``
group has one :class:`pof.batch.Batch` with a fixed device workspace.  The
loop works as follows:

* It admits waiting requests in arrival order as soon as their group has a
  free slot and a free workspace range (admission by device memory).
* It iterates every group that has IVPs in flight, each at its own pace on
  its own stream: as soon as a group's iteration has completed (an event
  query, no wait) its finished IVPs are retired, waiting requests are
  admitted into it and its next iteration is enqueued, so a group with short
  iterations never waits for one with long ones. Iterates stay in HBM
  between iterations.
* It retires each IVP the iteration it converges, copies its result to the
  host, and frees its range for the next request.
* A request larger than its group's whole workspace is answered at once with
  status ``"rejected"`` (no result).

An IVP's arithmetic does not depend on what else is in flight, so a result
served in the mix equals the same solve alone.
"""
import collections
import time

import numpy as np

from . import batch as _batch


class Record:
    __slots__ = ("rid", "arrival", "admitted", "done", "iterations", "status", "means")

    def __init__(self, rid, arrival):
        self.rid = rid
        self.arrival = arrival
        self.admitted = None
        self.done = None
        self.iterations = None
        self.status = None
        self.means = None


class Server:
    """One GPU's request loop.

    ``budget_bytes`` is the device memory shared by the groups. Each group
    created on first use gets ``budget_bytes / n_groups`` (``n_groups`` = the
    expected number of groups), sized into workspace entries for its state
    dimension, and ``slots`` IVP slots.
    """

    def __init__(self, budget_bytes=64 << 30, n_groups=12, slots=512, maxiters=100, calibrate=True,
                 keep_means=True):
        self.budget = int(budget_bytes)
        self.n_groups = int(n_groups)
        self.slots = int(slots)
        self.maxiters = int(maxiters)
        self.calibrate = calibrate
        self.keep_means = keep_means
        self.groups = {}

    def _group(self, d, order):
        key = (d, order)
        b = self.groups.get(key)
        if b is None:
            D = d * (order + 1)
            per = self.budget // self.n_groups
            unit = _batch.bytes_for(D, 1 << 20, 0) / float(1 << 20)
            entries = max(int(per / unit), 1 << 12)
            b = _batch.Batch(d, order, entries, self.slots)
            self.groups[key] = b
        return b

    def run(self, requests, arrivals, clock=time.perf_counter, on_done=None):
        """Serve ``requests[i]`` (``(f, y0, ts, order, init)``) arriving at ``arrivals[i]`` seconds.

        Returns one :class:`Record` per request, with times in seconds from the start.
        ``on_done(record, n_done)`` is called as each request completes.
        """
        n = len(requests)
        recs = [Record(i, float(arrivals[i])) for i in range(n)]
        order = np.argsort(arrivals, kind="stable")
        nxt = 0
        waiting = collections.deque()
        inflight = {}  # (group key, ticket) -> rid
        ndone = 0
        t0 = clock()
        while nxt < n or waiting or inflight:
            now = clock() - t0
            while nxt < n and recs[order[nxt]].arrival <= now:
                waiting.append(int(order[nxt]))
                nxt += 1
            # admission in arrival order; a request whose group is full or has an iteration in flight
            # waits (admission happens between a group's iterations)
            still = collections.deque()
            blocked = set()
            while waiting:
                rid = waiting.popleft()
                f, y0, ts, q, init = requests[rid]
                key = (f.d, q)
                if key in blocked:
                    still.append(rid)
                    continue
                g = self._group(f.d, q)
                if g.launched:
                    blocked.add(key)
                    still.append(rid)
                    continue
                try:
                    t = g.admit(f, y0, ts, init=init, calibrate=self.calibrate, maxiters=self.maxiters)
                except _batch.BatchFull:
                    if g.active == 0:
                        # larger than the group's whole workspace: it can never be admitted
                        r = recs[rid]
                        r.done = clock() - t0
                        r.status = "rejected"
                        ndone += 1
                        if on_done is not None:
                            on_done(r, ndone)
                        continue
                    blocked.add(key)
                    still.append(rid)
                    continue
                inflight[(key, t)] = rid
                recs[rid].admitted = clock() - t0
            waiting = still
            if not inflight:
                if nxt < n:
                    time.sleep(max(0.0, recs[order[nxt]].arrival - (clock() - t0)))
                continue
            for g in self.groups.values():
                if not g.launched and g.active:
                    g.launch()
            progressed = False
            for k, g in self.groups.items():
                if not g.launched or not g.ready():
                    continue
                progressed = True
                for t, (ys, info) in g.finish():
                    rid = inflight.pop((k, t))
                    r = recs[rid]
                    m = ys.mean.cpu().numpy()  # the result on the host ends the request
                    r.done = clock() - t0
                    r.iterations = info["iterations"]
                    r.status = info["status"]
                    if self.keep_means:
                        r.means = m
                    ndone += 1
                    if on_done is not None:
                        on_done(r, ndone)
            if not progressed:
                time.sleep(20e-6)
        return recs

    def close(self):
        for g in self.groups.values():
            g.close()
        self.groups = {}


def latency_stats(recs):
    lat = np.array([r.done - r.arrival for r in recs if r.done is not None and r.status != "rejected"])
    if lat.size == 0:
        return {"p50_ms": None, "p99_ms": None, "mean_ms": None}
    return {
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3),
    }


def fifo_plain(requests, arrivals, solve_one, clock=time.perf_counter, on_done=None):
    """The plain baseline: one request at a time, in arrival order, each solved alone.

    ``solve_one(request)`` runs one solve and returns its host means.  Returns
    the records (arrival, done) of a single-server FIFO queue run in real time.
    ``on_done(record, n_done)`` is called as each request completes.
    """
    n = len(requests)
    recs = [Record(i, float(arrivals[i])) for i in range(n)]
    order = np.argsort(arrivals, kind="stable")
    t0 = clock()
    for k, rid in enumerate(order):
        r = recs[int(rid)]
        now = clock() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        r.admitted = clock() - t0
        r.means = solve_one(requests[int(rid)])
        r.done = clock() - t0
        if on_done is not None:
            on_done(r, k + 1)
    return recs
