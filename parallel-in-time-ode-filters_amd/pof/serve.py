"""Continuous-batching request loop for IEKS solves on one GPU.

Requests (IVPs of any ``(d, order)``) arrive over time.  Each ``(d, order)``
group holds one or more :class:`pof.batch.Batch` workspaces.  The loop works
as follows:

* It admits waiting requests in arrival order as soon as a batch of their
  group has a free slot and a free workspace range and no iteration in flight.
  When no batch of the group can take a request, a new batch is created for
  it if the device's free memory (``hipMemGetInfo``, through
  ``torch.cuda.mem_get_info``) and the server's budget allow: admission by
  device memory, so a group's share grows with its load instead of being a
  fixed split.  A group's extra batches are freed once they empty.
* It iterates every batch that has IVPs in flight, each at its own pace on its
  own stream: as soon as a batch's iteration has completed (an event query, no
  wait) its finished IVPs are retired, waiting requests are admitted into it
  and its next iteration is enqueued, so a batch with short iterations never
  waits for one with long ones.  Iterates stay in HBM between iterations.
* It retires each IVP the iteration it converges, copies its result to the
  host, and frees its range for the next request.
* A request that no batch of its group can ever hold and for which no new
  batch can be created (budget or free memory, with no extra batch left to
  free) is answered at once with status ``"rejected"`` (no result).

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

    ``budget_bytes`` caps the device memory of the server's batches.  A group's
    batches are sized for ``budget_bytes / n_groups`` (at least 4096 entries, or
    the request's own size); the first batch of a group is always created (at
    the minimum size when the budget is smaller), further ones only while the
    budget and the device's free memory (less ``reserve_bytes``) allow.
    ``slots`` is the number of IVP slots per batch.
    """

    MIN_ENTRIES = 1 << 12

    def __init__(self, budget_bytes=64 << 30, n_groups=12, slots=512, maxiters=100, calibrate=True,
                 keep_means=True, reserve_bytes=2 << 30):
        self.budget = int(budget_bytes)
        self.n_groups = int(n_groups)
        self.slots = int(slots)
        self.maxiters = int(maxiters)
        self.calibrate = calibrate
        self.keep_means = keep_means
        self.reserve = int(reserve_bytes)
        self.groups = {}  # (d, order) -> [Batch, ...]; the first is kept, the others freed when empty
        self._bytes = {}  # id(batch) -> device bytes
        self._entries = {}  # id(batch) -> workspace entries
        self.used = 0
        self.batches_created = 0

    def _target_entries(self, D):
        per = self.budget // max(self.n_groups, 1)
        unit = _batch.bytes_for(D, 1 << 20, 0) / float(1 << 20)
        return max(int(per / unit), self.MIN_ENTRIES)

    def _new_batch(self, key, need):
        """A new batch for group `key` that holds an IVP of `need` entries, or None."""
        import torch

        d, order = key
        D = d * (order + 1)
        first = not self.groups.get(key)
        free = torch.cuda.mem_get_info()[0]
        for entries in (max(self._target_entries(D), need), max(self.MIN_ENTRIES, need)):
            nbytes = _batch.bytes_for(D, entries, self.slots)
            if nbytes + self.reserve > free:
                continue
            if not first and self.used + nbytes > self.budget:
                continue
            if first and entries > self.MIN_ENTRIES and self.used + nbytes > self.budget:
                continue
            b = _batch.Batch(d, order, entries, self.slots)
            self.groups.setdefault(key, []).append(b)
            self._bytes[id(b)] = nbytes
            self._entries[id(b)] = entries
            self.used += nbytes
            self.batches_created += 1
            return b
        return None

    def _admit(self, key, req):
        """Admit `req` into a batch of its group: ("ok", batch, ticket), ("wait",) or ("reject",)."""
        f, y0, ts, q, init = req
        need = _batch.entries_for(len(ts))
        blist = self.groups.get(key, [])
        busy = False
        for b in blist:
            if b.launched:
                busy = True
                continue
            if b.free_entries < need:
                continue
            try:
                return ("ok", b, b.admit(f, y0, ts, init=init, calibrate=self.calibrate, maxiters=self.maxiters))
            except _batch.BatchFull:
                continue
        b = self._new_batch(key, need)
        if b is not None:
            return ("ok", b, b.admit(f, y0, ts, init=init, calibrate=self.calibrate, maxiters=self.maxiters))
        # no batch now: wait while the group can still take it later (a range or slot frees up) or
        # memory can come back (an extra batch of any group empties); otherwise it never fits
        fits_some = any(self._capacity(b) >= need for b in blist)
        extras = any(len(v) > 1 for v in self.groups.values())
        if busy or fits_some or extras:
            return ("wait",)
        return ("reject",)

    def _capacity(self, b):
        """Largest IVP (in entries) the batch could hold empty."""
        return self._entries.get(id(b), 0)

    def _release_empty_extras(self):
        for key, blist in self.groups.items():
            keep = [blist[0]] if blist else []
            for b in blist[1:]:
                if b.active == 0 and not b.launched:
                    self.used -= self._bytes.pop(id(b), 0)
                    self._entries.pop(id(b), None)
                    b.close()
                else:
                    keep.append(b)
            self.groups[key] = keep

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
        inflight = {}  # (id(batch), ticket) -> rid
        ndone = 0
        t0 = clock()
        while nxt < n or waiting or inflight:
            now = clock() - t0
            while nxt < n and recs[order[nxt]].arrival <= now:
                waiting.append(int(order[nxt]))
                nxt += 1
            # admission in arrival order; a request that cannot be placed now waits (its group's
            # later requests wait behind it, other groups go on)
            still = collections.deque()
            blocked = set()
            while waiting:
                rid = waiting.popleft()
                f, y0, ts, q, init = requests[rid]
                key = (f.d, q)
                if key in blocked:
                    still.append(rid)
                    continue
                st = self._admit(key, requests[rid])
                if st[0] == "ok":
                    inflight[(id(st[1]), st[2])] = rid
                    recs[rid].admitted = clock() - t0
                elif st[0] == "wait":
                    blocked.add(key)
                    still.append(rid)
                else:
                    r = recs[rid]
                    r.done = clock() - t0
                    r.status = "rejected"
                    ndone += 1
                    if on_done is not None:
                        on_done(r, ndone)
            waiting = still
            if not inflight:
                if nxt < n:
                    time.sleep(max(0.0, recs[order[nxt]].arrival - (clock() - t0)))
                continue
            for blist in self.groups.values():
                for g in blist:
                    if not g.launched and g.active:
                        g.launch()
            progressed = False
            for blist in self.groups.values():
                for g in blist:
                    if not g.launched or not g.ready():
                        continue
                    progressed = True
                    for t, (ys, info) in g.finish():
                        rid = inflight.pop((id(g), t))
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
            if progressed:
                self._release_empty_extras()
            else:
                time.sleep(20e-6)
        return recs

    def close(self):
        for blist in self.groups.values():
            for g in blist:
                g.close()
        self.groups = {}
        self._bytes = {}
        self._entries = {}
        self.used = 0


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
