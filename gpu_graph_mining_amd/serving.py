"""Serving: a request loop over a resident graph database with continuous
batching.

Requests (the serving workload of the molecule-like database):
  * pattern-support queries -- a connected 3-12-edge subgraph sampled by a
    random walk from a database graph; answer: the number of database graphs
    that contain it (gm_support_queries);
  * mining jobs -- mine a random subset of 1,000-10,000 graphs at a min
    support of 5-20 %; answer: its frequent patterns and supports.

The server keeps the database resident on the GPU and runs two workers, each
with its own device handle (and so its own HIP stream): one answers support
queries, one runs mining jobs, so the two kinds of work overlap on the
device.  Continuous batching: requests that arrive while a batch runs wait
in the queue and are admitted into the next batch as soon as the worker is
free, as many as the admission budget allows.  Admission is by estimated
embedding memory:
  query  est = |D| * edges * 16 bytes   (16-byte rows, one per embedding
                                          per graph per edge, as a bound)
  mining est = subset graphs * 64 KiB
A batch takes queued requests in arrival order while the sum of their
estimates stays within `budget_bytes` (the first request is always taken).
Latency of a request = completion time - arrival time.
"""
from __future__ import annotations

import math
import queue
import random
import threading
import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

from . import GraphDB, mine, support, Batch

QUERY = "query"
MINING = "mining"


@dataclass
class Request:
    kind: str
    rid: int
    arrival: float = 0.0
    code: Optional[List[Tuple[int, int, int, int, int]]] = None  # query
    graphs: Optional[List[int]] = None                            # mining job
    minsup: int = 0
    est_bytes: int = 0
    done: float = 0.0
    result: object = None
    batch_size: int = 0


def random_walk_code(db_arrays: dict, rng: random.Random, n_edges: int, tries: int = 50):
    """A connected subgraph of n_edges edges, grown by a random walk in a
    random database graph (a random frontier edge of the visited part is
    added at each step), written as a DFS-style tuple list in the graph's
    labels: (from, to, fromlabel, elabel, tolabel) over walk-local vertex
    ids.  Returns None when no graph of the database has n_edges edges in
    one component after `tries` attempts."""
    vb, eb = db_arrays["graph_vbegin"], db_arrays["graph_ebegin"]
    vl, ef, et, el = db_arrays["vlabels"], db_arrays["efrom"], db_arrays["eto"], db_arrays["elabel"]
    n_graphs = len(vb) - 1
    for _ in range(tries):
        g = rng.randrange(n_graphs)
        e0, e1 = int(eb[g]), int(eb[g + 1])
        if e1 - e0 < n_edges:
            continue
        adj = {}
        for e in range(e0, e1):
            a, b = int(ef[e]), int(et[e])
            adj.setdefault(a, []).append((b, int(el[e]), e))
            adj.setdefault(b, []).append((a, int(el[e]), e))
        start = int(ef[rng.randrange(e0, e1)])
        local = {start: 0}
        used = set()
        tuples = []
        while len(tuples) < n_edges:
            frontier = [(a, b, lab, e) for a in local for (b, lab, e) in adj[a] if e not in used]
            if not frontier:
                break
            a, b, lab, e = frontier[rng.randrange(len(frontier))]
            used.add(e)
            if b not in local:
                local[b] = len(local)
            base = int(vb[g])
            tuples.append((local[a], local[b], int(vl[base + a]), lab, int(vl[base + b])))
        if len(tuples) == n_edges:
            return tuples
    return None


def make_requests(db: GraphDB, n: int, seed: int = 4, query_frac: float = 0.9,
                  min_edges: int = 3, max_edges: int = 12, min_jobs_graphs: int = 1000,
                  max_jobs_graphs: int = 10000, min_sup: float = 0.05, max_sup: float = 0.20) -> List[Request]:
    """A seeded request mix (no arrival times yet)."""
    rng = random.Random(seed)
    arrays = db.arrays()
    n_graphs = len(db)
    out = []
    for i in range(n):
        if rng.random() < query_frac:
            code = None
            while code is None:
                code = random_walk_code(arrays, rng, rng.randint(min_edges, max_edges))
            out.append(Request(QUERY, i, code=code, est_bytes=n_graphs * len(code) * 16))
        else:
            k = rng.randint(min(min_jobs_graphs, n_graphs), min(max_jobs_graphs, n_graphs))
            graphs = sorted(rng.sample(range(n_graphs), k))
            frac = rng.uniform(min_sup, max_sup)
            ms = max(1, int(math.ceil(frac * k - 1e-9)))
            out.append(Request(MINING, i, graphs=graphs, minsup=ms, est_bytes=k * 65536))
    return out


def run_alone(db: GraphDB, r: Request):
    """The result of one request on its own (the reference for the mix)."""
    if r.kind == QUERY:
        return support(db, [r.code])[0]
    sub = db.subset(r.graphs)
    try:
        pats, _ = mine(sub, r.minsup)
        return sorted((tuple(c), s) for c, s, _ in pats)
    finally:
        sub.close()


class Server:
    """Two workers (queries, mining jobs) over their own resident handles of
    one database; submit() enqueues, results land on the Request objects."""

    def __init__(self, db: GraphDB, budget_bytes: int = 8 << 30, max_query_batch: int = 4096,
                 max_mining_batch: int = 8):
        self.db_query = db
        self.db_query.upload()
        self.budget = budget_bytes
        self.max_q = max_query_batch
        self.max_m = max_mining_batch
        self.qq: "queue.Queue[Request]" = queue.Queue()
        self.mq: "queue.Queue[Request]" = queue.Queue()
        self.finished: List[Request] = []
        self._lock = threading.Lock()
        self._stop = threading.Event()
        self._errors: List[BaseException] = []
        self._workers = [threading.Thread(target=self._query_loop, daemon=True),
                         threading.Thread(target=self._mining_loop, daemon=True)]
        for w in self._workers:
            w.start()

    def submit(self, r: Request) -> None:
        (self.qq if r.kind == QUERY else self.mq).put(r)

    def _take(self, q: "queue.Queue[Request]", limit: int) -> List[Request]:
        """Block for one request, then admit queued ones while the memory
        estimate fits the budget."""
        while not self._stop.is_set():
            try:
                first = q.get(timeout=0.05)
                break
            except queue.Empty:
                continue
        else:
            return []
        batch, used = [first], first.est_bytes
        while len(batch) < limit:
            try:
                r = q.get_nowait()
            except queue.Empty:
                break
            if used + r.est_bytes > self.budget:
                q.put(r)  # admitted with the next batch
                break
            batch.append(r)
            used += r.est_bytes
        return batch

    def _finish(self, batch: List[Request]) -> None:
        now = time.perf_counter()
        with self._lock:
            for r in batch:
                r.done = now
                r.batch_size = len(batch)
                self.finished.append(r)

    def _query_loop(self) -> None:
        try:
            while not self._stop.is_set():
                batch = self._take(self.qq, self.max_q)
                if not batch:
                    continue
                sups = support(self.db_query, [r.code for r in batch])
                for r, s in zip(batch, sups):
                    r.result = s
                self._finish(batch)
        except BaseException as e:  # reported by close()
            self._errors.append(e)

    def _mining_loop(self) -> None:
        try:
            while not self._stop.is_set():
                batch = self._take(self.mq, self.max_m)
                if not batch:
                    continue
                subs = [self.db_query.subset(r.graphs) for r in batch]
                try:
                    b = Batch(subs)  # the jobs of one batch run in one schedule
                    try:
                        res, _ = b.mine([r.minsup for r in batch])
                    finally:
                        b.close()
                finally:
                    for s in subs:
                        s.close()
                for r, pats in zip(batch, res):
                    r.result = sorted((tuple(c), s) for c, s, _ in pats)
                self._finish(batch)
        except BaseException as e:
            self._errors.append(e)

    def close(self) -> None:
        self._stop.set()
        for w in self._workers:
            w.join()
        if self._errors:
            raise self._errors[0]


def percentile(xs: Sequence[float], p: float) -> float:
    if not xs:
        return float("nan")
    s = sorted(xs)
    k = max(0, min(len(s) - 1, int(math.ceil(p / 100.0 * len(s))) - 1))
    return s[k]


def serve(db: GraphDB, rate: float, n_requests: int, seed: int = 4, **server_kw) -> dict:
    """Poisson arrivals at `rate` requests/s of a seeded mix; returns latency
    percentiles (ms) per kind and overall, throughput, and the requests."""
    reqs = make_requests(db, n_requests, seed=seed)
    rng = random.Random(seed + 1)
    srv = Server(db, **server_kw)
    t0 = time.perf_counter()
    t = t0
    for r in reqs:
        t += rng.expovariate(rate)
        delay = t - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        r.arrival = time.perf_counter()
        srv.submit(r)
    while True:
        with srv._lock:
            if len(srv.finished) == len(reqs):
                break
        if srv._errors:
            break
        time.sleep(0.002)
    t_end = time.perf_counter()
    srv.close()
    out = {"rate_per_s": rate, "requests": len(reqs), "seconds": t_end - t0,
           "throughput_per_s": len(reqs) / (t_end - t0)}
    for kind in (QUERY, MINING, None):
        lat = [(r.done - r.arrival) * 1e3 for r in reqs if kind is None or r.kind == kind]
        key = kind or "all"
        out[key] = {"n": len(lat), "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
                    "mean_batch": (sum(r.batch_size for r in reqs if kind is None or r.kind == kind) /
                                   max(1, len(lat)))}
    out["_requests"] = reqs
    return out
