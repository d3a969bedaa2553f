"""Serving: a request loop over a resident graph database with continuous
batching.

Requests (the serving workload of the molecule-like database):
  * pattern-support queries -- a connected 3-12-edge subgraph sampled by a
    random walk from a database graph; answer: the number of database graphs
    that contain it (gm_support_queries);
  * mining jobs -- mine a random subset of 1,000-10,000 graphs at a min
    support of 5-20 %; answer: its frequent patterns and supports.

The server keeps the database resident on the GPU and runs several slots
per request kind, each a worker thread with its own device handle and HIP
stream, so that work of different slots overlaps on the device.  A query
slot answers a batch of queries in one batched projection over its own
resident copy; a mining slot gathers its jobs' subsets from the resident
copy on the device (gm_batch_create_subsets: no host packing, no upload of
graph data) and mines them in one schedule.

Continuous batching: a request that arrives while others run is admitted by
the next free slot of its kind as soon as device memory allows -- with the
queued requests behind it that still fit, up to the slot's batch limit.
Admission is by estimated embedding memory: a request's estimate is its
size in units (query: database graphs x code edges; mining job: edges of
its subset) times the device bytes per unit measured on the requests of its
kind completed so far (peak device bytes of the run less the resident
database), starting from a fixed guess; a batch is admitted while the
estimates of everything running fit the budget (a request that would not
fit on an idle device runs alone).  Latency = completion - arrival.
"""
from __future__ import annotations

import math
import queue
import random
import threading
import time
from dataclasses import dataclass
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
    units: int = 0        # query: database graphs x edges; mining: edges of the subset
    n_vertices: int = 0   # mining: the subset's size
    n_edges: int = 0
    est_bytes: int = 0    # the estimate it was admitted with
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
            out.append(Request(QUERY, i, code=code, units=n_graphs * len(code)))
        else:
            k = rng.randint(min(min_jobs_graphs, n_graphs), min(max_jobs_graphs, n_graphs))
            graphs = sorted(rng.sample(range(n_graphs), k))
            frac = rng.uniform(min_sup, max_sup)
            ms = max(1, int(math.ceil(frac * k - 1e-9)))
            vb, eb = arrays["graph_vbegin"], arrays["graph_ebegin"]
            nv = int(sum(int(vb[g + 1]) - int(vb[g]) for g in graphs))
            ne = int(sum(int(eb[g + 1]) - int(eb[g]) for g in graphs))
            out.append(Request(MINING, i, graphs=graphs, minsup=ms, units=max(ne, 1), n_vertices=nv, n_edges=ne))
    return out


def db_bytes(n_graphs: int, n_vertices: int, n_edges: int) -> int:
    """Device bytes of a packed database (graph and vertex offsets, packed
    adjacency, labels, job and graph ids)."""
    return 4 * (n_graphs + 1) + 4 * (n_vertices + 1) + 8 * n_edges + n_vertices + 4 * n_graphs + 4 * n_vertices


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


class MemoryGate:
    """Device memory admitted to running batches, against a budget."""

    def __init__(self, budget: int):
        self.budget = budget
        self.used = 0
        self.running = 0
        self.cv = threading.Condition()

    def admit_first(self, n: int, stop: threading.Event) -> bool:
        """Wait until n fits beside what runs (or nothing runs)."""
        with self.cv:
            while self.running > 0 and self.used + n > self.budget:
                if stop.is_set():
                    return False
                self.cv.wait(0.05)
            self.used += n
            self.running += 1
            return True

    def admit_more(self, n: int) -> bool:
        """Add n to an admitted batch if it fits now."""
        with self.cv:
            if self.used + n > self.budget:
                return False
            self.used += n
            return True

    def release(self, n: int) -> None:
        with self.cv:
            self.used -= n
            self.running -= 1
            self.cv.notify_all()


class Estimator:
    """Device bytes per unit of a request kind, learned from completed
    batches (peak device bytes less the resident database)."""

    def __init__(self, guess: float):
        self.per_unit = guess
        self.lock = threading.Lock()

    def estimate(self, r: Request) -> int:
        with self.lock:
            return int(r.units * self.per_unit) + 1

    def observe(self, units: int, measured_bytes: int) -> None:
        if units <= 0 or measured_bytes <= 0:
            return
        m = measured_bytes / units
        with self.lock:
            # follow the measurements, and never fall far below the latest
            self.per_unit = max(0.7 * self.per_unit + 0.3 * m, 0.5 * m)


class Server:
    """query_slots + mining_slots workers over resident handles of one
    database; submit() enqueues, results land on the Request objects."""

    def __init__(self, db: GraphDB, budget_bytes: int = 8 << 30, max_query_batch: int = 4096,
                 max_mining_batch: int = 8, query_slots: int = 2, mining_slots: int = 2):
        db.upload()
        c = db.counts()
        self.n_graphs = c["n_graphs"]
        self.resident_bytes = db_bytes(c["n_graphs"], c["n_vertices"], c["n_edges"])
        # one resident copy per query slot (a handle runs one call at a
        # time); mining slots gather their subsets from the first
        self.copies = [db] + [db.subset(list(range(self.n_graphs))) for _ in range(max(0, query_slots - 1))]
        for d in self.copies:
            d.upload()
        self.gate = MemoryGate(budget_bytes)
        self.est = {QUERY: Estimator(16.0), MINING: Estimator(4096.0)}
        self.max_q = max_query_batch
        self.max_m = max_mining_batch
        self.qq: "queue.Queue[Request]" = queue.Queue()
        self.mq: "queue.Queue[Request]" = queue.Queue()
        self.finished: List[Request] = []
        self._lock = threading.Lock()
        self._stop = threading.Event()
        self._errors: List[BaseException] = []
        self.query_slots, self.mining_slots = max(1, query_slots), max(1, mining_slots)
        self._workers = [threading.Thread(target=self._query_loop, args=(self.copies[i],), daemon=True)
                         for i in range(self.query_slots)]
        self._workers += [threading.Thread(target=self._mining_loop, daemon=True) for _ in range(self.mining_slots)]
        for w in self._workers:
            w.start()

    def submit(self, r: Request) -> None:
        (self.qq if r.kind == QUERY else self.mq).put(r)

    def _take(self, q: "queue.Queue[Request]", kind: str, limit: int) -> Tuple[List[Request], int]:
        """Block for one request and the memory it needs, then admit the
        queued ones behind it that fit now."""
        while not self._stop.is_set():
            try:
                first = q.get(timeout=0.05)
                break
            except queue.Empty:
                continue
        else:
            return [], 0
        first.est_bytes = self.est[kind].estimate(first)
        if not self.gate.admit_first(first.est_bytes, self._stop):
            return [], 0
        batch, used = [first], first.est_bytes
        while len(batch) < limit:
            try:
                r = q.get_nowait()
            except queue.Empty:
                break
            r.est_bytes = self.est[kind].estimate(r)
            if not self.gate.admit_more(r.est_bytes):
                # first in line for the next free slot (queue order is
                # arrival order within a kind; it goes back to the front)
                with q.mutex:
                    q.queue.appendleft(r)
                    q.unfinished_tasks += 1
                    q.not_empty.notify()
                break
            batch.append(r)
            used += r.est_bytes
        return batch, used

    def _finish(self, batch: List[Request]) -> None:
        now = time.perf_counter()
        with self._lock:
            for r in batch:
                r.done = now
                r.batch_size = len(batch)
                self.finished.append(r)

    def _query_loop(self, db: GraphDB) -> None:
        try:
            while not self._stop.is_set():
                batch, used = self._take(self.qq, QUERY, self.max_q)
                if not batch:
                    continue
                try:
                    sups, st = support(db, [r.code for r in batch], with_stats=True)
                finally:
                    self.gate.release(used)
                self.est[QUERY].observe(sum(r.units for r in batch), st["peak_device_bytes"] - self.resident_bytes)
                for r, s in zip(batch, sups):
                    r.result = s
                self._finish(batch)
        except BaseException as e:  # reported by close()
            self._errors.append(e)

    def _mining_loop(self) -> None:
        try:
            while not self._stop.is_set():
                batch, used = self._take(self.mq, MINING, self.max_m)
                if not batch:
                    continue
                try:
                    b = Batch.from_subsets(self.copies[0], [r.graphs for r in batch])
                    try:
                        res, st = b.mine([r.minsup for r in batch])
                    finally:
                        b.close()
                finally:
                    self.gate.release(used)
                sub_bytes = sum(db_bytes(len(r.graphs), r.n_vertices, r.n_edges) for r in batch)
                self.est[MINING].observe(sum(r.units for r in batch), st["peak_device_bytes"] - sub_bytes)
                for r, pats in zip(batch, res):
                    r.result = sorted((tuple(c), s) for c, s, _ in pats)
                self._finish(batch)
        except BaseException as e:
            self._errors.append(e)

    def close(self) -> None:
        self._stop.set()
        for w in self._workers:
            w.join()
        for d in self.copies[1:]:
            d.close()
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
           "throughput_per_s": len(reqs) / (t_end - t0),
           "slots": {"query": srv.query_slots, "mining": srv.mining_slots},
           "bytes_per_unit": {k: e.per_unit for k, e in srv.est.items()}}
    for kind in (QUERY, MINING, None):
        lat = [(r.done - r.arrival) * 1e3 for r in reqs if kind is None or r.kind == kind]
        key = kind or "all"
        out[key] = {"n": len(lat), "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
                    "mean_batch": (sum(r.batch_size for r in reqs if kind is None or r.kind == kind) /
                                   max(1, len(lat)))}
    out["_requests"] = reqs
    return out
