"""Serving on the GPU: every request answered by the continuously batched
server (queries and mining jobs overlapping on the device) equals the same
request run alone -- exactly (integer supports and pattern sets)."""
import pytest

import gpu_graph_mining_amd as gm
from gpu_graph_mining_amd import serving

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone():
    db = gm.GraphDB.molecule(4, 1500)
    out = serving.serve(db, rate=2000.0, n_requests=48, seed=9, budget_bytes=1 << 30)
    reqs = out["_requests"]
    assert len(reqs) == 48 and all(r.done >= r.arrival for r in reqs)
    kinds = {r.kind for r in reqs}
    assert serving.QUERY in kinds and serving.MINING in kinds
    assert max(r.batch_size for r in reqs if r.kind == serving.QUERY) > 1  # batching happened
    for r in reqs:
        print("request", r.rid, r.kind, flush=True)
        assert r.result == serving.run_alone(db, r)
    assert out["query"]["p99_ms"] >= out["query"]["p50_ms"] > 0


def test_subset_batch_equals_each_subset_alone():
    """Jobs gathered on the device from the resident database (overlapping
    subsets, one empty) give each subset's own results."""
    db = gm.GraphDB.molecule(4, 3000)
    db.upload()
    subsets = [list(range(0, 3000, 3)), list(range(1, 2000, 2)), [5, 7, 11, 13, 2999], [], list(range(2500, 3000))]
    mins = [40, 60, 2, 1, 25]
    b = gm.Batch.from_subsets(db, subsets)
    try:
        res, _ = b.mine(mins, collect_proj=True)
    finally:
        b.close()
    for ids, m, got in zip(subsets, mins, res):
        if not ids:
            assert got == []
            continue
        sub = db.subset(ids)
        want, _ = gm.mine(sub, m, collect_proj=True)
        sub.close()
        assert sorted((tuple(c), s, n) for c, s, n in got) == sorted((tuple(c), s, n) for c, s, n in want)


def test_served_results_equal_alone_on_the_full_database():
    """The 50,000-graph molecule database: queries and mining jobs admitted
    into concurrent slots under a small memory budget (requests queue for
    memory), each result equal to the request run alone."""
    from gpu_graph_mining_amd import workloads as wl
    db = wl.config4()
    out = serving.serve(db, rate=1000.0, n_requests=40, seed=21, budget_bytes=256 << 20)
    reqs = out["_requests"]
    assert {r.kind for r in reqs} == {serving.QUERY, serving.MINING}
    for r in reqs:
        print("request", r.rid, r.kind, flush=True)
        assert r.result == serving.run_alone(db, r)
    assert out["bytes_per_unit"][serving.QUERY] > 0 and out["bytes_per_unit"][serving.MINING] > 0
