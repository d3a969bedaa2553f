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
