"""The autotuner on the GPU: every candidate setting reports the default
setting's patterns (the sweep checks it and would raise), and the chosen
one is written to a configuration file that reads back."""
import pytest

import gpu_graph_mining_amd as gm
from gpu_graph_mining_amd import tune
from gpu_graph_mining_amd import workloads as wl

from helpers import case_db

pytestmark = pytest.mark.gpu


def test_tune_mine_small(golden, tmp_path):
    case = golden["config1_1k"]
    db = case_db(case)
    res = tune.tune_mine(db, case["minsup"], runs=1, shapes=((64, 7), (256, 8)), canon=(0, 2))
    assert len(res["table"]) == 4 and set(res["best"]) == set(tune.PARAMS)
    assert res["default_seconds"] is not None and res["best_seconds"] <= res["default_seconds"]
    tune.write_config(tmp_path / "t.csv", "config1_1k", res["best"])
    params = tune.read_config(tmp_path / "t.csv", "config1_1k")
    pats, _ = gm.mine(db, case["minsup"], **params)
    assert len(pats) == case["n_patterns"]


def test_tune_batch_small():
    jobs = wl.config3_jobs()[:6]
    b = gm.Batch([j.database() for j in jobs])
    res = tune.tune_batch(b, [j.minsup for j in jobs], runs=1, shapes=((128, 8), (256, 8)), canon=(0,))
    assert res["best"]["chunk_rows"] in (128, 256)
