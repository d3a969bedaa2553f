"""Mining under a device memory budget (gm_cfg.mem_budget_bytes): a level
whose children would not fit is written in pieces by parent pattern (or by
key within one pattern), each piece's patterns mined to the end before the
next -- the reference sizes its support batches by a budget the same way
(src/gspan_cuda/gspan_cuda_mult_block.cpp:10-112).  Whatever the budget,
the patterns, supports, |Proj|, E_child and E_cand are the unlimited run's
and the oracle's."""
import pytest

import gpu_graph_mining_amd as gm
from gpu_graph_mining_amd import workloads as wl
from oracle import py_oracle as po

from helpers import case_db, check_listing, listing

pytestmark = pytest.mark.gpu


def _key(pats):
    return sorted((tuple(c), s, n) for c, s, n in pats)


@pytest.mark.parametrize("name", ["chem_34", "config1_1k", "config4_2k", "labeled_3"])
def test_budget_pieces_give_the_same_results(golden, name):
    case = golden[name]
    db = case_db(case)
    full, s0 = gm.mine(db, case["minsup"], collect_proj=True)
    opats, ost = po.mine_arrays(db.arrays(), case["minsup"])
    peak = s0["peak_device_bytes"]
    for budget in (peak // 4, peak // 16, 1):
        print("case", name, "budget", budget, flush=True)
        part, sb = gm.mine(db, case["minsup"], collect_proj=True, mem_budget=budget)
        check_listing(listing(part), case)
        assert _key(part) == _key(full)
        for k in ("n_patterns", "e_child", "e_root", "e_cand"):
            assert sb[k] == s0[k], (budget, k)
        assert sb["e_child"] == ost["e_child"] and sb["e_cand"] == ost["e_cand"]
    # the smallest budget writes one child per piece: many more level steps
    assert sb["n_levels"] > s0["n_levels"]


def test_budget_with_the_host_check_and_speculation(golden):
    case = golden["chem_34"]
    db = case_db(case)
    full, _ = gm.mine(db, case["minsup"], collect_proj=True)
    for canon_host, speculate in ((1, 2), (2, 1), (1, 1)):
        part, _ = gm.mine(db, case["minsup"], collect_proj=True, mem_budget=20000, canon_host=canon_host,
                          speculate=speculate)
        assert _key(part) == _key(full), (canon_host, speculate)


def test_budget_in_a_batch():
    """Root pieces are whole jobs: each job's results equal its own run's."""
    jobs = wl.config3_jobs()[:12]
    dbs = [j.database() for j in jobs]
    mins = [j.minsup for j in jobs]
    batch = gm.Batch(dbs)
    full, _ = batch.mine(mins)
    part, _ = batch.mine(mins, mem_budget=200000)
    for a, b in zip(full, part):
        assert sorted((tuple(c), s) for c, s, _ in a) == sorted((tuple(c), s) for c, s, _ in b)


def test_negative_budget_is_rejected():
    db, _ = wl.config0(n_graphs=20)
    with pytest.raises(gm.GMError):
        gm.mine(db, 5, mem_budget=-1)
