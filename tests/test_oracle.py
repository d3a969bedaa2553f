"""Pins of the CPU oracle (oracle/cpu_gspan.cpp) -- CPU only.

The oracle is the checker of every GPU parity test, so it is itself checked
against: the reference's own output (golden.json, written by the reference
build in oracle/_ref), the reference's known-answer tests, and brute force.
"""
import random

import pytest

import gpu_graph_mining_amd as gm
from oracle import py_oracle as po

from helpers import brute_force_frequent, case_db, check_listing, code_graph, listing, monomorphisms

# src/globals/tests/dfs_ismin_tests.cpp:48-61
ISMIN_KAT = [
    ("(0 1 0 0 0);(1 2 -1 0 0);(2 0 -1 0 -1);(2 3 -1 0 0);(3 0 -1 0 -1)", True),
    ("(0 1 0 0 0);(1 2 -1 0 0)", True),
    ("(0 1 0 0 0);(1 2 -1 0 0);(2 0 -1 0 -1);(2 3 -1 0 0);(3 1 -1 0 -1)", False),
    ("(0 1 0 0 0);(1 2 0 0 0);(2 0 0 0 0);(1 3 0 0 0)", False),
    ("(0 1 0 0 0);(1 2 0 0 0);(2 0 0 0 0);(2 3 0 0 0)", True),
    ("(0 1 0 0 0);(1 2 0 0 0);(2 3 0 0 0)", True),
    ("(0 1 0 0 0);(1 2 0 0 0);(1 3 0 0 0)", True),
    ("(0 1 0 0 0);(1 2 0 0 0);(1 3 0 0 0);(1 4 0 0 0)", True),
    ("(0 1 0 0 0);(1 2 0 0 0);(2 3 0 0 0);(2 4 0 0 0);(1 5 0 0 0)", True),
    ("(0 1 0 0 0);(1 2 0 0 0);(2 3 0 0 0);(3 1 0 0 0)", False),
    ("(0 1 0 0 0);(1 2 0 0 0);(2 0 0 0 0);(1 3 0 0 0)", False),
]


def _fast_cases(golden):
    return [c for c in golden.values() if c["name"] not in ("chem_10",)]


def test_oracle_matches_reference_output(golden):
    """Every golden case: the restatement's sorted (code, support) listing
    equals the reference build's."""
    for case in _fast_cases(golden):
        db = case_db(case)
        pats, _ = po.mine_arrays(db.arrays(), case["minsup"])
        check_listing(listing(pats), case)


def test_oracle_chem10_emission_order_matches_reference(golden):
    """Byte-identical output order on Chemical_340 at support 10 (22758
    codes): the md5 of the reference CLI's output file (code-only lines in
    emission order) measured from the reference build."""
    import hashlib
    pats, st = po.mine_file(gm.__path__[0] + "/../tests/golden/Chemical_340.txt", "txt", 10)
    assert st["n_patterns"] == 22758
    blob = "".join(po.code_str(c) + "\n" for c, _, _ in pats).encode()
    assert hashlib.md5(blob).hexdigest() == "9c62661180405ea09080046d42ba51c0"
    check_listing(listing(pats), golden["chem_10"])


@pytest.mark.parametrize("code,expect", ISMIN_KAT)
def test_oracle_is_min_known_answers(code, expect):
    assert po.is_min(gm.DFSCode.read_from_str(code).wildcard()) is expect


def test_oracle_support_from_records_known_answer():
    """src/gspan_cuda/tests/compute_support_remapped_test.cpp:23-71: 14
    records with global vertex ids (graph = vertex // 10) and 5 DFS keys."""
    recs = [((2, 3, 0, 1, 0), 1), ((2, 3, 0, 2, 2), 2), ((2, 3, 0, 1, 2), 1), ((2, 0, 0, 1, 0), 0),
            ((2, 3, 0, 1, 0), 10), ((2, 3, 0, 1, 2), 10), ((2, 0, 0, 1, 0), 10), ((2, 3, 0, 1, 2), 20),
            ((2, 3, 0, 2, 2), 50), ((2, 3, 0, 1, 0), 50), ((2, 3, 0, 2, 1), 52), ((2, 0, 0, 1, 0), 52),
            ((2, 3, 0, 2, 1), 61), ((2, 0, 0, 1, 0), 60)]
    keys = [(2, 3, 0, 1, 0), (2, 3, 0, 2, 2), (2, 3, 0, 1, 2), (2, 3, 0, 2, 1), (2, 0, 0, 1, 0)]
    kidx = [keys.index(k) for k, _ in recs]
    gids = [v // 10 for _, v in recs]
    assert list(po.support_from_records(kidx, gids, len(keys))) == [3, 2, 3, 2, 4]


def _tiny_db(seed, n_graphs=7, nv=6, extra=2, nvl=2, nel=2):
    rng = random.Random(seed)
    graphs = []
    for _ in range(n_graphs):
        labels = [rng.randrange(nvl) for _ in range(nv)]
        edges = {}
        for v in range(1, nv):
            edges[(rng.randrange(v), v)] = rng.randrange(nel)
        for _ in range(extra):
            a, b = rng.sample(range(nv), 2)
            edges.setdefault((min(a, b), max(a, b)), rng.randrange(nel))
        graphs.append((labels, [(a, b, l) for (a, b), l in edges.items()]))
    return graphs


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_complete_and_unique_against_brute_force(seed):
    """On tiny databases the oracle reports every connected frequent
    subgraph with <= 4 edges exactly once (up to isomorphism), with its
    support, and |Proj(P)| equals the number of label-preserving
    monomorphisms of P into the database."""
    from helpers import canonical_form
    graphs = _tiny_db(seed)
    db = gm.GraphDB.from_graphs(graphs)
    minsup = 2
    pats, _ = po.mine_arrays(db.arrays(), minsup)
    small = [(c, s, n) for c, s, n in pats if len(c) <= 4]
    forms = {}
    for c, s, n in small:
        labels, edges = code_graph(c)
        f = canonical_form(labels, edges)
        assert f not in forms, "pattern reported twice"
        forms[f] = s
        gadjs = []
        for gl, ge in graphs:
            adj = {}
            for a, b, l in ge:
                adj[(a, b)] = l
                adj[(b, a)] = l
            gadjs.append((gl, adj))
        assert n == sum(monomorphisms(labels, edges, gl, adj) for gl, adj in gadjs)
    assert forms == brute_force_frequent(graphs, minsup, 4)


def test_oracle_candidate_counts_and_bytes_closed_form():
    """E_cand, E_child and the algorithmic bytes worked out by hand
    (B_ext(P) = sum over rows of 8k + sum over rightmost-path vertices v of
    (8 + 12 deg v), plus 12 E_cand(P); B_sup = 12 E_cand):
    - an edge a-b, labels 0 and 1: one root row (0 <= 1), k = 2, both
      vertices of degree 1 on the rightmost path, no extension:
      B_ext = 16 + 2 * 20 = 56, E_cand = 0;
    - a triangle with equal labels: 6 directed root rows, each with a
      forward extension from v1 and one from v0 (12 candidates; rows
      16 + 2 * 32 = 80 bytes); the path child (6 rows, k = 3, three
      degree-2 vertices: 24 + 3 * 32 = 120 bytes) has one backward
      candidate per row (6); the triangle (6 rows, 120 bytes) has none.
      E_cand = 18, B_ext = 480 + 144 + 720 + 72 + 720 = 2136, B_sup = 216,
      E_child = 6 + 6."""
    db = gm.GraphDB.from_arrays([0, 2], [0, 1], [0, 1], [0], [1], [0])
    _, st = po.mine_arrays(db.arrays(), 1)
    assert (st["e_cand"], st["b_ext"], st["b_sup"], st["e_child"], st["e_root"]) == (0, 56, 0, 0, 1)
    db = gm.GraphDB.from_arrays([0, 3], [0, 0, 0], [0, 3], [0, 1, 0], [1, 2, 2], [0, 0, 0])
    pats, st = po.mine_arrays(db.arrays(), 1)
    assert len(pats) == 3
    assert (st["e_cand"], st["b_ext"], st["b_sup"], st["e_child"], st["e_root"]) == (18, 2136, 216, 12, 6)
