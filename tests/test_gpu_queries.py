"""Pattern-support queries on the GPU (gm_support_queries) against mined
supports and the oracle: a query's support must equal the support the
reference's gSpan reports for the same pattern, whatever DFS code the query
is written in, alone or batched with others."""
import random

import pytest

import gpu_graph_mining_amd as gm
from oracle import py_oracle as po

from helpers import case_db

pytestmark = pytest.mark.gpu


def _scramble(code, rng):
    """The same pattern under a random vertex numbering and tuple order (a
    DFS code in general no longer minimal, nor even in DFS order)."""
    nv = 1 + max(max(t[0], t[1]) for t in code)
    perm = list(range(nv))
    rng.shuffle(perm)
    label = {}
    for f, t, fl, el, tl in code:
        label[f], label[t] = fl, tl
    out = []
    for f, t, fl, el, tl in code:
        a, b = perm[f], perm[t]
        if rng.random() < 0.5:
            a, b = b, a
        out.append((a, b, label[f] if a == perm[f] else label[t], el, label[t] if b == perm[t] else label[f]))
    rng.shuffle(out)
    return out


@pytest.mark.parametrize("name", ["chem_34", "config1_1k", "config4_2k"])
def test_queries_equal_mined_supports(golden, name):
    case = golden[name]
    db = case_db(case)
    pats, _ = gm.mine(db, case["minsup"])
    codes = [list(c) for c, _, _ in pats]
    want = [s for _, s, _ in pats]
    assert gm.support(db, codes) == want
    rng = random.Random(7)
    assert gm.support(db, [_scramble(c, rng) for c in codes]) == want


def test_queries_equal_oracle_at_support_one():
    """Every pattern of up to 4 edges of a small random database (oracle at
    min support 1), queried in scrambled form, batched and one at a time."""
    db = gm.GraphDB.random(11, 60, 6, 14, 1.3, 3, 2)
    opats, _ = po.mine_arrays(db.arrays(), 1, max_edges=4)
    rng = random.Random(3)
    codes = []
    for c, s, _ in opats:
        full = []
        lab = {}
        for f, t, fl, el, tl in c:
            if fl != -1:
                lab[f] = fl
            if tl != -1:
                lab[t] = tl
        for f, t, fl, el, tl in c:
            full.append((f, t, lab[f], el, lab[t]))
        codes.append(_scramble(full, rng))
    want = [s for _, s, _ in opats]
    got = gm.support(db, codes)
    assert got == want
    for i in range(0, len(codes), max(1, len(codes) // 12)):  # alone equals in the mix
        assert gm.support(db, [codes[i]]) == [want[i]]


def test_queries_absent_patterns_and_labels():
    tri = ([0, 1, 2], [(0, 1, 0), (1, 2, 0), (2, 0, 1)])
    db = gm.GraphDB.from_graphs([tri, tri, ([0, 1], [(0, 1, 0)])])
    got = gm.support(db, [
        [(0, 1, 0, 0, 1)],                     # in all three graphs
        [(0, 1, 0, 0, 1), (1, 2, 1, 0, 2)],    # in the two triangles
        [(0, 1, 0, 0, 1), (1, 2, 1, 0, 2), (2, 0, 2, 1, 0)],
        [(0, 1, 0, 5, 1)],                     # edge label absent from the database
        [(0, 1, 7, 0, 1)],                     # vertex label absent
        [(0, 1, 0, 0, 1), (1, 2, 1, 0, 0)],    # a path that does not occur
    ])
    assert got == [3, 2, 2, 0, 0, 0]
