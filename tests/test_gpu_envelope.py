"""Edges of the envelope on the GPU: a level with more frequent keys than
the first frequent list holds (the stage runs again with an exact list,
from cleared counters), patterns past 32 vertices (64-byte tuples, the wide
pattern metadata) and the pattern-size limit, which fails cleanly."""
import random

import pytest

import gpu_graph_mining_amd as gm
from oracle import py_oracle as po

pytestmark = pytest.mark.gpu


def test_frequent_list_overflow_gives_the_closed_form():
    """300 random labelled trees of 256 vertices over 250 vertex and 250
    edge labels: about 76,000 distinct labelled edges, each frequent at
    support 1, more than the 65,536 records of the root level's first list.
    The 1-edge patterns and supports are the closed form: every labelled
    edge (fromlabel <= tolabel) with the number of graphs that hold it."""
    rng = random.Random(7)
    vb, vl, eb, ef, et, el = [0], [], [0], [], [], []
    want = {}
    for g in range(300):
        labs = [rng.randrange(250) for _ in range(256)]
        vl += labs
        vb.append(len(vl))
        seen = set()
        for v in range(1, 256):
            u = rng.randrange(v)
            e = rng.randrange(250)
            ef.append(u)
            et.append(v)
            el.append(e)
            a, b = sorted((labs[u], labs[v]))
            seen.add((a, e, b))
        eb.append(len(ef))
        for t in seen:
            want[t] = want.get(t, 0) + 1
    assert len(want) > 65536
    db = gm.GraphDB.from_arrays(vb, vl, eb, ef, et, el)
    pats, st = gm.mine(db, 1, max_edges=1)
    got = {(c[0][2], c[0][3], c[0][4]): s for c, s, _ in pats}
    assert len(pats) == len(got) == len(want)
    assert got == want


def _paths_db(graphs):
    """Graphs given as (n, chords): a path 0-1-...-(n-1) with vertex labels
    0..n-1 (path edges label 0) plus the chords (label 1)."""
    vb, vl, eb, ef, et, el = [0], [], [0], [], [], []
    for n, chords in graphs:
        vl += list(range(n))
        vb.append(len(vl))
        for a, b in [(i, i + 1) for i in range(n - 1)] + list(chords):
            ef.append(a)
            et.append(b)
            el.append(0 if b == a + 1 else 1)
        eb.append(len(ef))
    return gm.GraphDB.from_arrays(vb, vl, eb, ef, et, el)


def _by_code(patterns):
    return {po.normalize(c): (s, n) for c, s, n in patterns}


def test_patterns_of_33_to_40_vertices_closed_form():
    """Two identical 40-vertex paths with distinct labels at support 2: the
    frequent patterns are the sub-paths, 780 of them, of up to 40 vertices
    (the levels past 32 vertices run with 64-byte tuples and the wide
    pattern metadata)."""
    n = 40
    db = _paths_db([(n, []), (n, [])])
    pats, st = gm.mine(db, 2)
    assert len(pats) == sum(n - k for k in range(1, n)) == 780
    assert max(len(c) for c, _, _ in pats) == n - 1
    assert all(s == 2 for _, s, _ in pats)


@pytest.mark.parametrize("canon_host", [0, 1, 2])
def test_patterns_past_32_vertices_match_the_oracle(canon_host):
    """Chorded 44-vertex paths (backward edges to pattern vertices 33 and 38,
    rightmost-path sources past vertex 32) and a plain path: every pattern,
    support and |Proj(P)| equal the oracle's, with the canonical check on
    the host, on the device (patterns past its 39 vertices go to the host)
    or adaptive."""
    ch = [(2, 5), (33, 36), (38, 41)]
    db = _paths_db([(44, ch), (44, ch), (44, [])])
    print("chorded paths, canon_host", canon_host, flush=True)
    pats, st = gm.mine(db, 2, collect_proj=True, canon_host=canon_host)
    opats, ost = po.mine_arrays(db.arrays(), 2)
    assert len(opats) == 14190
    assert max(max(max(t[0], t[1]) for t in c) + 1 for c, _, _ in opats) == 44
    assert _by_code(pats) == _by_code(opats)
    assert st["e_child"] == ost["e_child"] and st["e_cand"] == ost["e_cand"]
    if canon_host == 2:
        assert st["canon_host_keys"] > 0  # the 40-44-vertex candidates


def test_pattern_beyond_64_vertices_fails_cleanly():
    """Two identical 70-vertex paths with distinct labels: the frequent
    65-vertex path passes the limit; mining stops with GM_E_UNSUPPORTED
    and the handle stays usable."""
    n = 70
    db = _paths_db([(n, []), (n, [])])
    with pytest.raises(gm.GMError) as e:
        gm.mine(db, 2)
    assert e.value.code == gm.GM_E_UNSUPPORTED and "64 vertices" in str(e.value)
    pats, _ = gm.mine(db, 2, max_edges=63)  # patterns up to 64 vertices
    assert len(pats) == sum(n - k for k in range(1, 64))  # every sub-path of each length
