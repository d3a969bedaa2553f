"""Edges of the envelope on the GPU: a level with more frequent keys than
the first frequent list holds (the stage runs again with an exact list,
from cleared counters), and the pattern-size limit, which fails cleanly."""
import random

import pytest

import gpu_graph_mining_amd as gm

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


def test_pattern_beyond_32_vertices_fails_cleanly():
    """Two identical 40-vertex paths with distinct labels: the frequent
    34-vertex path passes the limit; mining stops with GM_E_UNSUPPORTED
    and the handle stays usable."""
    n = 40
    vb, vl, eb, ef, et, el = [0], [], [0], [], [], []
    for _ in range(2):
        vl += list(range(n))
        vb.append(len(vl))
        ef += list(range(n - 1))
        et += list(range(1, n))
        el += [0] * (n - 1)
        eb.append(len(ef))
    db = gm.GraphDB.from_arrays(vb, vl, eb, ef, et, el)
    with pytest.raises(gm.GMError) as e:
        gm.mine(db, 2)
    assert e.value.code == gm.GM_E_UNSUPPORTED and "32 vertices" in str(e.value)
    pats, _ = gm.mine(db, 2, max_edges=31)  # patterns up to 32 vertices
    assert len(pats) == sum(n - k for k in range(1, 32))  # every sub-path of each length
