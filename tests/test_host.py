"""Host-side tests of the product library (no GPU needed): the C ABI loads
and exports every declared symbol, readers, generators, the canonical check
and the error paths that fire before any device work."""
import random
import re
import struct
from pathlib import Path

import numpy as np
import pytest

import gpu_graph_mining_amd as gm
from gpu_graph_mining_amd import workloads as wl
from oracle import py_oracle as po

from helpers import GOLDEN

ROOT = Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "gspan_amd.h").read_text()
    declared = set(re.findall(r"\b(gm_[a-z_0-9]+)\s*\(", header))
    declared -= {"gm_pattern_cb", "gm_job_pattern_cb"}
    lib = gm.load_library()
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared == set(gm.EXPORTED_SYMBOLS)
    assert gm.version().startswith("gspan_amd")


def test_txt_reader_chemical_340_crlf():
    db = gm.GraphDB.load(GOLDEN / "Chemical_340.txt", "txt")
    c = db.counts()
    assert (c["n_graphs"], c["n_vertices"], c["n_edges"], c["max_vertices"]) == (340, 9189, 9317, 214)
    assert (c["n_vlabels"], c["n_elabels"]) == (66, 4)
    # the '-' prefixed filetype is accepted (src/globals/dbio.cpp:33-41)
    assert gm.GraphDB.load(GOLDEN / "Chemical_340.txt", "-txt").counts() == c


def test_txt_roundtrip(tmp_path):
    db, _ = wl.config0(n_graphs=50)
    p = tmp_path / "db.txt"
    db.write_txt(p)
    back = gm.GraphDB.load(p, "txt")
    a, b = db.arrays(), back.arrays()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_fsg_reader_maps_labels_per_graph(tmp_path):
    p = tmp_path / "db.fsg"
    p.write_text("t # 0\nv 0 C\nv 1 O\nu 0 1 s\nt # 1\nv 0 O\nv 1 C\nu 0 1 d\n")
    a = gm.GraphDB.load(p, "fsg").arrays()
    # labels are numbered in first-seen order within each graph
    # (src/globals/graph_types.cpp:122-157)
    assert list(a["vlabels"]) == [0, 1, 0, 1]
    assert list(a["elabel"]) == [0, 0]


def test_bin_reader_reference_fixtures():
    for f in ("labeled.bin", "unlabeled.bin"):
        db = gm.GraphDB.load(GOLDEN / f, "bin")
        assert db.counts()["n_graphs"] > 0


def _write_cudabin(path, arrays):
    """graph_database_cuda layout (src/globals/dbio.cpp:346-404)."""
    vb, vl, eb = arrays["graph_vbegin"], arrays["vlabels"], arrays["graph_ebegin"]
    ef, et, el = arrays["efrom"], arrays["eto"], arrays["elabel"]
    n = len(vb) - 1
    mvc = int(max(vb[i + 1] - vb[i] for i in range(n)))
    edges, elabs, voffs, vlabs = [], [], [], []
    for g in range(n):
        nv = vb[g + 1] - vb[g]
        adj = [[] for _ in range(nv)]
        for e in range(eb[g], eb[g + 1]):
            adj[ef[e]].append((et[e], el[e]))
            adj[et[e]].append((ef[e], el[e]))
        for v in range(mvc):
            if v < nv:
                voffs.append(len(edges))
                vlabs.append(int(vl[vb[g] + v]))
                for w, l in adj[v]:
                    edges.append(g * mvc + w)
                    elabs.append(int(l))
            else:
                voffs.append(-1)
                vlabs.append(-1)
    voffs.append(len(edges))
    with open(path, "wb") as f:
        f.write(struct.pack("<4i", n, len(edges), mvc, int(vb[-1])))
        for arr in (edges, elabs, vlabs, voffs):
            f.write(np.asarray(arr, dtype="<i4").tobytes())


def test_cudabin_reader(tmp_path):
    db, _ = wl.config0(n_graphs=30)
    a = db.arrays()
    p = tmp_path / "db.cudabin"
    _write_cudabin(p, a)
    b = gm.GraphDB.load(p, "cudabin").arrays()
    assert np.array_equal(a["graph_vbegin"], b["graph_vbegin"])
    assert np.array_equal(a["vlabels"], b["vlabels"])
    for g in range(len(a["graph_vbegin"]) - 1):
        sa = {(min(x, y), max(x, y), l) for x, y, l in zip(*(a[k][a["graph_ebegin"][g]:a["graph_ebegin"][g + 1]]
                                                             for k in ("efrom", "eto", "elabel")))}
        sb = {(min(x, y), max(x, y), l) for x, y, l in zip(*(b[k][b["graph_ebegin"][g]:b["graph_ebegin"][g + 1]]
                                                             for k in ("efrom", "eto", "elabel")))}
        assert sa == sb


def test_generators_are_deterministic():
    a = gm.GraphDB.planted(7, 200).sha256()
    assert a == gm.GraphDB.planted(7, 200).sha256()
    assert a != gm.GraphDB.planted(8, 200).sha256()
    m = gm.GraphDB.molecule(3, 100)
    c = m.counts()
    assert c["n_graphs"] == 100 and c["n_vlabels"] <= 8 and c["n_elabels"] <= 3
    arr = m.arrays()
    deg = np.zeros(c["n_vertices"], np.int64)
    for g in range(100):
        for e in range(arr["graph_ebegin"][g], arr["graph_ebegin"][g + 1]):
            deg[arr["graph_vbegin"][g] + arr["efrom"][e]] += 1
            deg[arr["graph_vbegin"][g] + arr["eto"][e]] += 1
    assert deg.max() <= 4


def test_config0_generator_shape():
    db, ms = wl.config0()
    c = db.counts()
    assert c["n_graphs"] == 1000 and ms == 100
    a = db.arrays()
    nv = np.diff(a["graph_vbegin"])
    ne = np.diff(a["graph_ebegin"])
    assert nv.min() >= 10 and nv.max() <= 30
    assert np.array_equal(ne, np.round(1.2 * nv).astype(np.int64))


def test_config3_jobs_shape():
    jobs = wl.config3_jobs()
    assert len(jobs) == 256
    assert all(200 <= j.n_graphs <= 20000 for j in jobs)
    assert all(0.02 <= j.min_support_frac <= 0.20 for j in jobs)
    assert jobs == wl.config3_jobs()


def _rand_code(rng, n_range=(2, 7), n_labels=3, extra=(0, 4)):
    n = rng.randint(*n_range)
    labels = [rng.randint(0, n_labels - 1) for _ in range(n)]
    edges = set()
    for v in range(1, n):
        edges.add((rng.randrange(v), v))
    for _ in range(rng.randint(*extra)):
        a, b = rng.sample(range(n), 2)
        edges.add((min(a, b), max(a, b)))
    el = {e: rng.randint(0, 1) for e in edges}
    adj = {v: [] for v in range(n)}
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    order, code, used = {}, [], set()
    start = rng.randrange(n)
    order[start] = 0

    def visit(u):
        for w in sorted([w for w in adj[u] if w in order and (min(u, w), max(u, w)) not in used],
                        key=lambda w: order[w]):
            used.add((min(u, w), max(u, w)))
            code.append((order[u], order[w], labels[u], el[(min(u, w), max(u, w))], labels[w]))
        nb = adj[u][:]
        rng.shuffle(nb)
        for w in nb:
            if w not in order:
                order[w] = len(order)
                used.add((min(u, w), max(u, w)))
                code.append((order[u], order[w], labels[u], el[(min(u, w), max(u, w))], labels[w]))
                visit(w)

    visit(start)
    return code


def test_canonical_check_matches_oracle_on_random_codes():
    """The product's minimum-DFS-code check (csrc/canonical.cpp) against the
    oracle's restatement of ismin.cpp on random DFS traversals of random
    labelled graphs (both minimal and non-minimal codes)."""
    rng = random.Random(11)
    n_min = 0
    for _ in range(4000):
        c = _rand_code(rng)
        got = gm.DFSCode(c).is_min()
        want = po.is_min(po.normalize(c))
        assert got == want, c
        n_min += want
    assert 300 < n_min < 3700


def _chorded_path_code(rng, n, n_chords, reverse=False):
    """A path 0-1-...-(n-1) with vertex labels 0..n-1 and chords (i, j),
    j > i + 1, all edge labels 0, walked along the path, each chord a
    backward edge at its later end (earlier ends in DFS order).  From vertex
    0 this is the minimum DFS code (the greedy growth takes the smallest new
    label first); from vertex n-1 (reverse) it is not."""
    chords = set()
    while len(chords) < n_chords:
        i = rng.randrange(n - 2)
        chords.add((i, rng.randrange(i + 2, n)))
    seq = list(range(n))[::-1] if reverse else list(range(n))
    pos = {v: k for k, v in enumerate(seq)}
    code = []
    for k in range(1, n):
        u, v = seq[k - 1], seq[k]
        code.append((k - 1, k, u, 0, v))
        back = sorted(pos[x] for (a, b) in chords for x, y in ((a, b), (b, a)) if y == v and pos[x] < k)
        code += [(k, t, v, 0, seq[t]) for t in back]
    return code


def test_canonical_check_on_large_patterns():
    """Patterns of 33-64 vertices and up to 128 edges (two-word edge masks
    in the check's embeddings) against the oracle: minimal codes of chorded
    paths, the same codes with two tuples' edge labels swapped, and random
    traversals of random graphs (rarely minimal)."""
    import sys
    rng = random.Random(5)
    old = sys.getrecursionlimit()
    sys.setrecursionlimit(10000)
    n_min = 0
    try:
        for i in range(60):
            n = rng.randint(33, 64)
            c = _chorded_path_code(rng, n, rng.randint(0, 128 - (n - 1)))
            assert gm.DFSCode(c).is_min() and po.is_min(po.normalize(c)), c
            n_min += 1
            d = _chorded_path_code(rng, n, rng.randint(0, 128 - (n - 1)), reverse=True)
            assert not gm.DFSCode(d).is_min() and not po.is_min(po.normalize(d)), d
        for i in range(60):
            c = _rand_code(rng, n_range=(33, 64), n_labels=(2 if i % 2 else 60), extra=(10, 60))
            assert gm.DFSCode(c).is_min() == po.is_min(po.normalize(c)), c
    finally:
        sys.setrecursionlimit(old)
    # past the limits: a clean error
    path = [(i, i + 1, 0, 0, 0) for i in range(64)]  # 65 vertices
    with pytest.raises(gm.GMError) as e:
        gm.DFSCode(path).is_min()
    assert e.value.code == gm.GM_E_UNSUPPORTED


def test_canonical_check_known_answers():
    from test_oracle import ISMIN_KAT
    for s, expect in ISMIN_KAT:
        assert gm.DFSCode.read_from_str(s).is_min() is expect, s


def _err(fn):
    with pytest.raises(gm.GMError) as e:
        fn()
    return e.value.code


def test_error_paths_before_device_work(tmp_path):
    ok = gm.GraphDB.from_graphs([([0, 1], [(0, 1, 0)])])
    assert _err(lambda: gm.mine(ok, 0)) == gm.GM_E_ARG
    loop = gm.GraphDB.from_graphs([([0, 1], [(0, 1, 0), (1, 1, 0)])])
    assert _err(lambda: gm.mine(loop, 1)) == gm.GM_E_UNSUPPORTED
    par = gm.GraphDB.from_graphs([([0, 1], [(0, 1, 0), (1, 0, 1)])])
    assert _err(lambda: gm.mine(par, 1)) == gm.GM_E_UNSUPPORTED
    big = gm.GraphDB.from_graphs([(list(range(2)) * 150, [(i, i + 1, 0) for i in range(299)])])
    assert _err(lambda: gm.mine(big, 1)) == gm.GM_E_UNSUPPORTED
    many = gm.GraphDB.from_graphs([(list(range(300)), [(i, i + 1, 0) for i in range(299)])])
    assert _err(lambda: gm.mine(many, 1)) == gm.GM_E_UNSUPPORTED
    assert _err(lambda: gm.GraphDB.load(tmp_path / "nope.txt", "txt")) == gm.GM_E_IO
    bad = tmp_path / "bad.txt"
    bad.write_text("t # 0\nv 0 1\ne 0 5 1\n")
    assert _err(lambda: gm.GraphDB.load(bad, "txt")) == gm.GM_E_ARG
    assert _err(lambda: gm.GraphDB.load(bad, "xml")) == gm.GM_E_ARG
    assert _err(lambda: gm.GraphDB.from_arrays([0, 2], [0, 1], [0, 1], [0], [7], [0])) == gm.GM_E_ARG
    assert _err(lambda: gm.DFSCode([(0, 0, 1, 1, 1), (0, 1, 1, 1, 1)]).is_min()) == gm.GM_E_ARG


def test_support_queries_reject_malformed_codes_before_device_work():
    """Disconnected patterns, missing labels and empty queries are argument
    errors raised on the host (no GPU needed)."""
    db = gm.GraphDB.from_graphs([([0, 1, 2], [(0, 1, 0), (1, 2, 0)])])
    for codes in ([[(0, 1, 0, 0, 1), (2, 3, 1, 0, 2)]],   # two components
                  [[(0, 1, 0, 0, -1)]],                   # forward tuple without its to-label
                  [[]]):
        with pytest.raises(gm.GMError) as e:
            gm.support(db, codes)
        assert e.value.code == gm.GM_E_ARG


def test_serving_request_mix_is_seeded_and_connected():
    from gpu_graph_mining_amd import serving
    db = gm.GraphDB.molecule(4, 300)
    a = serving.make_requests(db, 40, seed=5)
    b = serving.make_requests(db, 40, seed=5)
    assert [(r.kind, r.code, r.graphs, r.minsup) for r in a] == [(r.kind, r.code, r.graphs, r.minsup) for r in b]
    for r in a:
        if r.kind == serving.QUERY:
            assert 3 <= len(r.code) <= 12
            seen = {r.code[0][0]}
            for f, t, *_ in r.code:  # every tuple touches the part grown so far
                assert f in seen or t in seen
                seen |= {f, t}
        else:
            assert r.minsup >= 1 and len(r.graphs) == len(set(r.graphs))


def test_callback_guard_keeps_the_exception():
    def boom(*args):
        raise ValueError("x")

    g = gm._Guard(boom)
    assert g(1, 2) == 1  # stop mining
    with pytest.raises(ValueError):
        g.raise_pending()
    assert gm._Guard(lambda *a: 0)(1) == 0
