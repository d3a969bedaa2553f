"""On-disk database formats (host side, no GPU): the BIN writer and the
partitioned BIN read (src/globals/dbio.cpp:209-329), the cudabin writer
(dbio.cpp:346-370), round trips as in src/globals/dbio_test.cpp:53-131, and
the same handling of self-loops in every reader."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import gpu_graph_mining_amd as gm
from gpu_graph_mining_amd import workloads as wl
from oracle import py_oracle as po

from helpers import GOLDEN, check_listing
from test_host import _write_cudabin


def _edge_sets(a):
    out = []
    vb, eb = a["graph_vbegin"], a["graph_ebegin"]
    for g in range(len(vb) - 1):
        s = sorted((min(x, y), max(x, y), l) for x, y, l in
                   zip(*(a[k][eb[g]:eb[g + 1]] for k in ("efrom", "eto", "elabel"))))
        out.append(s)
    return out


def _same_db(a, b):
    assert np.array_equal(a["graph_vbegin"], b["graph_vbegin"])
    assert np.array_equal(a["vlabels"], b["vlabels"])
    assert _edge_sets(a) == _edge_sets(b)


def test_bin_roundtrip_and_idempotence(tmp_path):
    db, _ = wl.config0(n_graphs=60)
    p1, p2 = tmp_path / "a.bin", tmp_path / "b.bin"
    db.write(p1, "bin")
    back = gm.GraphDB.load(p1, "bin")
    _same_db(db.arrays(), back.arrays())
    # reading fixes the adjacency order; writing what was read reproduces
    # the file byte for byte
    back.write(p2, "bin")
    again = gm.GraphDB.load(p2, "bin")
    for k, v in back.arrays().items():
        assert np.array_equal(v, again.arrays()[k]), k
    again.write(tmp_path / "c.bin", "-bin")
    assert p2.read_bytes() == (tmp_path / "c.bin").read_bytes()


def test_bin_writer_reproduces_reference_fixtures(tmp_path):
    """The reference's own BIN files, read and written back: same graphs,
    and the record layout (sizes, offsets, edge ids) of the reference's
    writer -- the re-written file equals the fixture byte for byte once its
    adjacency order is the one the reader produces."""
    for f in ("labeled.bin", "unlabeled.bin"):
        db = gm.GraphDB.load(GOLDEN / f, "bin")
        out = tmp_path / f
        db.write(out, "bin")
        back = gm.GraphDB.load(out, "bin")
        for k, v in db.arrays().items():
            assert np.array_equal(v, back.arrays()[k]), (f, k)
        ref = (GOLDEN / f).read_bytes()
        mine = out.read_bytes()
        assert len(mine) == len(ref), f  # same record sizes, hence same offsets
        assert mine[:8 * (1 + db.counts()["n_graphs"])] == ref[:8 * (1 + db.counts()["n_graphs"])], f


@pytest.mark.skipif(not po.REF_CLI.exists(), reason="reference build (oracle/_ref) not present")
def test_bin_writer_read_by_reference_build(tmp_path, golden):
    """A BIN file written here, mined by the reference's own reader and
    gSpan (built from its sources), gives the reference's listing of the
    same database read as txt."""
    case = golden["chem_68"]
    db = gm.GraphDB.load(GOLDEN / "Chemical_340.txt", "txt")
    p = tmp_path / "chem.bin"
    db.write(p, "bin")
    out = tmp_path / "out.tsv"
    subprocess.run([str(po.REF_CLI), "bin", str(p), str(case["minsup"]), str(out)], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    check_listing(sorted(out.read_text().splitlines()), case)


def test_bin_partitioned_read(tmp_path):
    db, _ = wl.config0(n_graphs=47)
    p = tmp_path / "db.bin"
    db.write(p, "bin")
    full = gm.GraphDB.load(p, "bin").arrays()
    for parts in (1, 3, 5):
        sizes, seen = [], []
        for part in range(parts):
            a = gm.GraphDB.load_part(p, parts, part).arrays()
            sizes.append(len(a["graph_vbegin"]) - 1)
            seen.append(a)
        # n / parts graphs per part, the last one takes the rest
        assert sizes[:-1] == [47 // parts] * (parts - 1) and sum(sizes) == 47
        vl = np.concatenate([a["vlabels"] for a in seen])
        assert np.array_equal(vl, full["vlabels"])
        assert sum((_edge_sets(a) for a in seen), []) == _edge_sets(full)
    with pytest.raises(gm.GMError):
        gm.GraphDB.load_part(p, 3, 3)


def test_cudabin_writer_matches_padded_layout(tmp_path):
    """The cudabin writer equals an independent restatement of
    graph_database_cuda::create_from_host_representation
    (src/globals/cuda_graph_types.cpp:65-122) byte for byte."""
    db, _ = wl.config0(n_graphs=40)
    p, q = tmp_path / "mine.cudabin", tmp_path / "restated.cudabin"
    db.write(p, "cudabin")
    _write_cudabin(q, db.arrays())
    assert p.read_bytes() == q.read_bytes()
    _same_db(db.arrays(), gm.GraphDB.load(p, "cudabin").arrays())


def test_self_loops_read_alike_in_every_format(tmp_path):
    """A self-loop survives every reader (one edge per pair of adjacency
    entries), so the same database is rejected by the device path the same
    way whatever its file format (the envelope is simple graphs)."""
    t = tmp_path / "loop.txt"
    t.write_text("t # 0\nv 0 1\nv 1 2\ne 0 1 3\ne 1 1 4\n")
    db = gm.GraphDB.load(t, "txt")
    for ft in ("bin", "cudabin"):
        p = tmp_path / f"loop.{ft}"
        db.write(p, ft)
        back = gm.GraphDB.load(p, ft)
        assert _edge_sets(back.arrays()) == _edge_sets(db.arrays()) == [[(0, 1, 3), (1, 1, 4)]]
        with pytest.raises(gm.GMError) as e:
            gm.mine(back, 1)
        assert e.value.code == gm.GM_E_UNSUPPORTED and "self-loop" in str(e.value)


def test_empty_database_bin_write_rejected(tmp_path):
    empty = gm.GraphDB.from_arrays([0], [], [0], [], [], [])
    with pytest.raises(gm.GMError):
        empty.write(tmp_path / "e.bin", "bin")
