"""The entry point that allocates nothing (gm_mine_noalloc): every buffer
of its level steps comes from the library's static device array.  Checked
against gm_mine, the entry point it is a variant of, on two calls in a row
-- the second with a smaller database -- and on an input whose level steps
need more than the array holds (a clear GM_E_OOM)."""
import pytest

import gpu_graph_mining_amd as gm
from gpu_graph_mining_amd import workloads as wl

from helpers import GOLDEN, listing

pytestmark = pytest.mark.gpu


def _same(db, minsup):
    want, wst = gm.mine(db, minsup, collect_proj=True)
    got, gst = gm.mine_noalloc(db, minsup, collect_proj=True)
    assert sorted((tuple(c), s, n) for c, s, n in got) == sorted((tuple(c), s, n) for c, s, n in want)
    for k in ("n_patterns", "e_child", "e_root", "e_cand"):
        assert gst[k] == wst[k], k
    assert 0 < gst["peak_device_bytes"] <= 64 << 20
    return listing(got)


def test_no_alloc_equals_mine_on_two_calls():
    big = gm.GraphDB.load(GOLDEN / "Chemical_340.txt", "txt")
    big.upload()
    first = _same(big, 34)
    small = big.subset(list(range(0, 340, 3)))  # 114 graphs
    small.upload()
    second = _same(small, 12)
    assert first != second
    # the array keeps the last call's contents: a third call is still exact
    assert _same(big, 34) == first


def test_no_alloc_needs_an_uploaded_database():
    db, ms = wl.config0(n_graphs=50)
    with pytest.raises(gm.GMError) as e:
        gm.mine_noalloc(db, ms)
    assert e.value.code == gm.GM_E_ARG


def test_no_alloc_rejects_work_beyond_its_array():
    db, ms = wl.config1(n_graphs=10000)
    db.upload()
    with pytest.raises(gm.GMError) as e:
        gm.mine_noalloc(db, ms, count_only=True)
    assert e.value.code == gm.GM_E_OOM and "static workspace" in str(e.value)
    # and the next call is unaffected
    small, ms0 = wl.config0(n_graphs=200)
    small.upload()
    _same(small, ms0)
