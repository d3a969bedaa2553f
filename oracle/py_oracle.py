"""ctypes binding of the CPU oracle (oracle/cpu_gspan.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the product.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path
from typing import List, Tuple

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "oracle" / "build" / "libcpu_gspan.so"
CLI = ROOT / "oracle" / "build" / "cpu_gspan"
REF_CLI = ROOT / "oracle" / "_ref" / "gspan_ref"


class orc_stats(C.Structure):
    _fields_ = [("n_patterns", C.c_uint64), ("e_child", C.c_uint64), ("e_cand", C.c_uint64),
                ("e_root", C.c_uint64), ("b_alg", C.c_uint64), ("n_expanded", C.c_uint64),
                ("max_edges", C.c_uint64), ("seconds", C.c_double), ("b_ext", C.c_uint64),
                ("b_sup", C.c_uint64), ("b_child", C.c_uint64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


CB = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.c_uint32, C.c_uint64)
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(ROOT), "-f", "oracle/Makefile"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = C.CDLL(str(LIB))
        i32p = C.POINTER(C.c_int32)
        l.orc_mine.restype = C.c_int
        l.orc_mine.argtypes = [C.c_int32, i32p, i32p, i32p, i32p, i32p, i32p, C.c_int32, CB, C.c_void_p,
                               C.POINTER(orc_stats)]
        l.orc_mine_capped.restype = C.c_int
        l.orc_mine_capped.argtypes = [C.c_int32, i32p, i32p, i32p, i32p, i32p, i32p, C.c_int32, C.c_int32, CB,
                                      C.c_void_p, C.POINTER(orc_stats)]
        l.orc_mine_sampled.restype = C.c_int
        l.orc_mine_sampled.argtypes = [C.c_int32, i32p, i32p, i32p, i32p, i32p, i32p, C.c_int32, C.c_int32, CB,
                                       C.c_void_p, C.POINTER(orc_stats)]
        l.orc_mine_file.restype = C.c_int
        l.orc_mine_file.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, CB, C.c_void_p, C.POINTER(orc_stats)]
        l.orc_is_min.restype = C.c_int
        l.orc_is_min.argtypes = [i32p, C.c_int32]
        l.orc_support_from_records.restype = C.c_int
        l.orc_support_from_records.argtypes = [i32p, i32p, C.c_int64, C.c_int32, i32p]
        _lib = l
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


Pattern = Tuple[Tuple[Tuple[int, int, int, int, int], ...], int, int]


def mine_arrays(arrays: dict, minsup: int, collect: bool = True, max_edges: int = 0) -> Tuple[List[Pattern], dict]:
    """Mine a database given in gm_db_export's array form.  Patterns are
    (code in the -1 convention, support, |Proj|) in emission order;
    max_edges > 0 stops growth beyond that many edges."""
    a = {k: np.ascontiguousarray(v, dtype=np.int32) for k, v in arrays.items()}
    out: List[Pattern] = []

    def cb(_u, d, n, sup, nproj):
        if collect:
            out.append((tuple(tuple(d[5 * i + j] for j in range(5)) for i in range(n)), int(sup), int(nproj)))

    ccb = CB(cb)
    st = orc_stats()
    rc = lib().orc_mine_capped(len(a["graph_vbegin"]) - 1, _p(a["graph_vbegin"]), _p(a["vlabels"]),
                               _p(a["graph_ebegin"]), _p(a["efrom"]), _p(a["eto"]), _p(a["elabel"]), int(minsup),
                               int(max_edges), ccb, None, C.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_mine failed: {rc}")
    return out, st.as_dict()


def mine_sampled(arrays: dict, minsup: int, root_stride: int) -> dict:
    """Stats of a run that grows only every root_stride-th 1-edge root (in
    the oracle's root order), each subtree in full: a bounded sample of the
    same workload for timing the CPU baseline."""
    a = {k: np.ascontiguousarray(v, dtype=np.int32) for k, v in arrays.items()}
    ccb = CB(lambda *args: None)
    st = orc_stats()
    rc = lib().orc_mine_sampled(len(a["graph_vbegin"]) - 1, _p(a["graph_vbegin"]), _p(a["vlabels"]),
                                _p(a["graph_ebegin"]), _p(a["efrom"]), _p(a["eto"]), _p(a["elabel"]), int(minsup),
                                int(root_stride), ccb, None, C.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_mine_sampled failed: {rc}")
    return st.as_dict()


def mine_file(path, filetype: str, minsup: int) -> Tuple[List[Pattern], dict]:
    out: List[Pattern] = []

    def cb(_u, d, n, sup, nproj):
        out.append((tuple(tuple(d[5 * i + j] for j in range(5)) for i in range(n)), int(sup), int(nproj)))

    ccb = CB(cb)
    st = orc_stats()
    rc = lib().orc_mine_file(str(path).encode(), filetype.encode(), int(minsup), ccb, None, C.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_mine_file failed: {rc}")
    return out, st.as_dict()


def is_min(code) -> bool:
    flat = np.ascontiguousarray(np.array(code, dtype=np.int32).reshape(-1))
    rc = lib().orc_is_min(_p(flat), len(code))
    if rc < 0:
        raise ValueError("bad code")
    return rc == 1


def support_from_records(keys, graphs, n_keys: int) -> np.ndarray:
    k = np.ascontiguousarray(keys, dtype=np.int32)
    g = np.ascontiguousarray(graphs, dtype=np.int32)
    out = np.zeros(n_keys, np.int32)
    if lib().orc_support_from_records(_p(k), _p(g), len(k), n_keys, _p(out)) != 0:
        raise ValueError("bad records")
    return out


def code_str(code) -> str:
    """to_string(false) of a code tuple (src/globals/dfs_code.cpp:199-230)."""
    return ";".join("(%d %d %d %d %d)" % tuple(d) for d in code)


def normalize(code) -> Tuple[Tuple[int, int, int, int, int], ...]:
    """Fully labelled -> the -1 convention of the sequential gSpan."""
    return tuple((d[0], d[1], d[2] if i == 0 else -1, d[3], d[4] if d[0] < d[1] else -1)
                 for i, d in enumerate(code))
