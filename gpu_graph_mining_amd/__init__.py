"""MI355X-native gSpan frequent-subgraph miner.

Python host mirror of the reference's mining-driver API over the C ABI of
``_lib/libgspan_amd.so`` (declared in ``include/gspan_amd.h``):

* ``GraphDB``             -- a graph database (dbio readers, generators);
                             types::graph_database_t + graph_database_cuda
                             (src/globals/dbio.hpp:8-37, cuda_graph_types.hpp)
* ``DFS`` / ``DFSCode``   -- DFS tuples and codes with the reference's string
                             form (src/globals/dfs_code.cpp:109-118, 199-230)
* ``graph_output`` sinks  -- graph_counting_output / graph_file_output
                             (src/globals/graph_output.hpp:9-52)
* ``gspan_cuda``          -- set_database / set_min_support /
                             set_graph_output / set_exec_configurator / run /
                             delete_database_from_device
                             (src/gspan_cuda/gspan_cuda.hpp:282-292)
* ``mine`` / ``mine_batch`` -- one-call helpers returning (code, support, |Proj|).

Every call runs the HIP path; when the native library is missing the import
of this module fails loudly (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Callable, Iterable, List, Optional, Sequence, Tuple

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("GM_LIB_PATH", _PKG / "_lib" / "libgspan_amd.so"))

GM_OK = 0
GM_E_ARG = -1
GM_E_IO = -2
GM_E_OOM = -3
GM_E_DEVICE = -4
GM_E_UNSUPPORTED = -5
GM_E_CALLBACK = -6
GM_LABELS_FULL = 0
GM_LABELS_WILDCARD = 1

EXPORTED_SYMBOLS = (
    "gm_version", "gm_last_error", "gm_cfg_init", "gm_device_count", "gm_db_create", "gm_db_load",
    "gm_db_write_txt", "gm_db_counts", "gm_db_export", "gm_db_upload", "gm_db_destroy", "gm_mine",
    "gm_mine_batch", "gm_dfs_code_is_min", "gm_gen_random", "gm_gen_planted", "gm_gen_molecule", "gm_db_subset",
    "gm_batch_create", "gm_batch_mine", "gm_batch_destroy", "gm_support_queries", "gm_support_query",
    "gm_db_write", "gm_db_load_part", "gm_mine_noalloc", "gm_dfs_codes_are_min", "gm_batch_create_subsets",
)


class gm_cfg(C.Structure):
    _fields_ = [
        ("label_style", C.c_int32), ("max_edges", C.c_int32), ("canon_threads", C.c_int32),
        ("block_threads", C.c_int32), ("mem_budget_bytes", C.c_int64), ("collect_proj", C.c_int32),
        ("candidate_lists", C.c_int32), ("path_pruning", C.c_int32), ("speculate", C.c_int32), ("stage_events", C.c_int32),
        ("chunk_rows", C.c_int32), ("table_bits", C.c_int32), ("canon_host", C.c_int32),
    ]


class gm_stats(C.Structure):
    _fields_ = [
        ("n_patterns", C.c_uint64), ("e_child", C.c_uint64), ("e_root", C.c_uint64), ("e_cand", C.c_uint64),
        ("n_candidates", C.c_uint64), ("n_levels", C.c_uint64), ("seconds_total", C.c_double),
        ("seconds_device", C.c_double), ("seconds_canon", C.c_double), ("seconds_ext_kernel", C.c_double),
        ("ext_kernel_launches", C.c_uint64), ("peak_device_bytes", C.c_uint64), ("split_passes", C.c_uint64),
        ("max_graph_rows", C.c_uint64), ("max_table_fill", C.c_uint64),
        ("pruned", C.c_uint64), ("seconds_stage_events", C.c_double), ("canon_host_keys", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


PATTERN_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.c_uint32, C.c_uint64)
JOB_PATTERN_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.c_int32, C.c_uint32,
                             C.c_uint64)

_lib: Optional[C.CDLL] = None


def load_library(path: Optional[os.PathLike] = None) -> C.CDLL:
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(f"native library {p} is missing: run `python -m gpu_graph_mining_amd.build`")
    lib = C.CDLL(str(p))
    P = C.POINTER
    i32p, dbp = P(C.c_int32), C.c_void_p
    sig = {
        "gm_version": (C.c_char_p, []),
        "gm_last_error": (C.c_char_p, []),
        "gm_cfg_init": (None, [P(gm_cfg)]),
        "gm_device_count": (C.c_int, []),
        "gm_db_create": (C.c_int, [C.c_int32, i32p, i32p, i32p, i32p, i32p, i32p, P(dbp)]),
        "gm_db_load": (C.c_int, [C.c_char_p, C.c_char_p, P(dbp)]),
        "gm_db_write_txt": (C.c_int, [dbp, C.c_char_p]),
        "gm_db_write": (C.c_int, [dbp, C.c_char_p, C.c_char_p]),
        "gm_db_load_part": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, P(dbp)]),
        "gm_db_counts": (C.c_int, [dbp, P(C.c_int64)]),
        "gm_db_export": (C.c_int, [dbp, i32p, i32p, i32p, i32p, i32p, i32p]),
        "gm_db_upload": (C.c_int, [dbp]),
        "gm_db_destroy": (None, [dbp]),
        "gm_mine": (C.c_int, [dbp, C.c_int32, P(gm_cfg), PATTERN_CB, C.c_void_p, P(gm_stats)]),
        "gm_mine_noalloc": (C.c_int, [dbp, C.c_int32, P(gm_cfg), PATTERN_CB, C.c_void_p, P(gm_stats)]),
        "gm_mine_batch": (C.c_int, [C.c_int32, P(dbp), i32p, P(gm_cfg), JOB_PATTERN_CB, C.c_void_p, P(gm_stats)]),
        "gm_dfs_code_is_min": (C.c_int, [i32p, C.c_int32]),
        "gm_dfs_codes_are_min": (C.c_int, [C.c_int32, C.c_int32, i32p, P(C.c_uint8)]),
        "gm_gen_random": (C.c_int, [C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_int32,
                                    P(dbp)]),
        "gm_gen_planted": (C.c_int, [C.c_uint64, C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_int32, C.c_int32,
                                     C.c_double, P(dbp)]),
        "gm_gen_molecule": (C.c_int, [C.c_uint64, C.c_int32, C.c_double, C.c_int32, C.c_int32, P(C.c_double),
                                      C.c_int32, P(C.c_double), P(dbp)]),
        "gm_db_subset": (C.c_int, [dbp, i32p, C.c_int32, P(dbp)]),
        "gm_batch_create": (C.c_int, [C.c_int32, P(dbp), P(dbp)]),
        "gm_batch_create_subsets": (C.c_int, [dbp, C.c_int32, i32p, i32p, P(dbp)]),
        "gm_batch_mine": (C.c_int, [dbp, i32p, P(gm_cfg), JOB_PATTERN_CB, C.c_void_p, P(gm_stats)]),
        "gm_batch_destroy": (None, [dbp]),
        "gm_support_queries": (C.c_int, [dbp, C.c_int32, i32p, i32p, P(gm_cfg), P(C.c_uint32), P(gm_stats)]),
        "gm_support_query": (C.c_int, [dbp, i32p, C.c_int32, P(C.c_uint32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class GMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class _Guard:
    """Wraps a pattern callback so that no Python exception is lost in
    ctypes (which would print it and return 0): the exception is kept, the
    callback returns 1 (stop mining) and raise_pending() re-raises it after
    the library call returned."""

    def __init__(self, fn):
        self.fn = fn
        self.error: Optional[BaseException] = None

    def __call__(self, *args):
        try:
            return self.fn(*args)
        except BaseException as e:  # noqa: BLE001 -- re-raised by raise_pending
            if self.error is None:
                self.error = e
            return 1

    def raise_pending(self) -> None:
        if self.error is not None:
            raise self.error


def _check(rc: int) -> None:
    if rc != GM_OK:
        raise GMError(rc, load_library().gm_last_error().decode(errors="replace"))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def device_count() -> int:
    return int(load_library().gm_device_count())


def version() -> str:
    return load_library().gm_version().decode()


# --------------------------------------------------------------------- DFS
class DFS(tuple):
    """(from, to, fromlabel, elabel, tolabel) -- src/globals/dfs_code.hpp:16-68."""

    def __new__(cls, frm, to, fromlabel, elabel, tolabel):
        return super().__new__(cls, (int(frm), int(to), int(fromlabel), int(elabel), int(tolabel)))

    frm = property(lambda s: s[0])
    to = property(lambda s: s[1])
    fromlabel = property(lambda s: s[2])
    elabel = property(lambda s: s[3])
    tolabel = property(lambda s: s[4])

    def is_forward(self) -> bool:
        return self[0] < self[1]

    def to_string(self, print_edge_type: bool = True) -> str:
        body = "(%d %d %d %d %d)" % tuple(self)
        return (("F" if self.is_forward() else "B") + body) if print_edge_type else body

    @staticmethod
    def parse_from_string(s: str) -> "DFS":
        parts = s.strip().strip("FB").strip("()").split()
        if len(parts) != 5:
            raise ValueError("error: could not read dfs code from string")
        return DFS(*map(int, parts))


class DFSCode(tuple):
    """A sequence of DFS tuples -- src/globals/dfs_code.hpp:150-211."""

    def __new__(cls, tuples: Iterable[Sequence[int]]):
        return super().__new__(cls, tuple(DFS(*t) for t in tuples))

    def to_string(self, print_edge_type: bool = True) -> str:
        return ";".join(d.to_string(print_edge_type) for d in self)

    @staticmethod
    def read_from_str(s: str) -> "DFSCode":
        return DFSCode(DFS.parse_from_string(x) for x in s.split(";") if x.strip())

    def wildcard(self) -> "DFSCode":
        """The sequential gSpan's labelling: fromlabel -1 after the first
        tuple, tolabel -1 on backward tuples (src/gspan/gspan.cpp:253, 266)."""
        return DFSCode((d[0], d[1], d[2] if i == 0 else -1, d[3], d[4] if d[0] < d[1] else -1)
                       for i, d in enumerate(self))

    def build_rmpath(self) -> List[int]:
        rm, old_from = [], -1
        for i in range(len(self) - 1, -1, -1):
            d = self[i]
            if d[0] < d[1] and (not rm or old_from == d[1]):
                rm.append(i)
                old_from = d[0]
        return rm

    def is_min(self) -> bool:
        """Minimum-DFS-code check (src/globals/dfs_ismin.cpp:23-160)."""
        flat = _i32(np.array(self, dtype=np.int32).reshape(-1))
        rc = load_library().gm_dfs_code_is_min(_ptr(flat), len(self))
        if rc < 0:
            _check(rc)
        return rc == 1


def codes_are_min(codes: Sequence[Sequence[Sequence[int]]]) -> List[bool]:
    """Minimum-DFS-code check of many codes on the device (one wave per
    code; gm_dfs_codes_are_min), grouped by length; the same answers as
    DFSCode.is_min."""
    codes = [tuple(tuple(int(x) for x in t) for t in c) for c in codes]
    out = [False] * len(codes)
    by_len: dict = {}
    for i, c in enumerate(codes):
        by_len.setdefault(len(c), []).append(i)
    lib = load_library()
    for n, idx in by_len.items():
        if n == 0:
            raise GMError(GM_E_ARG, "empty code")
        flat = _i32(np.array([codes[i] for i in idx], dtype=np.int32).reshape(-1))
        res = np.zeros(len(idx), dtype=np.uint8)
        _check(lib.gm_dfs_codes_are_min(len(idx), n, _ptr(flat), res.ctypes.data_as(C.POINTER(C.c_uint8))))
        for k, i in enumerate(idx):
            out[i] = bool(res[k])
    return out


# ------------------------------------------------------------------ GraphDB
class GraphDB:
    """A graph database held by the native library (host copy + lazily
    uploaded device layout)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    # construction ----------------------------------------------------------
    @classmethod
    def from_arrays(cls, graph_vbegin, vlabels, graph_ebegin, efrom, eto, elabel) -> "GraphDB":
        vb, vl, eb, ef, et, el = map(_i32, (graph_vbegin, vlabels, graph_ebegin, efrom, eto, elabel))
        h = C.c_void_p()
        _check(load_library().gm_db_create(len(vb) - 1, _ptr(vb), _ptr(vl), _ptr(eb), _ptr(ef), _ptr(et), _ptr(el),
                                           C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_graphs(cls, graphs: Sequence[Tuple[Sequence[int], Sequence[Tuple[int, int, int]]]]) -> "GraphDB":
        """graphs: [(vertex_labels, [(from, to, elabel), ...]), ...]"""
        vb, vl, eb, ef, et, el = [0], [], [0], [], [], []
        for labels, edges in graphs:
            vl.extend(labels)
            vb.append(len(vl))
            for a, b, l in edges:
                ef.append(a)
                et.append(b)
                el.append(l)
            eb.append(len(ef))
        return cls.from_arrays(vb, vl, eb, ef, et, el)

    @classmethod
    def from_codes(cls, codes: Iterable[str]) -> "GraphDB":
        """One graph per DFS-code string, as DFSCode::toGraph builds it
        (src/globals/dfs_code.cpp:245-265)."""
        graphs = []
        for s in codes:
            code = DFSCode.read_from_str(s)
            nv = max(max(d[0], d[1]) for d in code) + 1
            labels = [0] * nv
            edges = []
            for d in code:
                if d[2] != -1:
                    labels[d[0]] = d[2]
                if d[4] != -1:
                    labels[d[1]] = d[4]
                edges.append((d[0], d[1], d[3]))
            graphs.append((labels, edges))
        return cls.from_graphs(graphs)

    @classmethod
    def load(cls, path: os.PathLike, filetype: str = "txt") -> "GraphDB":
        h = C.c_void_p()
        _check(load_library().gm_db_load(str(path).encode(), filetype.encode(), C.byref(h)))
        return cls(h.value)

    @classmethod
    def load_part(cls, path: os.PathLike, total_parts: int, part: int) -> "GraphDB":
        """Part `part` of `total_parts` of a BIN file (dbio::read_database_bin
        with parts, src/globals/dbio.cpp:275-329)."""
        h = C.c_void_p()
        _check(load_library().gm_db_load_part(str(path).encode(), int(total_parts), int(part), C.byref(h)))
        return cls(h.value)

    @classmethod
    def random(cls, seed: int, n_graphs: int, vmin: int, vmax: int, edge_factor: float = 1.2, n_vlabels: int = 5,
               n_elabels: int = 3) -> "GraphDB":
        h = C.c_void_p()
        _check(load_library().gm_gen_random(seed, n_graphs, vmin, vmax, edge_factor, n_vlabels, n_elabels,
                                            C.byref(h)))
        return cls(h.value)

    @classmethod
    def planted(cls, seed: int, n_graphs: int, mean_edges: float = 30.0, n_seed_patterns: int = 50,
                mean_pattern_edges: float = 4.0, n_vlabels: int = 5, n_elabels: int = 5,
                merge_prob: float = 0.5) -> "GraphDB":
        h = C.c_void_p()
        _check(load_library().gm_gen_planted(seed, n_graphs, mean_edges, n_seed_patterns, mean_pattern_edges,
                                             n_vlabels, n_elabels, merge_prob, C.byref(h)))
        return cls(h.value)

    @classmethod
    def molecule(cls, seed: int, n_graphs: int, mean_vertices: float = 25.0, max_degree: int = 4,
                 vlabel_probs=(.70, .12, .10, .03, .02, .01, .01, .01), elabel_probs=(.85, .12, .03)) -> "GraphDB":
        vp = (C.c_double * len(vlabel_probs))(*vlabel_probs)
        ep = (C.c_double * len(elabel_probs))(*elabel_probs)
        h = C.c_void_p()
        _check(load_library().gm_gen_molecule(seed, n_graphs, mean_vertices, max_degree, len(vlabel_probs), vp,
                                              len(elabel_probs), ep, C.byref(h)))
        return cls(h.value)

    def subset(self, ids: Sequence[int]) -> "GraphDB":
        a = _i32(ids)
        h = C.c_void_p()
        _check(load_library().gm_db_subset(self._h, _ptr(a), len(a), C.byref(h)))
        return GraphDB(h.value)

    # inspection ----------------------------------------------------------------
    def counts(self) -> dict:
        c = (C.c_int64 * 6)()
        _check(load_library().gm_db_counts(self._h, c))
        keys = ("n_graphs", "n_vertices", "n_edges", "max_vertices", "n_vlabels", "n_elabels")
        return dict(zip(keys, map(int, c)))

    def __len__(self) -> int:
        return self.counts()["n_graphs"]

    def arrays(self) -> dict:
        c = self.counts()
        g, nv, ne = c["n_graphs"], c["n_vertices"], c["n_edges"]
        out = {"graph_vbegin": np.zeros(g + 1, np.int32), "vlabels": np.zeros(nv, np.int32),
               "graph_ebegin": np.zeros(g + 1, np.int32), "efrom": np.zeros(ne, np.int32),
               "eto": np.zeros(ne, np.int32), "elabel": np.zeros(ne, np.int32)}
        _check(load_library().gm_db_export(self._h, *(_ptr(out[k]) for k in
                                                       ("graph_vbegin", "vlabels", "graph_ebegin", "efrom", "eto",
                                                        "elabel"))))
        return out

    def sha256(self) -> str:
        import hashlib
        h = hashlib.sha256()
        for k, v in self.arrays().items():
            h.update(k.encode())
            h.update(v.astype("<i4").tobytes())
        return h.hexdigest()

    def write_txt(self, path: os.PathLike) -> None:
        _check(load_library().gm_db_write_txt(self._h, str(path).encode()))

    def write(self, path: os.PathLike, filetype: str = "txt") -> None:
        """Write as txt, bin (dbio::write_database_bin) or cudabin
        (dbio::write_database_cudabin)."""
        _check(load_library().gm_db_write(self._h, str(path).encode(), filetype.encode()))

    def upload(self) -> None:
        _check(load_library().gm_db_upload(self._h))

    def close(self) -> None:
        if self._h and self._h.value:
            load_library().gm_db_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- outputs
class graph_output:
    """Pattern sink (src/globals/graph_output.hpp:9-16)."""

    def output_graph(self, code: DFSCode, support: int = 0) -> None:
        raise NotImplementedError

    def get_count(self) -> float:
        raise NotImplementedError

    def to_string(self) -> str:
        return type(self).__name__


class graph_counting_output(graph_output):
    def __init__(self):
        self.counter = 0.0

    def output_graph(self, code, support=0):
        self.counter += 1.0

    def get_count(self):
        return self.counter


def dfs_order_key(code) -> tuple:
    """Sort key of the sequential gSpan's report order (depth first; the
    children of a pattern in the order of its extension maps,
    src/gspan/gspan.cpp:251-271: backward tuples by `to` then edge label,
    before forward tuples by `from` descending, then edge label, then
    to-label; roots by (fromlabel, elabel, tolabel)).  A pattern precedes
    its extensions, so sorting by this key gives the pre-order."""
    key = []
    for i, d in enumerate(code):
        if i == 0:
            key.append((d[2], d[3], d[4]))
        elif d[0] > d[1]:
            key.append((0, d[1], d[3], 0))
        else:
            key.append((1, -d[0], d[3], d[4]))
    return tuple(key)


class graph_file_output(graph_output):
    """One ``code.to_string(false)`` line per pattern
    (src/globals/graph_output.cpp:90-95), as they come or (dfs_order=True)
    all at close() in the sequential gSpan's report order, which with
    wildcard labels makes the file byte-identical to that program's."""

    def __init__(self, fname: os.PathLike, dfs_order: bool = False):
        self.filename = str(fname)
        self._f = open(self.filename, "w")
        self.count = 0.0
        self.dfs_order = dfs_order
        self._held: list = []

    def output_graph(self, code, support=0):
        if self.dfs_order:
            self._held.append(code)
        else:
            self._f.write(code.to_string(False) + "\n")
        self.count += 1.0

    def get_count(self):
        return self.count

    def close(self):
        if self._f.closed:
            return
        if self.dfs_order:
            for c in sorted(self._held, key=dfs_order_key):
                self._f.write(DFSCode(c).to_string(False) + "\n")
            self._held = []
        self._f.close()

    def to_string(self):
        return f"graph_file_output, filename: {self.filename}"


class graph_collecting_output(graph_output):
    """Keeps (code, support, |Proj|) triples in emission order."""

    def __init__(self):
        self.patterns: List[Tuple[DFSCode, int, int]] = []

    def output_graph(self, code, support=0, n_proj=0):
        self.patterns.append((code, int(support), int(n_proj)))

    def get_count(self):
        return float(len(self.patterns))


# ------------------------------------------------------------------- mining
def _make_cfg(label_style=GM_LABELS_FULL, max_edges=0, canon_threads=0, collect_proj=False, candidate_lists=True,
              block_threads=0, path_pruning=True, speculate=0, stage_events=False, chunk_rows=0,
              table_bits=0, canon_host=0, mem_budget=0) -> gm_cfg:
    cfg = gm_cfg()
    load_library().gm_cfg_init(C.byref(cfg))
    cfg.label_style = label_style
    cfg.max_edges = max_edges
    cfg.canon_threads = canon_threads
    cfg.collect_proj = 1 if collect_proj else 0
    cfg.candidate_lists = 0 if candidate_lists else 1
    cfg.block_threads = block_threads
    cfg.path_pruning = 0 if path_pruning else 1
    cfg.speculate = speculate
    cfg.stage_events = 1 if stage_events else 0
    cfg.chunk_rows = chunk_rows
    cfg.table_bits = table_bits
    cfg.canon_host = int(canon_host)
    cfg.mem_budget_bytes = int(mem_budget)
    return cfg


def mine(db: GraphDB, minsup: int, *, label_style: int = GM_LABELS_FULL, max_edges: int = 0,
         collect_proj: bool = False, canon_threads: int = 0,
         on_pattern: Optional[Callable[[DFSCode, int, int], Optional[bool]]] = None, count_only: bool = False,
         candidate_lists: bool = True, block_threads: int = 0, path_pruning: bool = True,
         speculate: int = 0, stage_events: bool = False, chunk_rows: int = 0, table_bits: int = 0,
         canon_host: int = 0, mem_budget: int = 0):
    """Mine db on the GPU.  Returns (patterns, stats) where patterns is a
    list of (DFSCode, support, |Proj|) in emission order (empty when
    on_pattern is given; it then receives each pattern, and returning False
    stops mining; empty with count_only, which reports nothing back and
    only fills the statistics).  candidate_lists=False makes the write pass
    enumerate extensions again instead of reading the support pass's lists;
    block_threads fixes the workgroup size (64, 128 or 256; 0 = automatic);
    path_pruning=False keeps the candidates whose new edge forms an
    infrequent 2-edge path; speculate (0 adaptive, 1 never, 2 always)
    controls whether a level's children are written before their canonical
    check, which runs on the host while the device writes them or, on the
    levels written after the check, on the device (canon_host=0); 1 keeps it
    on the host, 2 on the device; the results are the same either way;
    mem_budget bounds the device bytes of the embedding rows and candidate
    lists held at once (levels are then mined in pieces; 0 = unlimited);
    stage_events=True times
    every level's extension + support stage with HIP events on the mining
    stream (stats["seconds_stage_events"])."""
    lib = load_library()
    out: list = []

    def cb(_user, dfs, n, sup, nproj):
        code = DFSCode(tuple(dfs[5 * i + j] for j in range(5)) for i in range(n))
        if on_pattern is not None:
            return 1 if on_pattern(code, int(sup), int(nproj)) is False else 0
        out.append((code, int(sup), int(nproj)))
        return 0

    guard = _Guard(cb)
    ccb = PATTERN_CB() if count_only else PATTERN_CB(guard)  # PATTERN_CB() is a NULL pointer
    cfg = _make_cfg(label_style, max_edges, canon_threads, collect_proj, candidate_lists, block_threads,
                    path_pruning, speculate, stage_events, chunk_rows, table_bits, canon_host, mem_budget)
    st = gm_stats()
    rc = lib.gm_mine(db._h, int(minsup), C.byref(cfg), ccb, None, C.byref(st))
    guard.raise_pending()
    _check(rc)
    return out, st.as_dict()


def mine_noalloc(db: GraphDB, minsup: int, *, label_style: int = GM_LABELS_FULL, collect_proj: bool = False,
                 count_only: bool = False):
    """mine() through gm_mine_noalloc: no device allocation, every buffer of
    the level steps from the library's static 64 MiB device array (the
    database must be uploaded first).  Returns (patterns, stats) like
    mine()."""
    lib = load_library()
    out: list = []

    def cb(_user, dfs, n, sup, nproj):
        out.append((DFSCode(tuple(dfs[5 * i + j] for j in range(5)) for i in range(n)), int(sup), int(nproj)))
        return 0

    guard = _Guard(cb)
    ccb = PATTERN_CB() if count_only else PATTERN_CB(guard)
    cfg = _make_cfg(label_style, 0, 0, collect_proj)
    st = gm_stats()
    rc = lib.gm_mine_noalloc(db._h, int(minsup), C.byref(cfg), ccb, None, C.byref(st))
    guard.raise_pending()
    _check(rc)
    return out, st.as_dict()


def mine_batch(dbs: Sequence[GraphDB], minsups: Sequence[int], *, label_style: int = GM_LABELS_FULL,
               collect_proj: bool = False, canon_threads: int = 0, count_only: bool = False):
    """Mine several independent databases in one batched schedule.  Returns
    (per_job_patterns, stats); with count_only nothing is reported back."""
    lib = load_library()
    res: List[list] = [[] for _ in dbs]

    def cb(_user, job, dfs, n, sup, nproj):
        res[job].append((DFSCode(tuple(dfs[5 * i + j] for j in range(5)) for i in range(n)), int(sup), int(nproj)))
        return 0

    guard = _Guard(cb)
    ccb = JOB_PATTERN_CB() if count_only else JOB_PATTERN_CB(guard)
    handles = (C.c_void_p * len(dbs))(*[d._h.value for d in dbs])
    ms = _i32(minsups)
    cfg = _make_cfg(label_style, 0, canon_threads, collect_proj)
    st = gm_stats()
    rc = lib.gm_mine_batch(len(dbs), handles, _ptr(ms), C.byref(cfg), ccb, None, C.byref(st))
    guard.raise_pending()
    _check(rc)
    return res, st.as_dict()


def support(db: GraphDB, codes: Sequence[Sequence[Sequence[int]]], *, with_stats: bool = False):
    """Support (number of graphs containing the pattern) of each query code
    -- any valid DFS code of a connected, fully labelled pattern -- from one
    batched projection over the resident database (gm_support_queries).
    Returns a list of ints (and the run statistics with with_stats)."""
    lib = load_library()
    n = len(codes)
    lens = np.array([len(c) for c in codes], dtype=np.int32)
    flat = np.array([v for c in codes for t in c for v in t], dtype=np.int32)
    if flat.size != 5 * int(lens.sum()):
        raise ValueError("every DFS tuple needs 5 integers")
    out = np.zeros(max(n, 1), dtype=np.uint32)
    cfg = _make_cfg()
    st = gm_stats()
    _check(lib.gm_support_queries(db._h, n, _ptr(flat) if flat.size else None, _ptr(lens) if n else None,
                                  C.byref(cfg), out.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(st)))
    res = [int(x) for x in out[:n]]
    return (res, st.as_dict()) if with_stats else res


class Batch:
    """Databases of several jobs resident on the device in one packed
    layout (gm_batch_create); mine() runs all jobs in one schedule."""

    def __init__(self, dbs: Sequence[GraphDB]):
        lib = load_library()
        handles = (C.c_void_p * len(dbs))(*[d._h.value for d in dbs])
        h = C.c_void_p()
        _check(lib.gm_batch_create(len(dbs), handles, C.byref(h)))
        self._h = h
        self.n_jobs = len(dbs)

    @classmethod
    def from_subsets(cls, db: GraphDB, subsets: Sequence[Sequence[int]]) -> "Batch":
        """Jobs that are subsets of the resident database `db` (graph ids
        per job, in order; they may overlap), gathered on the device from
        db's own device copy (gm_batch_create_subsets)."""
        lib = load_library()
        begin = [0]
        for ids in subsets:
            begin.append(begin[-1] + len(ids))
        flat = _i32([g for ids in subsets for g in ids] or [0])
        b = _i32(begin)
        h = C.c_void_p()
        _check(lib.gm_batch_create_subsets(db._h, len(subsets), _ptr(b), _ptr(flat), C.byref(h)))
        self = cls.__new__(cls)
        self._h = h
        self.n_jobs = len(subsets)
        return self

    def mine(self, minsups: Sequence[int], *, label_style: int = GM_LABELS_FULL, collect_proj: bool = False,
             canon_threads: int = 0, count_only: bool = False, path_pruning: bool = True, speculate: int = 0,
             canon_host: int = 0, mem_budget: int = 0, chunk_rows: int = 0, table_bits: int = 0):
        lib = load_library()
        res: List[list] = [[] for _ in range(self.n_jobs)]

        def cb(_user, job, dfs, n, sup, nproj):
            res[job].append((DFSCode(tuple(dfs[5 * i + j] for j in range(5)) for i in range(n)), int(sup),
                             int(nproj)))
            return 0

        guard = _Guard(cb)
        ccb = JOB_PATTERN_CB() if count_only else JOB_PATTERN_CB(guard)
        ms = _i32(minsups)
        cfg = _make_cfg(label_style, 0, canon_threads, collect_proj, path_pruning=path_pruning, speculate=speculate,
                        canon_host=canon_host, mem_budget=mem_budget, chunk_rows=chunk_rows, table_bits=table_bits)
        st = gm_stats()
        rc = lib.gm_batch_mine(self._h, _ptr(ms), C.byref(cfg), ccb, None, C.byref(st))
        guard.raise_pending()
        _check(rc)
        return res, st.as_dict()

    def close(self) -> None:
        if self._h and self._h.value:
            load_library().gm_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class gspan_cuda:
    """The reference's driver class (src/gspan_cuda/gspan_cuda.hpp:53-293)
    over the HIP path: same method names, same call order."""

    def __init__(self):
        self._db: Optional[GraphDB] = None
        self._minsup = -1
        self._out: Optional[graph_output] = None
        self._cfg: Optional[dict] = None
        self.stats: dict = {}

    def set_database(self, db: GraphDB, convert_to_cuda: bool = True) -> None:
        self._db = db
        if convert_to_cuda:
            db.upload()

    def set_min_support(self, minsup: int) -> None:
        self._minsup = int(minsup)

    def set_graph_output(self, gout: graph_output) -> None:
        self._out = gout

    def set_exec_configurator(self, cfg: Optional[dict]) -> None:
        """Tuning options (label_style, max_edges, canon_threads); replaces
        cudapp::cuda_configurator."""
        self._cfg = dict(cfg or {})

    def run(self) -> None:
        if self._db is None or self._out is None:
            raise RuntimeError("database and graph output must be set before run()")
        if self._cfg is None:
            raise RuntimeError("exec configurator must be set before run()")  # gspan_cuda.cpp:127
        out = self._out
        collect = isinstance(out, graph_collecting_output)

        def on_pattern(code, sup, nproj):
            if collect:
                out.output_graph(code, sup, nproj)
            else:
                out.output_graph(code, sup)

        _, self.stats = mine(self._db, self._minsup, on_pattern=on_pattern, **self._cfg)

    def delete_database_from_device(self) -> None:
        if self._db is not None:
            self._db.close()
            self._db = None


__all__ = [
    "GraphDB", "Batch", "DFS", "DFSCode", "graph_output", "graph_counting_output", "graph_file_output",
    "dfs_order_key",
    "graph_collecting_output", "gspan_cuda", "mine", "mine_batch", "load_library", "device_count", "version",
    "GMError", "codes_are_min", "GM_LABELS_FULL", "GM_LABELS_WILDCARD", "EXPORTED_SYMBOLS",
]
