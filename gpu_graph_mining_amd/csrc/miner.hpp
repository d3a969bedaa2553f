// miner.hpp -- device database and the level-synchronous mining driver.
#pragma once
#include <hip/hip_runtime.h>

#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <vector>

#include "gspan_amd.h"
#include "device.hpp"
#include "gm_internal.hpp"

namespace gm {

void hip_check(hipError_t e, const char *what);

// A fixed device region (the static workspace of the entry point that
// allocates nothing) used as two stacks: side 0 grows up from the bottom,
// side 1 down from the top.  The rows of a level live on side (level id
// parity, the Workspace's double-buffer index); while a level step runs,
// its other buffers (scratch) are stacked on the side that holds its rows,
// and the other side is emptied for the children's rows.  Everything on one
// stream, so a region is only reused by work ordered after its last use.
struct Arena {
  char *base = nullptr;
  size_t cap = 0;
  size_t lo = 0, hi = 0;   // side 0 used [0, lo), side 1 used [hi, cap)
  size_t lo0 = 0;          // side 0's origin (after the call's permanent buffers)
  int scratch_side = 0;    // where scratch goes during the current level
  size_t peak = 0;         // most bytes in use at once
  void *take(int side, size_t bytes);  // Error(GM_E_OOM) when the stacks meet
  void reset(int side) {
    if (side == 0)
      lo = lo0;
    else
      hi = cap;
  }
};

// Growable device allocation (grows only; freed with the owner).  With an
// arena, the memory is taken from it instead of hipMalloc: from side
// `side` (0 / 1), or from the current scratch side (side < 0).
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  Arena *arena = nullptr;
  int side = -1;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release();
  void *ensure(size_t bytes);
  template <class T>
  T *as() const { return static_cast<T *>(p); }
};

// Growable pinned host allocation, mapped into the device's address space:
// the level step's copy kernels read and write it on the compute queue.
// With an arena: ordinary host memory `p` and a device copy `dev` taken from
// the arena, moved by explicit copies.
struct HostBuf {
  void *p = nullptr;
  void *dev = nullptr;  // the same memory as the device addresses it (mapped), or its device copy
  size_t cap = 0;
  Arena *arena = nullptr;
  std::vector<char> host;  // arena mode: the host side
  HostBuf() = default;
  HostBuf(const HostBuf &) = delete;
  HostBuf &operator=(const HostBuf &) = delete;
  ~HostBuf();
  void *ensure(size_t bytes);
  template <class T>
  T *as() const { return static_cast<T *>(p); }
};

// One level's rows (pid/rgraph/vt/rbound/rlist) and its pattern upload
// (metadata, path masks, codes).
struct RowSet {
  DevBuf pid, rgraph, vt, rbound, rlist, upl;
  std::vector<DevBuf *> bufs() { return {&pid, &rgraph, &vt, &rbound, &rlist, &upl}; }
  size_t bytes() const { return pid.cap + rgraph.cap + vt.cap + rbound.cap + rlist.cap + upl.cap; }
};

// Device buffers of the level steps, kept across mining calls on one
// database so that repeated calls allocate nothing once warm.
struct Workspace {
  DevBuf minsup, sup, acc, accb, freq, tiles, rccnt, cur_rows, cur_list, cursor, fail, chunks, chunk_keys, cand;
  // row sets: a level's rows stay in one while its children are written to
  // another; two alternate, more are held while levels are mined in pieces
  // under a memory budget
  std::deque<RowSet> rows{2};
  RowSet &row(int i) {
    while ((int)rows.size() <= i) rows.emplace_back();
    return rows[(size_t)i];
  }
  HostBuf up, down;
  hipEvent_t ev[2] = {nullptr, nullptr};  // stage timing (MineOptions::stage_events)
  // the children's upload runs on a copy engine beside the write pass:
  // its stream (default priority) and the event the next level waits on
  hipStream_t side = nullptr;
  hipEvent_t up_done = nullptr;
  // the packed counters hold zeros outside a level step (k_reduce clears
  // what it reads); false after a call that ended early
  bool acc_zero = false;
  // every buffer from this region (no device allocation, no pinned host
  // memory, no events); null: ordinary allocations
  Arena *arena = nullptr;
  void use_arena(Arena *ar) {
    arena = ar;
    for (DevBuf *b : scratch_bufs()) b->arena = ar;
    for (int i = 0; i < 2; ++i)
      for (DevBuf *b : row_bufs(i)) {
        b->arena = ar;
        b->side = i;
      }
    up.arena = ar;
    down.arena = ar;
  }
  // Arena workspace, at the start of a level step whose rows are on side
  // `cur`: the scratch buffers are taken again on that side, the other side
  // is emptied for the children's rows.
  void arena_level(int cur) {
    arena->reset(cur ^ 1);
    arena->scratch_side = cur;
    for (DevBuf *b : scratch_bufs()) {
      b->p = nullptr;
      b->cap = 0;
    }
    for (DevBuf *b : row_bufs(cur ^ 1)) {
      b->p = nullptr;
      b->cap = 0;
    }
    for (HostBuf *h : {&up, &down}) {
      h->dev = nullptr;
      h->p = nullptr;
      h->cap = 0;
    }
  }
  std::vector<DevBuf *> scratch_bufs() {
    return {&sup, &acc, &accb, &freq, &tiles, &rccnt, &cur_rows, &cur_list, &cursor, &fail, &chunks, &chunk_keys, &cand};
  }
  std::vector<DevBuf *> row_bufs(int i) { return row(i).bufs(); }
  ~Workspace() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    if (up_done) (void)hipEventDestroy(up_done);
    if (side) (void)hipStreamDestroy(side);
  }
};

// The packed database of device.hpp on the GPU, for one or more jobs.
struct DeviceDB {
  int32_t n_graphs = 0, n_jobs = 1;
  int64_t n_vertices = 0, n_adj = 0;
  int32_t max_graph_v = 0, max_graph_adj = 0;  // largest graph: vertices, adjacency entries
  int32_t NVL = 0, NEL = 0;
  std::vector<int32_t> vl_values, el_values;  // dense label -> original
  std::vector<int32_t> job_first_graph;       // n_jobs + 1
  std::vector<int64_t> job_first_vertex;      // n_jobs + 1 (the root level's rows of each job)
  DevBuf gvbase, vadj, adj, vlab, gjob, vgraph;
  size_t bytes = 0;
  // mining state; one mining call at a time per database
  std::unique_ptr<Workspace> ws;
  std::mutex mu;
};

// Host image of the device layout (device.hpp).
struct PackedDB {
  int32_t max_graph_v = 0, max_graph_adj = 0;
  std::vector<int32_t> gvbase, vadj, gjob, vgraph;
  std::vector<uint32_t> adj;
  std::vector<uint8_t> vlab;
};

// Validate the databases (one job each) for the device path and pack them:
// dense labels, packed adjacency.  Throws Error(GM_E_UNSUPPORTED) for graphs
// over 256 vertices, self-loops, parallel edges or more than 255 labels.
void pack_db(const std::vector<const HostDB *> &dbs, DeviceDB &out, PackedDB &pk);
// Copy a packed database to the GPU.
void upload_db(const PackedDB &pk, DeviceDB &out, hipStream_t s);
// A database of jobs whose graphs are subsets of a resident one, built on
// the device (no host packing, no upload of graph data): job j holds graphs
// ids[begin[j] .. begin[j+1]) of `src` (whose host copy gives their sizes),
// in that order, with src's dense labels.
void gather_db(const DeviceDB &src, const HostDB &src_host, int32_t n_jobs, const int32_t *begin, const int32_t *ids,
               DeviceDB &out, hipStream_t s);

// Reported pattern: job, fully labelled code in original labels, support,
// |Proj| (0 unless collected).
// Called per frequent pattern, level by level, with its code in the
// database's own labels; an empty function reports nothing (the patterns
// are still counted).
typedef std::function<bool(int32_t job, const DFS *code, int32_t n, uint32_t sup, uint64_t n_proj)> ReportFn;

// Pattern-support queries as a trie of minimum DFS codes (dense labels):
// node 0 is the empty code; a node's children extend its code by one tuple;
// `ends` lists the queries whose code ends at the node.
struct QueryTrie {
  struct Node {
    std::vector<std::pair<DFS, int32_t>> kids;
    std::vector<int32_t> ends;
  };
  std::vector<Node> nodes{Node()};
  int32_t add(const std::vector<DFS> &code, int32_t query);  // returns the final node
  int32_t child(int32_t node, const DFS &t) const;           // -1 when absent
};

struct MineOptions {
  std::vector<int32_t> minsup;  // per job
  int32_t max_edges = 0;
  int32_t canon_threads = 0;
  bool collect_proj = false;
  bool path_pruning = true;     // k_level drops candidates with an infrequent 2-edge sub-path
  int32_t speculate = 0;        // 0 adaptive, 1 never, 2 always (gm_cfg::speculate)
  bool stage_events = false;    // time the extension + support stages with HIP events
  int32_t chunk_rows = 0;       // launch shape (gm_cfg::chunk_rows); 0 = default
  int32_t table_bits = 0;       // launch shape (gm_cfg::table_bits); 0 = default
  // device bytes for the levels' rows and candidate lists (gm_cfg::
  // mem_budget_bytes); a level whose children would pass it is written in
  // pieces by parent pattern, each piece's subtree mined before the next
  // (0: unlimited)
  int64_t mem_budget = 0;
  // minimum-code check of a level's frequent keys: 0 on the device for the
  // levels whose children the host would check before writing them (no
  // speculation), else on the host pool overlapping the write; 1 always on
  // the host; 2 always on the device (k_canon_level)
  int32_t canon_mode = 0;
  // Query mode: instead of mining, follow the codes of `queries` (every
  // prefix of a minimum code is one) and store each query's support in
  // query_support; the report callback is not called.
  const QueryTrie *queries = nullptr;
  std::vector<uint32_t> *query_support = nullptr;
};

// Runs the level-synchronous miner; returns false when the report callback
// asked to stop.  `ws`: the workspace to use instead of the database's own
// (the arena-backed one of the entry point that allocates nothing).
bool mine(DeviceDB &db, const MineOptions &opt, const ReportFn &report, gm_stats *stats, hipStream_t s,
          Workspace *ws = nullptr);

}  // namespace gm
