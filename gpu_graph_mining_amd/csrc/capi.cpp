// capi.cpp -- the extern "C" boundary (include/gspan_amd.h).
//
// Every entry point converts C++ exceptions into GM_E_* codes and keeps the
// message in a thread-local buffer for gm_last_error(); nothing throws
// across the boundary (the reference throws std::runtime_error and aborts on
// launch failures, src/cuda_pp/cuda_tools.hpp:10-18, kernel_execution.hpp:63-68).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>

#include "gspan_amd.h"
#include "gm_internal.hpp"
#include "miner.hpp"
#include "kernels.hpp"

struct gm_db {
  gm::HostDB host;
  std::unique_ptr<gm::DeviceDB> dev;
  hipStream_t stream = nullptr;
  ~gm_db() {
    dev.reset();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct gm_batch {
  gm::DeviceDB dev;
  hipStream_t stream = nullptr;
  ~gm_batch() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

thread_local std::string g_last_error;

template <class F>
int guard(F &&f) {
  try {
    g_last_error.clear();
    f();
    return GM_OK;
  } catch (const gm::Error &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "host allocation failed";
    return GM_E_OOM;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return GM_E_ARG;
  } catch (...) {
    g_last_error = "unknown error";
    return GM_E_ARG;
  }
}

void ensure_stream(gm_db *db) {
  if (!db->stream) gm::hip_check(hipStreamCreateWithFlags(&db->stream, hipStreamNonBlocking), "hipStreamCreate");
}

void ensure_device(gm_db *db) {
  if (!db->dev) {
    std::unique_ptr<gm::DeviceDB> d(new gm::DeviceDB());
    std::vector<const gm::HostDB *> one{&db->host};
    gm::PackedDB pk;
    gm::pack_db(one, *d, pk);  // validates before any HIP call
    ensure_stream(db);
    gm::upload_db(pk, *d, db->stream);
    db->dev = std::move(d);
  }
}

gm_db *new_db(gm::HostDB &&h) {
  gm_db *db = new gm_db();
  db->host = std::move(h);
  return db;
}

void flatten(const gm::DFS *code, int32_t n, int style, std::vector<int32_t> &out) {
  out.resize((size_t)n * 5);
  for (int32_t i = 0; i < n; ++i) {
    const gm::DFS &d = code[i];
    out[5 * i + 0] = d.from;
    out[5 * i + 1] = d.to;
    out[5 * i + 2] = (style == GM_LABELS_WILDCARD && i > 0) ? -1 : d.fromlabel;
    out[5 * i + 3] = d.elabel;
    out[5 * i + 4] = (style == GM_LABELS_WILDCARD && d.from > d.to) ? -1 : d.tolabel;
  }
}

gm::MineOptions options(const gm_cfg *cfg) {
  gm::MineOptions o;
  if (cfg) {
    o.max_edges = cfg->max_edges;
    o.canon_threads = cfg->canon_threads;
    o.collect_proj = cfg->collect_proj != 0;
    o.path_pruning = cfg->path_pruning == 0;
    if (cfg->speculate < 0 || cfg->speculate > 2) throw gm::Error(GM_E_ARG, "speculate must be 0, 1 or 2");
    o.speculate = cfg->speculate;
    o.stage_events = cfg->stage_events != 0;
    if (cfg->chunk_rows != 0 &&
        (cfg->chunk_rows < 64 || cfg->chunk_rows > 4096 || (cfg->chunk_rows & (cfg->chunk_rows - 1))))
      throw gm::Error(GM_E_ARG, "chunk_rows must be 0 or a power of two in [64, 4096]");
    if (cfg->table_bits != 0 && (cfg->table_bits < 7 || cfg->table_bits > 11))
      throw gm::Error(GM_E_ARG, "table_bits must be 0 or in [7, 11]");
    // (k_write takes a chunk's records in one table of 2^8 entries; larger
    // k_level tables let chunks leave more, which k_write then counts in
    // split passes)
    if (cfg->mem_budget_bytes < 0) throw gm::Error(GM_E_ARG, "mem_budget_bytes must be >= 0");
    o.mem_budget = cfg->mem_budget_bytes;
    o.chunk_rows = cfg->chunk_rows;
    o.table_bits = cfg->table_bits;
    if (cfg->canon_host < 0 || cfg->canon_host > 2) throw gm::Error(GM_E_ARG, "canon_host must be 0, 1 or 2");
    o.canon_mode = cfg->canon_host;
    if (cfg->block_threads != 0 && cfg->block_threads != 64 && cfg->block_threads != 128 &&
        cfg->block_threads != 256)
      throw gm::Error(GM_E_ARG, "block_threads must be 0, 64, 128 or 256");
  }
  return o;
}

}  // namespace

extern "C" {

const char *gm_version(void) { return "gspan_amd 0.1.0 (gfx950)"; }

const char *gm_last_error(void) { return g_last_error.c_str(); }

void gm_cfg_init(gm_cfg *cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->label_style = GM_LABELS_FULL;
}

int gm_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int gm_db_create(int32_t n_graphs, const int32_t *graph_vbegin, const int32_t *vlabels, const int32_t *graph_ebegin,
                 const int32_t *efrom, const int32_t *eto, const int32_t *elabel, gm_db **out) {
  return guard([&]() {
    if (!out || n_graphs < 0 || !graph_vbegin || !graph_ebegin) throw gm::Error(GM_E_ARG, "null argument");
    *out = nullptr;
    gm::HostDB h;
    h.n_graphs = n_graphs;
    h.vbegin.assign(graph_vbegin, graph_vbegin + n_graphs + 1);
    h.ebegin.assign(graph_ebegin, graph_ebegin + n_graphs + 1);
    const int64_t nv = h.vbegin.back(), ne = h.ebegin.back();
    if (nv < 0 || ne < 0 || (nv > 0 && !vlabels) || (ne > 0 && (!efrom || !eto || !elabel)))
      throw gm::Error(GM_E_ARG, "null array");
    h.vlabel.assign(vlabels, vlabels + nv);
    h.efrom.assign(efrom, efrom + ne);
    h.eto.assign(eto, eto + ne);
    h.elabel.assign(elabel, elabel + ne);
    h.check_consistent();
    *out = new_db(std::move(h));
  });
}

int gm_db_load(const char *path, const char *filetype, gm_db **out) {
  return guard([&]() {
    if (!path || !filetype || !out) throw gm::Error(GM_E_ARG, "null argument");
    *out = nullptr;
    std::string ft(filetype);
    if (!ft.empty() && ft[0] == '-') ft = ft.substr(1);
    gm::HostDB h;
    if (ft == "txt")
      gm::read_txt(path, false, h);
    else if (ft == "fsg")
      gm::read_txt(path, true, h);
    else if (ft == "bin")
      gm::read_bin(path, h);
    else if (ft == "cudabin")
      gm::read_cudabin(path, h);
    else
      throw gm::Error(GM_E_ARG, "unknown filetype");
    h.check_consistent();
    *out = new_db(std::move(h));
  });
}

int gm_db_write_txt(const gm_db *db, const char *path) {
  return guard([&]() {
    if (!db || !path) throw gm::Error(GM_E_ARG, "null argument");
    gm::write_txt(db->host, path);
  });
}

int gm_db_write(const gm_db *db, const char *path, const char *filetype) {
  return guard([&]() {
    if (!db || !path || !filetype) throw gm::Error(GM_E_ARG, "null argument");
    std::string ft(filetype);
    if (!ft.empty() && ft[0] == '-') ft = ft.substr(1);
    if (ft == "txt")
      gm::write_txt(db->host, path);
    else if (ft == "bin")
      gm::write_bin(db->host, path);
    else if (ft == "cudabin")
      gm::write_cudabin(db->host, path);
    else
      throw gm::Error(GM_E_ARG, "unknown filetype");
  });
}

int gm_db_load_part(const char *path, int32_t total_parts, int32_t part, gm_db **out) {
  return guard([&]() {
    if (!path || !out) throw gm::Error(GM_E_ARG, "null argument");
    *out = nullptr;
    gm::HostDB h;
    gm::read_bin_part(path, total_parts, part, h);
    h.check_consistent();
    *out = new_db(std::move(h));
  });
}

int gm_db_counts(const gm_db *db, int64_t counts[6]) {
  return guard([&]() {
    if (!db || !counts) throw gm::Error(GM_E_ARG, "null argument");
    const gm::HostDB &h = db->host;
    std::vector<int32_t> v(h.vlabel), e(h.elabel);
    std::sort(v.begin(), v.end());
    std::sort(e.begin(), e.end());
    counts[0] = h.n_graphs;
    counts[1] = h.n_vertices();
    counts[2] = h.n_edges();
    counts[3] = h.max_vertices();
    counts[4] = std::unique(v.begin(), v.end()) - v.begin();
    counts[5] = std::unique(e.begin(), e.end()) - e.begin();
  });
}

int gm_db_export(const gm_db *db, int32_t *graph_vbegin, int32_t *vlabels, int32_t *graph_ebegin, int32_t *efrom,
                 int32_t *eto, int32_t *elabel) {
  return guard([&]() {
    if (!db) throw gm::Error(GM_E_ARG, "null argument");
    const gm::HostDB &h = db->host;
    if (graph_vbegin) std::copy(h.vbegin.begin(), h.vbegin.end(), graph_vbegin);
    if (vlabels) std::copy(h.vlabel.begin(), h.vlabel.end(), vlabels);
    if (graph_ebegin) std::copy(h.ebegin.begin(), h.ebegin.end(), graph_ebegin);
    if (efrom) std::copy(h.efrom.begin(), h.efrom.end(), efrom);
    if (eto) std::copy(h.eto.begin(), h.eto.end(), eto);
    if (elabel) std::copy(h.elabel.begin(), h.elabel.end(), elabel);
  });
}

int gm_db_upload(gm_db *db) {
  return guard([&]() {
    if (!db) throw gm::Error(GM_E_ARG, "null argument");
    ensure_device(db);
  });
}

void gm_db_destroy(gm_db *db) { delete db; }

int gm_db_subset(const gm_db *db, const int32_t *ids, int32_t n, gm_db **out) {
  return guard([&]() {
    if (!db || !out || n < 0 || (n > 0 && !ids)) throw gm::Error(GM_E_ARG, "null argument");
    *out = nullptr;
    const gm::HostDB &h = db->host;
    gm::HostDB r;
    std::vector<int32_t> lab, ef, et, el;
    for (int32_t i = 0; i < n; ++i) {
      int32_t g = ids[i];
      if (g < 0 || g >= h.n_graphs) throw gm::Error(GM_E_ARG, "graph id out of range");
      lab.assign(h.vlabel.begin() + h.vbegin[g], h.vlabel.begin() + h.vbegin[g + 1]);
      ef.assign(h.efrom.begin() + h.ebegin[g], h.efrom.begin() + h.ebegin[g + 1]);
      et.assign(h.eto.begin() + h.ebegin[g], h.eto.begin() + h.ebegin[g + 1]);
      el.assign(h.elabel.begin() + h.ebegin[g], h.elabel.begin() + h.ebegin[g + 1]);
      r.add_graph(lab, ef, et, el);
    }
    *out = new_db(std::move(r));
  });
}

int gm_mine(gm_db *db, int32_t abs_minsup, const gm_cfg *cfg, gm_pattern_cb cb, void *user, gm_stats *stats) {
  return guard([&]() {
    if (!db) throw gm::Error(GM_E_ARG, "null database");
    if (abs_minsup < 1) throw gm::Error(GM_E_ARG, "absolute min support < 1");
    ensure_device(db);
    gm::MineOptions o = options(cfg);
    o.minsup = {abs_minsup};
    const int style = cfg ? cfg->label_style : GM_LABELS_FULL;
    std::vector<int32_t> flat;
    gm::ReportFn rep;
    if (cb)
      rep = [&](int32_t, const gm::DFS *code, int32_t n, uint32_t sup, uint64_t nproj) {
        flatten(code, n, style, flat);
        return cb(user, flat.data(), n, sup, nproj) == 0;
      };
    if (!gm::mine(*db->dev, o, rep, stats, db->stream)) throw gm::Error(GM_E_CALLBACK, "callback stopped mining");
  });
}

int gm_mine_noalloc(gm_db *db, int32_t abs_minsup, const gm_cfg *cfg, gm_pattern_cb cb, void *user,
                    gm_stats *stats) {
  return guard([&]() {
    if (!db) throw gm::Error(GM_E_ARG, "null database");
    if (abs_minsup < 1) throw gm::Error(GM_E_ARG, "absolute min support < 1");
    if (!db->dev) throw gm::Error(GM_E_ARG, "gm_mine_noalloc needs a database already on the device (gm_db_upload)");
    void *base = nullptr;
    size_t bytes = 0;
    gm::hip_check(gm::static_scratch(&base, &bytes), "static workspace");
    gm::MineOptions o = options(cfg);
    o.minsup = {abs_minsup};
    o.stage_events = false;  // events are allocations
    const int style = cfg ? cfg->label_style : GM_LABELS_FULL;
    std::vector<int32_t> flat;
    gm::ReportFn rep;
    if (cb)
      rep = [&](int32_t, const gm::DFS *code, int32_t n, uint32_t sup, uint64_t nproj) {
        flatten(code, n, style, flat);
        return cb(user, flat.data(), n, sup, nproj) == 0;
      };
    // one call at a time per device: every call shares the device's array
    static std::mutex mu[64];
    int dev = 0;
    gm::hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(mu[dev & 63]);
    gm::Arena arena;
    arena.base = static_cast<char *>(base);
    arena.cap = bytes;
    arena.hi = bytes;
    gm::Workspace ws;
    ws.use_arena(&arena);
    // the call's permanent buffer (the min supports) at the bottom of side 0
    ws.minsup.arena = &arena;
    ws.minsup.side = 0;
    ws.minsup.ensure(sizeof(int32_t));
    arena.lo0 = arena.lo;
    if (!gm::mine(*db->dev, o, rep, stats, db->stream, &ws)) throw gm::Error(GM_E_CALLBACK, "callback stopped mining");
  });
}

int gm_support_queries(gm_db *db, int32_t n_queries, const int32_t *dfs5, const int32_t *n_edges, const gm_cfg *cfg,
                       uint32_t *supports, gm_stats *stats) {
  return guard([&]() {
    if (!db || n_queries < 0 || (n_queries > 0 && (!dfs5 || !n_edges || !supports)))
      throw gm::Error(GM_E_ARG, "null argument");
    // minimum codes first: malformed queries fail before any device work
    std::vector<std::vector<gm::DFS>> mins(n_queries);
    std::vector<gm::DFS> code;
    const int32_t *row = dfs5;
    for (int32_t q = 0; q < n_queries; ++q) {
      supports[q] = 0;
      const int32_t n = n_edges[q];
      if (n < 1) throw gm::Error(GM_E_ARG, "query " + std::to_string(q) + " has no edges");
      code.resize(n);
      for (int32_t i = 0; i < n; ++i, row += 5) code[i] = gm::DFS{row[0], row[1], row[2], row[3], row[4]};
      gm::min_dfs_code(code.data(), n, mins[q]);
      int32_t nv = 0;
      for (const gm::DFS &t : mins[q]) nv = std::max(nv, std::max(t.from, t.to) + 1);
      if (nv > gm::kMaxPatV) throw gm::Error(GM_E_UNSUPPORTED, "query pattern with more than 64 vertices");
    }
    ensure_device(db);
    const gm::DeviceDB &d = *db->dev;
    auto dense = [](const std::vector<int32_t> &values, int32_t v) {
      auto it = std::lower_bound(values.begin(), values.end(), v);
      return (it == values.end() || *it != v) ? -1 : (int32_t)(it - values.begin());
    };
    gm::QueryTrie trie;
    for (int32_t q = 0; q < n_queries; ++q) {
      bool present = true;  // every label occurs in the database
      for (gm::DFS &t : mins[q]) {
        t.fromlabel = dense(d.vl_values, t.fromlabel);
        t.tolabel = dense(d.vl_values, t.tolabel);
        t.elabel = dense(d.el_values, t.elabel);
        if (t.fromlabel < 0 || t.tolabel < 0 || t.elabel < 0) present = false;
      }
      if (present) trie.add(mins[q], q);
    }
    if (n_queries == 0 || trie.nodes.size() == 1) {
      if (stats) std::memset(stats, 0, sizeof(*stats));
      return;
    }
    if (d.n_jobs != 1) throw gm::Error(GM_E_ARG, "queries run against a single database");
    gm::MineOptions o = options(cfg);
    o.minsup = {1};
    o.queries = &trie;
    std::vector<uint32_t> res(n_queries, 0);
    o.query_support = &res;
    gm::mine(*db->dev, o, gm::ReportFn(), stats, db->stream);
    std::memcpy(supports, res.data(), sizeof(uint32_t) * n_queries);
  });
}

int gm_support_query(gm_db *db, const int32_t *dfs5, int32_t n_edges, uint32_t *support) {
  return gm_support_queries(db, 1, dfs5, &n_edges, nullptr, support, nullptr);
}

int gm_batch_create(int32_t n_jobs, gm_db *const *dbs, gm_batch **out) {
  return guard([&]() {
    if (n_jobs < 1 || !dbs || !out) throw gm::Error(GM_E_ARG, "null argument");
    *out = nullptr;
    std::vector<const gm::HostDB *> hs;
    for (int32_t j = 0; j < n_jobs; ++j) {
      if (!dbs[j]) throw gm::Error(GM_E_ARG, "null database");
      hs.push_back(&dbs[j]->host);
    }
    std::unique_ptr<gm_batch> b(new gm_batch());
    gm::PackedDB pk;
    gm::pack_db(hs, b->dev, pk);
    gm::hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
    gm::upload_db(pk, b->dev, b->stream);
    *out = b.release();
  });
}

int gm_batch_create_subsets(gm_db *db, int32_t n_jobs, const int32_t *job_begin, const int32_t *graph_ids,
                            gm_batch **out) {
  return guard([&]() {
    if (!db || n_jobs < 1 || !job_begin || !out || (job_begin[n_jobs] > 0 && !graph_ids))
      throw gm::Error(GM_E_ARG, "null argument");
    *out = nullptr;
    ensure_device(db);
    std::unique_ptr<gm_batch> b(new gm_batch());
    gm::hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
    // the source's upload is complete (ensure_device synchronises its stream)
    gm::gather_db(*db->dev, db->host, n_jobs, job_begin, graph_ids, b->dev, b->stream);
    *out = b.release();
  });
}

int gm_batch_mine(gm_batch *batch, const int32_t *abs_minsup, const gm_cfg *cfg, gm_job_pattern_cb cb, void *user,
                  gm_stats *stats) {
  return guard([&]() {
    if (!batch || !abs_minsup) throw gm::Error(GM_E_ARG, "null argument");
    const int32_t n_jobs = batch->dev.n_jobs;
    for (int32_t j = 0; j < n_jobs; ++j)
      if (abs_minsup[j] < 1) throw gm::Error(GM_E_ARG, "absolute min support < 1");
    gm::MineOptions o = options(cfg);
    o.minsup.assign(abs_minsup, abs_minsup + n_jobs);
    const int style = cfg ? cfg->label_style : GM_LABELS_FULL;
    std::vector<int32_t> flat;
    gm::ReportFn rep;
    if (cb)
      rep = [&](int32_t job, const gm::DFS *code, int32_t n, uint32_t sup, uint64_t nproj) {
        flatten(code, n, style, flat);
        return cb(user, job, flat.data(), n, sup, nproj) == 0;
      };
    if (!gm::mine(batch->dev, o, rep, stats, batch->stream))
      throw gm::Error(GM_E_CALLBACK, "callback stopped mining");
  });
}

void gm_batch_destroy(gm_batch *batch) { delete batch; }

int gm_mine_batch(int32_t n_jobs, gm_db *const *dbs, const int32_t *abs_minsup, const gm_cfg *cfg,
                  gm_job_pattern_cb cb, void *user, gm_stats *stats) {
  gm_batch *b = nullptr;
  int rc = gm_batch_create(n_jobs, dbs, &b);
  if (rc != GM_OK) return rc;
  rc = gm_batch_mine(b, abs_minsup, cfg, cb, user, stats);
  gm_batch_destroy(b);
  return rc;
}

namespace {

// A code as the check takes it: every tuple carries the labels its vertices
// get anywhere in the code (a wildcard code completed from its labelled
// tuples).
void complete_code(const int32_t *dfs5, int32_t n_edges, gm::DFS *c) {
  std::vector<int32_t> vl(n_edges + 2, -1);
  for (int32_t i = 0; i < n_edges; ++i) {
    const int32_t *d = dfs5 + 5 * i;
    c[i] = gm::DFS{d[0], d[1], d[2], d[3], d[4]};
    int32_t m = std::max(d[0], d[1]);
    if (d[0] < 0 || d[1] < 0 || m >= (int32_t)vl.size()) throw gm::Error(GM_E_ARG, "malformed code");
    if (d[2] != -1) vl[d[0]] = d[2];
    if (d[4] != -1) vl[d[1]] = d[4];
  }
  for (int32_t i = 0; i < n_edges; ++i) {
    c[i].fromlabel = vl[c[i].from];
    c[i].tolabel = vl[c[i].to];
  }
}

}  // namespace

int gm_dfs_code_is_min(const int32_t *dfs5, int32_t n_edges) {
  int result = 0;
  int rc = guard([&]() {
    if (!dfs5 || n_edges < 1) throw gm::Error(GM_E_ARG, "empty code");
    std::vector<gm::DFS> c(n_edges);
    complete_code(dfs5, n_edges, c.data());
    result = gm::dfs_code_is_min(c.data(), n_edges) ? 1 : 0;
  });
  return rc == GM_OK ? result : rc;
}

int gm_dfs_codes_are_min(int32_t n_codes, int32_t n_edges, const int32_t *dfs5, uint8_t *out) {
  return guard([&]() {
    if (n_codes < 0 || n_edges < 1 || (n_codes > 0 && (!dfs5 || !out))) throw gm::Error(GM_E_ARG, "bad arguments");
    if (n_codes == 0) return;
    static_assert(sizeof(gm::DFS) == 5 * sizeof(int32_t), "DFS layout");
    const size_t tuples = (size_t)n_codes * n_edges;
    std::vector<gm::DFS> codes(tuples);
    for (int32_t q = 0; q < n_codes; ++q) complete_code(dfs5 + (size_t)5 * n_edges * q, n_edges, &codes[(size_t)q * n_edges]);
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1) throw gm::Error(GM_E_DEVICE, "no HIP device");
    hipStream_t s = nullptr;
    void *dcodes = nullptr, *dout = nullptr;
    auto release = [&]() {
      if (dcodes) (void)hipFree(dcodes);
      if (dout) (void)hipFree(dout);
      if (s) (void)hipStreamDestroy(s);
    };
    try {
      gm::hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
      gm::hip_check(hipMalloc(&dcodes, sizeof(gm::DFS) * tuples), "hipMalloc");
      gm::hip_check(hipMalloc(&dout, (size_t)n_codes), "hipMalloc");
      gm::hip_check(hipMemcpyAsync(dcodes, codes.data(), sizeof(gm::DFS) * tuples, hipMemcpyHostToDevice, s), "upload codes");
      gm::CanonArgs c;
      std::memset(&c, 0, sizeof(c));
      c.codes = static_cast<const int32_t *>(dcodes);
      c.n = n_codes;
      c.len = n_edges;
      c.out = static_cast<uint8_t *>(dout);
      gm::hip_check(gm::launch_canon_list(c, s), "k_canon_list");
      gm::hip_check(hipMemcpyAsync(out, dout, (size_t)n_codes, hipMemcpyDeviceToHost, s), "download answers");
      gm::hip_check(hipStreamSynchronize(s), "k_canon_list");
    } catch (...) {
      release();
      throw;
    }
    release();
    // codes the device left undecided are checked here
    for (int32_t q = 0; q < n_codes; ++q) {
      if (out[q] == gm::kCanonBad) throw gm::Error(GM_E_ARG, "malformed DFS code");
      if (out[q] == gm::kCanonHost) out[q] = gm::dfs_code_is_min(&codes[(size_t)q * n_edges], n_edges) ? 1 : 0;
    }
  });
}

int gm_gen_random(uint64_t seed, int32_t n_graphs, int32_t vmin, int32_t vmax, double edge_factor, int32_t n_vlabels,
                  int32_t n_elabels, gm_db **out) {
  return guard([&]() {
    if (!out) throw gm::Error(GM_E_ARG, "null argument");
    gm::HostDB h;
    gm::gen_random(seed, n_graphs, vmin, vmax, edge_factor, n_vlabels, n_elabels, h);
    *out = new_db(std::move(h));
  });
}

int gm_gen_planted(uint64_t seed, int32_t n_graphs, double mean_edges, int32_t n_seed_patterns,
                   double mean_pattern_edges, int32_t n_vlabels, int32_t n_elabels, double merge_prob, gm_db **out) {
  return guard([&]() {
    if (!out) throw gm::Error(GM_E_ARG, "null argument");
    gm::HostDB h;
    gm::gen_planted(seed, n_graphs, mean_edges, n_seed_patterns, mean_pattern_edges, n_vlabels, n_elabels, merge_prob,
                    h);
    *out = new_db(std::move(h));
  });
}

int gm_gen_molecule(uint64_t seed, int32_t n_graphs, double mean_vertices, int32_t max_degree, int32_t n_vlabels,
                    const double *vlabel_probs, int32_t n_elabels, const double *elabel_probs, gm_db **out) {
  return guard([&]() {
    if (!out || !vlabel_probs || !elabel_probs || n_vlabels < 1 || n_elabels < 1)
      throw gm::Error(GM_E_ARG, "null argument");
    gm::HostDB h;
    gm::gen_molecule(seed, n_graphs, mean_vertices, max_degree,
                     std::vector<double>(vlabel_probs, vlabel_probs + n_vlabels),
                     std::vector<double>(elabel_probs, elabel_probs + n_elabels), h);
    *out = new_db(std::move(h));
  });
}

}  // extern "C"
