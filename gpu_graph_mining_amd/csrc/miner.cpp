// miner.cpp -- device database upload and the level-synchronous driver.
//
// The reference grows the pattern tree depth-first on the host and launches
// a handful of kernels per pattern, each followed by a device sync
// (src/gspan_cuda/gspan_cuda_no_sort.cpp:275-489; kernel_execution.hpp:61-71).
// This driver grows it level by level instead: one level step extends the
// embeddings of *every* pattern of the current size in the same launches,
// so each step costs a fixed number of kernels and one host round trip
// (the frequent-key list down, the chosen children up) no matter how many
// patterns the level holds.  Patterns are independent subtrees, so the
// set of reported (code, support) pairs is the one the depth-first oracle
// reports (src/gspan/gspan.cpp:172-274); only the emission order differs.
//
// Level step for patterns P_1..P_n with embeddings in graph-major rows:
//   1. k_expand    : every row's extension candidates, listed (row-parallel);
//      k_support   : distinct-graph support and embedding count of every
//                    (pattern, extension key), one workgroup per graph;
//      k_reduce    : the keys whose support reached their job's min support;
//   2. host        : decode each frequent key into a child code, keep the
//                    minimal ones (canonical.cpp; on speculative levels every
//                    frequent key becomes a child and the check runs while
//                    the device writes), number the children in key order;
//   3. children    : count, place and write the child rows (row-parallel
//                    passes where every candidate is listed, per-graph passes
//                    otherwise), sized exactly by the embedding counts
//                    k_support gathered.
#include "miner.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

#include "kernels.hpp"

namespace gm {

void hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    throw Error(GM_E_OOM, std::string(what) + ": " + hipGetErrorString(e));
  throw Error(GM_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

void *Arena::take(int side, size_t bytes) {
  if (side < 0) side = scratch_side;
  size_t at;
  if (side == 0) {
    at = (lo + 255) & ~size_t(255);
    if (at + bytes > hi) at = cap;  // meets side 1
    else lo = at + bytes;
  } else {
    at = bytes <= hi ? (hi - bytes) & ~size_t(255) : cap;
    if (at < lo || at >= cap) at = cap;
    else hi = at;
  }
  if (at >= cap)
    throw Error(GM_E_OOM, "the level steps need more than the " + std::to_string(cap >> 20) +
                              " MiB static workspace (" + std::to_string((lo + (cap - hi) + bytes) >> 20) +
                              " MiB asked)");
  peak = std::max(peak, lo + (cap - hi));
  return base + at;
}

void DevBuf::release() {
  if (p && !arena) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

void *DevBuf::ensure(size_t bytes) {
  if (bytes <= cap && p) return p;
  if (arena) {  // exact sizes (whole 256-byte blocks); the stacks are emptied level by level
    release();
    cap = (std::max<size_t>(bytes, 256) + 255) & ~size_t(255);
    p = arena->take(side, cap);
    return p;
  }
  const size_t want = (std::max<size_t>({bytes, cap + cap / 2, 256}) + 255) & ~size_t(255);  // fills take 8-byte words
  release();
  hip_check(hipMalloc(&p, want), "hipMalloc");
  cap = want;
  return p;
}

HostBuf::~HostBuf() {
  if (p && !arena) (void)hipHostFree(p);
}

void *HostBuf::ensure(size_t bytes) {
  if (bytes <= cap && p) return p;
  if (arena) {
    host.resize(std::max<size_t>(bytes, 256));
    p = host.data();
    dev = arena->take(-1, host.size());
    cap = host.size();
    return p;
  }
  const size_t want = std::max<size_t>({bytes, cap + cap / 2, 4096});
  if (p) (void)hipHostFree(p);
  p = nullptr;
  dev = nullptr;
  cap = 0;
  hip_check(hipHostMalloc(&p, want, hipHostMallocMapped | hipHostMallocCoherent), "hipHostMalloc");
  cap = want;
  hip_check(hipHostGetDevicePointer(&dev, p, 0), "hipHostGetDevicePointer");
  return p;
}

namespace {

template <class T>
void upload(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
  b.ensure(sizeof(T) * std::max<size_t>(v.size(), 1));
  if (!v.empty()) hip_check(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s), "upload");
}

int64_t round4(int64_t x) { return (x + 3) & ~int64_t(3); }

}  // namespace

int32_t QueryTrie::add(const std::vector<DFS> &code, int32_t query) {
  int32_t node = 0;
  for (const DFS &t : code) {
    int32_t next = child(node, t);
    if (next < 0) {
      next = (int32_t)nodes.size();
      nodes[node].kids.emplace_back(t, next);
      nodes.emplace_back();
    }
    node = next;
  }
  nodes[node].ends.push_back(query);
  return node;
}

int32_t QueryTrie::child(int32_t node, const DFS &t) const {
  for (const auto &k : nodes[node].kids) {
    const DFS &a = k.first;
    if (a.from == t.from && a.to == t.to && a.fromlabel == t.fromlabel && a.elabel == t.elabel &&
        a.tolabel == t.tolabel)
      return k.second;
  }
  return -1;
}

void pack_db(const std::vector<const HostDB *> &dbs, DeviceDB &out, PackedDB &pk) {
  out.n_jobs = (int32_t)dbs.size();
  std::vector<int32_t> vls, els;
  int64_t nv = 0, ne = 0;
  int64_t ng = 0;
  for (const HostDB *db : dbs) {
    db->check_consistent();
    vls.insert(vls.end(), db->vlabel.begin(), db->vlabel.end());
    els.insert(els.end(), db->elabel.begin(), db->elabel.end());
    nv += db->n_vertices();
    ne += db->n_edges();
    ng += db->n_graphs;
  }
  if (ng > INT32_MAX - 1 || nv > INT32_MAX - 1 || 2 * ne > INT32_MAX - 1)
    throw Error(GM_E_UNSUPPORTED, "database too large for 32-bit offsets");
  if (ng >= (1ll << 24)) throw Error(GM_E_UNSUPPORTED, "more than 16777215 graphs (24-bit support counters)");
  std::sort(vls.begin(), vls.end());
  vls.erase(std::unique(vls.begin(), vls.end()), vls.end());
  std::sort(els.begin(), els.end());
  els.erase(std::unique(els.begin(), els.end()), els.end());
  if (vls.size() > 255 || els.size() > 255)
    throw Error(GM_E_UNSUPPORTED, "more than 255 distinct vertex or edge labels");
  out.vl_values = vls;
  out.el_values = els;
  out.NVL = std::max<int32_t>(1, (int32_t)vls.size());
  out.NEL = std::max<int32_t>(1, (int32_t)els.size());
  auto vmap = [&](int32_t l) { return (int32_t)(std::lower_bound(vls.begin(), vls.end(), l) - vls.begin()); };
  auto emap = [&](int32_t l) { return (int32_t)(std::lower_bound(els.begin(), els.end(), l) - els.begin()); };

  std::vector<int32_t> &gvbase = pk.gvbase, &vadj = pk.vadj, &gjob = pk.gjob, &vgraph = pk.vgraph;
  std::vector<uint32_t> &adj = pk.adj;
  std::vector<uint8_t> &vlab = pk.vlab;
  pk.max_graph_v = pk.max_graph_adj = 0;
  gvbase.clear();
  vadj.clear();
  gjob.clear();
  vgraph.clear();
  adj.clear();
  vlab.clear();
  gvbase.reserve(ng + 1);
  vadj.reserve(nv + 1);
  adj.reserve(2 * ne);
  vlab.reserve(nv);
  vgraph.reserve(nv);
  gjob.reserve(ng);
  out.job_first_graph.assign(1, 0);
  out.job_first_vertex.assign(1, 0);
  std::vector<int32_t> deg, fill, dl;
  std::vector<std::pair<int32_t, int32_t>> pairs;
  int32_t job = 0;
  for (const HostDB *db : dbs) {
    for (int32_t g = 0; g < db->n_graphs; ++g) {
      const int32_t v0 = db->vbegin[g], n = db->vbegin[g + 1] - v0;
      const int32_t e0 = db->ebegin[g], e1 = db->ebegin[g + 1];
      if (n > kMaxGraphV)
        throw Error(GM_E_UNSUPPORTED, "graph " + std::to_string(g) + " has more than 256 vertices");
      pairs.clear();
      for (int32_t e = e0; e < e1; ++e) {
        int32_t a = db->efrom[e], b = db->eto[e];
        if (a == b) throw Error(GM_E_UNSUPPORTED, "self-loop in graph " + std::to_string(g));
        pairs.emplace_back(std::min(a, b), std::max(a, b));
      }
      std::sort(pairs.begin(), pairs.end());
      if (std::adjacent_find(pairs.begin(), pairs.end()) != pairs.end())
        throw Error(GM_E_UNSUPPORTED, "parallel edges in graph " + std::to_string(g));
      gvbase.push_back((int32_t)vlab.size());
      gjob.push_back(job);
      dl.resize(n);
      for (int32_t v = 0; v < n; ++v) {
        dl[v] = vmap(db->vlabel[v0 + v]);
        vlab.push_back((uint8_t)dl[v]);
        vgraph.push_back((int32_t)gjob.size() - 1);
      }
      // adjacency in input edge order, both directions, as Graph::read
      // builds it (src/globals/graph_types.cpp:94-107)
      deg.assign(n, 0);
      for (int32_t e = e0; e < e1; ++e) {
        deg[db->efrom[e]]++;
        deg[db->eto[e]]++;
      }
      const int32_t abase = (int32_t)adj.size();
      pk.max_graph_v = std::max(pk.max_graph_v, n);
      pk.max_graph_adj = std::max(pk.max_graph_adj, 2 * (e1 - e0));
      fill.assign(n, 0);
      int32_t acc = abase;
      for (int32_t v = 0; v < n; ++v) {
        vadj.push_back(acc);
        fill[v] = acc - abase;
        acc += deg[v];
      }
      adj.resize(acc);
      // entry: neighbour | elabel << 8 | neighbour label << 16 | neighbour
      // degree << 24 (a simple graph of <= 256 vertices has degree <= 255)
      for (int32_t e = e0; e < e1; ++e) {
        int32_t a = db->efrom[e], b = db->eto[e];
        uint32_t el = (uint32_t)emap(db->elabel[e]);
        adj[abase + fill[a]++] = (uint32_t)b | (el << 8) | ((uint32_t)dl[b] << 16) | ((uint32_t)deg[b] << 24);
        adj[abase + fill[b]++] = (uint32_t)a | (el << 8) | ((uint32_t)dl[a] << 16) | ((uint32_t)deg[a] << 24);
      }
    }
    ++job;
    out.job_first_graph.push_back((int32_t)gvbase.size());
    out.job_first_vertex.push_back((int64_t)vlab.size());
  }
  gvbase.push_back((int32_t)vlab.size());
  vadj.push_back((int32_t)adj.size());
  out.n_graphs = (int32_t)ng;
  out.n_vertices = nv;
  out.n_adj = (int64_t)adj.size();
  out.max_graph_v = pk.max_graph_v;
  out.max_graph_adj = pk.max_graph_adj;
}

void upload_db(const PackedDB &pk, DeviceDB &out, hipStream_t s) {
  upload(out.gvbase, pk.gvbase, s);
  upload(out.vadj, pk.vadj, s);
  upload(out.adj, pk.adj, s);
  upload(out.vlab, pk.vlab, s);
  upload(out.gjob, pk.gjob, s);
  upload(out.vgraph, pk.vgraph, s);
  hip_check(hipStreamSynchronize(s), "upload database");
  out.bytes = 4 * pk.gvbase.size() + 4 * pk.vadj.size() + 4 * pk.adj.size() + pk.vlab.size() + 4 * pk.gjob.size() +
              4 * pk.vgraph.size();
}

void gather_db(const DeviceDB &src, const HostDB &h, int32_t n_jobs, const int32_t *begin, const int32_t *ids,
               DeviceDB &out, hipStream_t s) {
  if (n_jobs < 1 || begin[0] != 0) throw Error(GM_E_ARG, "job ranges must start at 0");
  const int32_t n_sel = begin[n_jobs];
  std::vector<int32_t> sel(n_sel), dvb(n_sel), dab(n_sel), djob(n_sel);
  int64_t nv = 0, na = 0;
  int32_t max_v = 0, max_a = 0;
  out.n_jobs = n_jobs;
  out.job_first_graph.assign(1, 0);
  out.job_first_vertex.assign(1, 0);
  for (int32_t j = 0; j < n_jobs; ++j) {
    if (begin[j + 1] < begin[j]) throw Error(GM_E_ARG, "job ranges must not decrease");
    for (int32_t t = begin[j]; t < begin[j + 1]; ++t) {
      const int32_t g = ids[t];
      if (g < 0 || g >= h.n_graphs) throw Error(GM_E_ARG, "graph id out of range");
      const int32_t n = h.vbegin[g + 1] - h.vbegin[g], a = 2 * (h.ebegin[g + 1] - h.ebegin[g]);
      sel[t] = g;
      dvb[t] = (int32_t)nv;
      dab[t] = (int32_t)na;
      djob[t] = j;
      nv += n;
      na += a;
      max_v = std::max(max_v, n);
      max_a = std::max(max_a, a);
    }
    out.job_first_graph.push_back(begin[j + 1]);
    out.job_first_vertex.push_back(nv);
  }
  if (n_sel >= (1 << 24)) throw Error(GM_E_UNSUPPORTED, "more than 16777215 graphs (24-bit support counters)");
  out.n_graphs = n_sel;
  out.n_vertices = nv;
  out.n_adj = na;
  out.max_graph_v = max_v;
  out.max_graph_adj = max_a;
  out.NVL = src.NVL;
  out.NEL = src.NEL;
  out.vl_values = src.vl_values;
  out.el_values = src.el_values;
  out.gvbase.ensure(4 * (size_t)(n_sel + 1));
  out.vadj.ensure(4 * (size_t)(nv + 1));
  out.adj.ensure(4 * (size_t)std::max<int64_t>(na, 1));
  out.vlab.ensure((size_t)std::max<int64_t>(nv, 1));
  out.gjob.ensure(4 * (size_t)std::max<int32_t>(n_sel, 1));
  out.vgraph.ensure(4 * (size_t)std::max<int64_t>(nv, 1));
  out.bytes = 4 * (size_t)(n_sel + 1) + 4 * (size_t)(nv + 1) + 4 * (size_t)na + (size_t)nv + 4 * (size_t)n_sel +
              4 * (size_t)nv;
  if (n_sel == 0) {
    const int32_t zero = 0;
    hip_check(hipMemcpyAsync(out.gvbase.p, &zero, 4, hipMemcpyHostToDevice, s), "gather");
    hip_check(hipMemcpyAsync(out.vadj.p, &zero, 4, hipMemcpyHostToDevice, s), "gather");
    hip_check(hipStreamSynchronize(s), "gather");
    return;
  }
  DevBuf idx;  // sel, dvb, dab, djob
  idx.ensure(16 * (size_t)n_sel);
  int32_t *d = idx.as<int32_t>();
  hip_check(hipMemcpyAsync(d, sel.data(), 4 * (size_t)n_sel, hipMemcpyHostToDevice, s), "gather");
  hip_check(hipMemcpyAsync(d + n_sel, dvb.data(), 4 * (size_t)n_sel, hipMemcpyHostToDevice, s), "gather");
  hip_check(hipMemcpyAsync(d + 2 * (size_t)n_sel, dab.data(), 4 * (size_t)n_sel, hipMemcpyHostToDevice, s), "gather");
  hip_check(hipMemcpyAsync(d + 3 * (size_t)n_sel, djob.data(), 4 * (size_t)n_sel, hipMemcpyHostToDevice, s), "gather");
  GatherArgs g;
  g.gvbase = src.gvbase.as<int32_t>();
  g.vadj = src.vadj.as<int32_t>();
  g.adj = src.adj.as<uint32_t>();
  g.vlab = src.vlab.as<uint8_t>();
  g.sel = d;
  g.dvb = d + n_sel;
  g.dab = d + 2 * (size_t)n_sel;
  g.djob = d + 3 * (size_t)n_sel;
  g.n_sel = n_sel;
  g.out_gvbase = out.gvbase.as<int32_t>();
  g.out_vadj = out.vadj.as<int32_t>();
  g.out_gjob = out.gjob.as<int32_t>();
  g.out_vgraph = out.vgraph.as<int32_t>();
  g.out_adj = out.adj.as<uint32_t>();
  g.out_vlab = out.vlab.as<uint8_t>();
  hip_check(launch_gather_graphs(g, s), "k_gather_graphs");
  hip_check(hipStreamSynchronize(s), "k_gather_graphs");  // before the index buffer is freed
}

namespace {

// Growable host array without value-initialisation, reused across levels
// (its contents are not kept when it grows).
template <class T>
struct Arr {
  std::unique_ptr<T[]> p;
  int64_t n = 0, cap = 0;
  void resize(int64_t k) {
    if (k > cap) {
      const int64_t want = std::max<int64_t>(k, cap + cap / 2);
      p.reset(new T[want]);
      cap = want;
    }
    n = k;
  }
  T *data() { return p.get(); }
  const T *data() const { return p.get(); }
  T &operator[](int64_t i) { return p[i]; }
  const T &operator[](int64_t i) const { return p[i]; }
};

// The patterns of one level, flat.  Level-synchronous growth adds one
// tuple per level, so every code of a level has the same length.
struct Level {
  int len = 0;         // tuples per code
  int64_t n = 0;       // patterns
  int max_nv = 0;      // most vertices of a pattern
  Arr<DFS> code;       // n * len tuples, dense labels, fully labelled
  Arr<uint8_t> vlab;   // kMaxPatV dense vertex labels per pattern
  Arr<uint8_t> nv;     // vertices per pattern
  Arr<int32_t> job;
  Arr<uint32_t> sup;
  Arr<int32_t> node;   // query mode: trie node of the code
  Arr<uint8_t> alive;  // the code is minimal (a dead pattern's rows are not extended)
  Arr<int64_t> row0;   // n + 1: pattern i's rows are [row0[i], row0[i + 1]) (pattern-major rows)
  void resize(int64_t k, int l) {
    n = k;
    len = l;
    row0.resize(k + 1);
    code.resize(k * l);
    vlab.resize(k * kMaxPatV);
    nv.resize(k);
    job.resize(k);
    sup.resize(k);
    node.resize(k);
    alive.resize(k);
  }
  const DFS *c(int64_t i) const { return code.data() + i * len; }
  DFS *c(int64_t i) { return code.data() + i * len; }
  const uint8_t *vl(int64_t i) const { return vlab.data() + i * kMaxPatV; }
  uint8_t *vl(int64_t i) { return vlab.data() + i * kMaxPatV; }
};

template <int V>
PatMetaT<V> make_meta(const DFS *code, int len, const uint8_t *vlab, int nv, int32_t job) {
  PatMetaT<V> m;
  std::memset(&m, 0, sizeof(m));
  m.job = job;
  m.nv = (uint8_t)nv;
  m.minlabel = vlab[0];
  m.maxlabel = vlab[nv - 1];
  // rightmost path (src/globals/dfs_code.cpp:75-90)
  int old_from = -1, nrm = 0;
  for (int i = len - 1; i >= 0; --i) {
    const DFS &d = code[i];
    if (d.from < d.to && (nrm == 0 || old_from == d.to)) {
      m.rm[nrm++] = (uint32_t)d.from | ((uint32_t)d.elabel << 8) | ((uint32_t)vlab[d.to] << 16);
      old_from = d.from;
    }
  }
  m.nrm = (uint8_t)nrm;
  uint64_t bwd_mask = 0, bwd_from = 0;
  for (int i = 0; i < len; ++i)
    if (code[i].from == nv - 1 && code[i].to < code[i].from) bwd_mask |= 1ull << code[i].to;
  // backward targets: the from vertices of rightmost-path edges q >= 1
  for (int q = 1; q < nrm; ++q) {
    const int f = (int)(m.rm[q] & 0xFF);
    bwd_from |= 1ull << f;
    m.rmq[f] = (uint8_t)q;
  }
  m.bwd_mask = (uint32_t)bwd_mask;
  m.bwd_mask_hi = (uint32_t)(bwd_mask >> 32);
  m.bwd_from = (uint32_t)bwd_from;
  m.bwd_from_hi = (uint32_t)(bwd_from >> 32);
  return m;
}

// One frequent key as k_reduce appends it: key, support, embeddings, and
// the sum of the candidates' item counts (the children's list entries).
struct Freq {
  unsigned long long key, sup, proj, items;
};

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double sec() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

// Persistent host worker pool for the canonical checks (one per process,
// created on first use; workers sleep between level steps).
class Pool {
 public:
  static Pool &get() {
    static Pool p;
    return p;
  }
  // Run fn(i) for i in [0, n) on up to `threads` threads (the caller helps);
  // below `serial_below` items the caller runs them alone (waking the
  // workers costs tens of microseconds).  The call returns once every item
  // is done: a worker that wakes after the items are all taken is not
  // waited for.  Callers on different threads (concurrent mining calls on
  // different handles) take turns: the pool runs one function at a time.
  void run(int64_t n, int threads, const std::function<void(int64_t)> &fn, int64_t serial_below = 64) {
    if (n <= 0) return;
    if (threads <= 1 || n < serial_below) {
      for (int64_t i = 0; i < n; ++i) fn(i);
      return;
    }
    std::lock_guard<std::mutex> one_at_a_time(run_m_);
    auto job = std::make_shared<Job>();
    job->fn = &fn;
    job->n = n;
    // items per grab: enough that the threads do not queue on the shared
    // counters (two atomics per grab), small enough to balance
    job->grain = std::max<int64_t>(32, n / ((int64_t)threads * 16));
    {
      std::lock_guard<std::mutex> lk(m_);
      while ((int)workers_.size() < threads - 1) workers_.emplace_back([this]() { loop(); });
      job_ = job;
      ++gen_;
    }
    cv_.notify_all();
    job->work();
    {
      std::unique_lock<std::mutex> lk(job->m);
      job->cv.wait(lk, [&]() { return job->finished.load() == job->n; });
    }
    if (job->err) std::rethrow_exception(job->err);
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : workers_) t.join();
  }

 private:
  // One run: items are taken `grain` at a time; `finished` counts the items
  // done (or given up after an exception), so the caller waits for items,
  // not for threads.  A worker holds the job it woke for, so one that runs
  // late finds every item taken and never calls fn after the run returned.
  struct Job {
    const std::function<void(int64_t)> *fn = nullptr;
    int64_t n = 0, grain = 32;
    std::atomic<int64_t> next{0}, finished{0};
    std::mutex m, err_m;
    std::condition_variable cv;
    std::exception_ptr err;
    void done(int64_t k) {
      if (finished.fetch_add(k) + k == n) {
        std::lock_guard<std::mutex> lk(m);
        cv.notify_all();
      }
    }
    void work() {
      for (;;) {
        const int64_t i = next.fetch_add(grain);
        if (i >= n) break;
        const int64_t e = std::min(n, i + grain);
        try {
          for (int64_t j = i; j < e; ++j) (*fn)(j);
        } catch (...) {
          {
            std::lock_guard<std::mutex> lk(err_m);
            if (!err) err = std::current_exception();
          }
          const int64_t left = next.exchange(n);  // items nobody will take
          if (left < n) done(n - left);
        }
        done(e - i);
      }
    }
  };
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      std::shared_ptr<Job> job;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&]() { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        job = job_;
      }
      job->work();
    }
  }
  std::mutex run_m_, m_;
  std::condition_variable cv_;
  std::vector<std::thread> workers_;
  std::shared_ptr<Job> job_;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// Device-side counters at the head of the frequent-key buffer: frequent
// keys, failed chunks, error bits, then kSpread
// lines of spread statistics.  The level-start clear zeroes all but the
// written-row slots (k_write of the previous level adds to them; the host
// compares their running total with the rows it asked for).
struct Counters {
  uint32_t fcount, nfail;
  int32_t err, pad0;
  unsigned long long pad[6];
  unsigned long long spread[kSpread * kSpreadStride];
};
static_assert(sizeof(Counters) == 64 + 8 * kSpread * kSpreadStride, "counter block layout");

// Grow a workspace buffer; the stream is drained first because the old
// allocation may still be in use by queued kernels.
void *grow(DevBuf &b, size_t bytes, hipStream_t s) {
  if (bytes <= b.cap && b.p) return b.p;
  // (arena regions are only reused by work queued later on the stream)
  if (!b.arena) hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  return b.ensure(bytes);
}

// Grow the packed counters and clear them (their zero invariant covers all
// of the buffer); a fill kernel, not hipMemsetAsync (which graph capture
// does not replay reliably).
void grow_zero(DevBuf &b, size_t bytes, hipStream_t s) {
  if (bytes <= b.cap && b.p) return;
  grow(b, bytes, s);
  hip_check(launch_fill8(b.p, b.cap, s), "clear counters");
}

// Device -> host staging: the copy kernel into mapped memory, or (arena
// workspace) an explicit copy into the host side.
void stage_down(Workspace &w, HostBuf &hb, const void *src, size_t bytes, hipStream_t s) {
  if (w.arena)
    hip_check(hipMemcpyAsync(hb.p, src, bytes, hipMemcpyDeviceToHost, s), "download");
  else
    hip_check(launch_copy8(hb.dev, src, bytes, s), "download");
}

// Host -> device staging for kernels that read hb.dev: nothing to do for
// mapped memory; the arena workspace copies the bytes [off, off + bytes).
void stage_up(Workspace &w, HostBuf &hb, size_t off, size_t bytes, hipStream_t s) {
  if (w.arena && bytes) {
    hip_check(hipMemcpyAsync(static_cast<char *>(hb.dev) + off, static_cast<const char *>(hb.p) + off, bytes,
                             hipMemcpyHostToDevice, s),
              "upload");
    hip_check(hipStreamSynchronize(s), "upload");  // the host side is rewritten next level
  }
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Division by a run-time constant d with one 64 x 64 -> 128-bit multiply by
// M = ceil(2^64 / d) (Lemire, Kaser and Kurz, "Faster remainder by direct
// computation", 2019), exact for every dividend below 2^32; larger ones
// divide the ordinary way.  The host decodes every frequent key (parent
// pattern, source vertex, labels) with a handful of these per level.
struct FastDiv {
  uint64_t m = 0;
  uint64_t d = 1;
  FastDiv() = default;
  explicit FastDiv(uint64_t d_) : m(d_ > 1 && d_ < (1ull << 32) ? ~0ull / d_ + 1 : 0), d(d_) {}
  uint64_t div(uint64_t a) const {
    if (d == 1) return a;
    if (m == 0 || a >= (1ull << 32)) return a / d;
    return (uint64_t)(((unsigned __int128)m * a) >> 64);
  }
  uint64_t mod(uint64_t a) const { return a - div(a) * d; }
};

// children's uploads from this size on go by DMA on the workspace's side stream
constexpr size_t kSideUploadBytes = 1 << 20;

}  // namespace

bool mine(DeviceDB &db, const MineOptions &opt, const ReportFn &report, gm_stats *stats, hipStream_t s,
          Workspace *ws) {
  std::lock_guard<std::mutex> lock(db.mu);
  Timer total;
  const int32_t NEL = db.NEL, NVL = db.NVL;
  int threads = opt.canon_threads > 0 ? opt.canon_threads
                                      : (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
  gm_stats st;
  std::memset(&st, 0, sizeof(st));
  if (!ws && !db.ws) db.ws.reset(new Workspace());
  Workspace &w = ws ? *ws : *db.ws;
  if (!w.acc_zero) {
    if (w.acc.p) hip_check(launch_fill8(w.acc.p, w.acc.cap, s), "clear counters");
    if (w.accb.p) hip_check(launch_fill8(w.accb.p, w.accb.cap, s), "clear counters");
  }
  w.acc_zero = false;  // until this call ends normally

  grow(w.minsup, sizeof(int32_t) * opt.minsup.size(), s);
  hip_check(hipMemcpyAsync(w.minsup.p, opt.minsup.data(), sizeof(int32_t) * opt.minsup.size(),
                           hipMemcpyHostToDevice, s), "upload min supports");

  double ext_ms = 0, dev_sec = 0, canon_sec = 0, ev_stage_ms = 0;
  uint64_t ext_launches = 0;
  size_t peak = db.bytes;

  StepArgs a;
  std::memset(&a, 0, sizeof(a));
  a.gvbase = db.gvbase.as<int32_t>();
  a.vadj = db.vadj.as<int32_t>();
  a.adj = db.adj.as<uint32_t>();
  a.vlab = db.vlab.as<uint8_t>();
  a.gjob = db.gjob.as<int32_t>();
  a.vgraph = db.vgraph.as<int32_t>();
  a.NEL = NEL;
  a.NVL = NVL;
  int dev = 0, wall_khz = 1;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, dev), "hipDeviceGetAttribute");
  if (wall_khz <= 0) wall_khz = 100000;

  Level N;                // patterns of the next level
  Arr<Freq> frall, fr;
  Arr<int64_t> kidx;      // per frequent key: child index, -1 if dropped
  Arr<DFS> ktup;          // per frequent key: the tuple it adds
  Arr<uint8_t> knv;       // per frequent key: the child's vertices
  Arr<int32_t> knode;     // per frequent key (query mode): trie node
  Arr<uint8_t> kalive;    // per frequent key: the child's code is minimal
  Arr<uint64_t> nproj;    // per child: embeddings reported
  bool stopped = false;
  unsigned long long written_expect = 0;  // child rows asked of every k_write so far
  uint32_t last_fcount = 0;
  int table_bits = opt.table_bits > 0 ? opt.table_bits : 8;
  const int max_chunk_rows = opt.chunk_rows > 0 ? opt.chunk_rows : kChunkRows;
  const Counters *dctr_last = nullptr;
  // Path pruning: t2[((job * NVL) + centre label) * AR + arm] is the mask of
  // the arms (elabel * NVL + neighbour label) that form a frequent 2-edge
  // path with `arm` around a vertex of that label -- the job's frequent
  // 2-edge patterns, known once the level of 2-edge children is selected.
  // Every connected sub-pattern of a frequent pattern is frequent, so an
  // extension whose new edge and an existing edge at either end form an
  // infrequent path is dropped before it is listed (same results, fewer
  // candidates).  Off in query mode (its 2-edge level follows the queries
  // only) and when NEL * NVL > 64 (the masks are 64-bit).
  const int AR = NEL * NVL;
  const bool prune = opt.path_pruning && !opt.queries && AR <= 64;
  std::vector<uint64_t> t2;
  bool t2_ready = false;
  bool t2_partial = false;  // the 2-edge level was written in pieces: no path table
  const bool want_codes = opt.canon_mode != 1 && !opt.queries;
  double dead_share = 0;  // the last level's rows of non-minimal children / its child rows
  // GM_LEVEL_TRACE=1: one line per level step on stderr (sizes, timings)
  const char *trace_env = getenv("GM_LEVEL_TRACE");
  const bool trace_levels = trace_env && trace_env[0] == '1';
  // root "patterns" are the jobs: their rows are the vertices of their graphs
  const std::vector<int64_t> &job_row0 = db.job_first_vertex;

  // query mode: record the supports of the queries that end at `node`; the
  // child is kept (its rows written) only when longer queries continue
  auto query_reached = [&](int32_t node, uint32_t sup) {
    const QueryTrie::Node &nd = opt.queries->nodes[node];
    for (int32_t q : nd.ends) (*opt.query_support)[q] = sup;  // each query ends at one node
    return !nd.kids.empty();
  };
  std::vector<DFS> orig;  // a reported code in the database's own labels
  bool counters_fresh = true;  // the counter block must be zeroed (first level, or it moved)

  // The levels being mined, the current one last.  A level is normally
  // expanded whole: its children become the next level and it is dropped.
  // Under a memory budget, a level whose children would not fit is expanded
  // in pieces by parent pattern (or by key within one pattern), and each
  // piece's children are mined to the end before the next piece: the stack
  // then holds one level per depth, each in its own row set.
  struct Frame {
    Level L;                 // patterns (the root's "patterns" are the jobs)
    int level = 0;           // 0 = root stage
    int buf = 0;             // row set of its rows and pattern upload
    int S = 0;               // tuple stride
    int64_t n_rows = 0;
    uint64_t list_cap = 1;   // list entries of its rows
    size_t mask_off = 0;     // path masks in its pattern upload (stride 0: none)
    int mask_stride = 0;
    size_t code_off = 0, vlab_off = 0;  // packed codes and vertex labels for the device check
    bool codes_ready = false;
    int64_t p_next = 0;      // first pattern whose children are not all written
    uint64_t key_next = 0;   // first key (relative to p_next) not written: a pattern split over pieces
    int64_t p_counted = 0;   // patterns whose candidates are in e_cand
    int64_t p_end_next = 0;  // > 0: the next piece ends there (its children were measured)
  };
  std::deque<Frame> frames(1);
  frames.back().list_cap = (uint64_t)std::max<int64_t>(db.n_adj, 1);  // root lists: 2E
  const int64_t budget = w.arena ? 0 : opt.mem_budget;
  double piece_bytes_per_row = 0;  // child bytes per parent row of the last piece (estimate)
  // device bytes a level's rows take (rows with their counts) and a list entry
  auto rows_bytes = [](int64_t rows, int S_) { return (int64_t)rows * (22 + S_); };

  for (;;) {
    if (frames.empty()) break;
    Frame &Fr = frames.back();
    const int level = Fr.level;
    const bool root = level == 0;
    Level &L = Fr.L;
    if (Fr.p_next >= (root ? db.n_jobs : L.n)) {
      frames.pop_back();
      continue;
    }
    const int cur = Fr.buf;
    const int S = Fr.S;
    const int64_t n_rows = Fr.n_rows;
    const uint64_t list_cap = Fr.list_cap;
    const size_t mask_off = Fr.mask_off, code_off = Fr.code_off, vlab_off = Fr.vlab_off;
    const int mask_stride = Fr.mask_stride;
    const bool codes_ready = Fr.codes_ready;
    if (w.arena) {
      w.arena_level(cur);
      counters_fresh = true;  // the counter block is taken again
    }
    Timer tdev;
    const int64_t P = root ? db.n_jobs : L.n;
    const int64_t K = root ? (int64_t)NVL * NEL * NVL : (int64_t)S * NEL * (NVL + 1);
    const int64_t rows_now = root ? db.n_vertices : n_rows;
    auto prow = [&](int64_t p) { return root ? job_row0[p] : L.row0[p]; };
    // launches over pattern ranges with range * K < 2^32 (32-bit table keys)
    const int64_t prange = std::max<int64_t>(1, (int64_t)0xFFFFFFFEll / K);
    // the piece: patterns [p_begin, p_end); under a budget, as many as the
    // last piece's child bytes per parent row say will fit
    const int64_t p_begin = Fr.p_next;
    int64_t p_end = P;
    const int64_t p_end_measured = Fr.p_end_next;
    Fr.p_end_next = 0;
    int64_t held = 0;  // rows of the levels on the stack, and this level's lists
    for (const Frame &f : frames) held += rows_bytes(f.n_rows, f.S);
    held += (int64_t)list_cap * 8;
    const int64_t avail = budget - held;
    if (Fr.key_next > 0) {
      // a pattern split by key stays a piece of its own until it is done
      // (its candidates are counted once, with its first piece)
      p_end = p_begin + 1;
    } else if (p_end_measured > 0) {
      p_end = p_end_measured;
    } else if (budget > 0 && piece_bytes_per_row > 0) {
      const double rows_fit = std::max(0.0, (double)avail / piece_bytes_per_row);
      int64_t lo = p_begin + 1, hi = P;  // the last p_end whose rows fit, at least p_begin + 1
      while (lo < hi) {
        const int64_t mid = (lo + hi + 1) / 2;
        if ((double)(prow(mid) - prow(p_begin)) <= rows_fit)
          lo = mid;
        else
          hi = mid - 1;
      }
      p_end = lo;
    }
    // the counters cover the piece's keys only: (p - p_begin) * K + key,
    // at most kPieceKeys of them (a level with more is cut into pieces)
    p_end = std::min<int64_t>(p_end, p_begin + std::max<int64_t>(1, kPieceKeys / K));
    const int64_t nkeys = (p_end - p_begin) * K;
    const FastDiv fdK((uint64_t)K), fdNEL((uint64_t)NEL), fdNVL((uint64_t)NVL);
    auto parent_of = [&](uint64_t key) { return (int64_t)fdK.div(key) + p_begin; };
    const int64_t prow_lo = prow(p_begin), prow_hi = prow(p_end);
    // chunk length: kChunkRows, or shorter so that the piece has about
    // kLevelWaves chunks (a power of two, at least one block)
    int chunk_rows = max_chunk_rows;
    while (chunk_rows > kMinChunkRows && (prow_hi - prow_lo) / chunk_rows < kLevelWaves) chunk_rows /= 2;
    int64_t total_chunks = 0;
    for (int64_t lo = p_begin; lo < p_end; lo += prange)
      total_chunks += ceil_div(prow(std::min<int64_t>(p_end, lo + prange)) - prow(lo), chunk_rows);
    // counter replicas: chunk i counts into replica i mod rep, so that the
    // keys every chunk hits (the root's) are not one hot address (atomics
    // on one address serialise at ~100 ns each); up to 64 MB of them
    // (one replica in the static workspace of the entry point that
    // allocates nothing)
    int rep = 1;
    while (!w.arena && rep < 64 && (int64_t)rep * 2 * nkeys * 16 <= (64ll << 20) && (int64_t)rep * 2 * 32 <= total_chunks)
      rep *= 2;
    grow(w.sup, sizeof(int32_t) * std::max<int64_t>(nkeys, 1), s);
    grow(w.tiles, sizeof(uint32_t) * std::max<int64_t>(reduce_tiles(nkeys), 1), s);
    grow_zero(w.acc, sizeof(unsigned long long) * std::max<int64_t>(nkeys * rep, 1), s);
    grow_zero(w.accb, sizeof(unsigned long long) * std::max<int64_t>(nkeys * rep, 1), s);
    grow(w.chunks, sizeof(ChunkInfo) * std::max<int64_t>(total_chunks, 1), s);
    grow(w.fail, sizeof(uint32_t) * std::max<int64_t>(total_chunks, 1), s);
    grow(w.chunk_keys, sizeof(uint4) * level_chunk_keys(table_bits) * std::max<int64_t>(total_chunks, 1), s);
    grow(w.rccnt, sizeof(uint16_t) * std::max<int64_t>(rows_now, 1), s);
    grow(w.cand, sizeof(uint2) * std::max<uint64_t>(list_cap, 1), s);
    uint32_t fcap = (uint32_t)std::min<int64_t>(std::max<int64_t>(nkeys, 1), std::max<int64_t>(1 << 16, (p_end - p_begin) * 64));
    // the children would pass max_edges: none is kept
    const bool past_max = opt.max_edges > 0 && (root ? 1 : L.len + 1) > opt.max_edges;
    // the minimum-code check of this level's frequent keys runs on the
    // device when asked to, or when the host would check them before the
    // write (no speculation: the check is then on the critical path)
    const bool spec_wanted = opt.speculate == 2 || (opt.speculate == 0 && dead_share <= 0.3);
    const bool canon_dev = !root && want_codes && codes_ready && !past_max &&
                           (opt.canon_mode == 2 || (opt.canon_mode == 0 && !spec_wanted));
    Counters *dctr = nullptr;
    Counters cnt;
    const char *hbuf = nullptr;
    double ms = 0;
    unsigned long long ecand = 0, max_fill = 0, pruned = 0, failed = 0;
    uint32_t guess = 0;  // frequent keys in the first download
    // the stage; rerun with a longer frequent-key list when it overflowed
    for (int attempt = 0;; ++attempt) {
      const size_t freq_bytes = sizeof(Counters) + sizeof(Freq) * fcap;
      {
        // a new block (the allocator may hand back the address it freed)
        const size_t cap0 = w.freq.cap;
        if (grow(w.freq, freq_bytes, s) != (const void *)dctr_last || w.freq.cap != cap0) counters_fresh = true;
      }
      dctr = w.freq.as<Counters>();
      if (counters_fresh) {
        // written rows of an earlier level live in the old block: checked already
        hip_check(launch_fill8(dctr, sizeof(Counters), s), "clear counters");
        written_expect = 0;
        counters_fresh = false;
      }
      dctr_last = dctr;
      a.rccnt = w.rccnt.as<uint16_t>();
      a.cand = w.cand.as<uint2>();
      a.chunks = w.chunks.as<ChunkInfo>();
      a.fail_list = w.fail.as<uint32_t>();
      a.fail_count = &dctr->nfail;
      a.acc = w.acc.as<unsigned long long>();
      a.accb = w.accb.as<unsigned long long>();
      a.rep = rep;
      a.table_bits = table_bits;
      a.chunk_rows = chunk_rows;
      a.chunk_keys = w.chunk_keys.as<uint4>();
      a.chunk_key_cap = level_chunk_keys(table_bits);
      a.spread = dctr->spread;
      a.error_flag = &dctr->err;
      a.sup = w.sup.as<int32_t>();
      a.nkeys = nkeys;
      a.p_base = (int32_t)p_begin;
      a.K = K;
      a.minsup_job = w.minsup.as<int32_t>();
      a.freq = reinterpret_cast<unsigned long long *>(dctr + 1);
      a.freq_cap = fcap;
      a.freq_count = &dctr->fcount;
      a.tile_count = w.tiles.as<uint32_t>();
      a.pid = root ? nullptr : w.row(cur).pid.as<int32_t>();
      a.rgraph = root ? nullptr : w.row(cur).rgraph.as<int32_t>();
      a.vt = root ? nullptr : w.row(cur).vt.as<uint8_t>();
      a.rbound = root ? nullptr : w.row(cur).rbound.as<uint32_t>();
      a.rlist = root ? nullptr : w.row(cur).rlist.as<unsigned long long>();
      a.S = S;
      a.meta = root ? nullptr : reinterpret_cast<const PatMeta *>(w.row(cur).upl.as<char>());
      a.mask_stride = root ? 0 : mask_stride;
      a.pmask = a.mask_stride ? reinterpret_cast<const unsigned long long *>(w.row(cur).upl.as<char>() + mask_off)
                              : nullptr;
      hip_check(launch_clear_counters(a, s), "clear counters");
      const bool time_stage = opt.stage_events && !root && !w.arena;
      if (time_stage) {
        for (hipEvent_t &e : w.ev)
          if (!e) hip_check(hipEventCreate(&e), "hipEventCreate");
        hip_check(hipEventRecord(w.ev[0], s), "hipEventRecord");
      }
      int64_t chunk0 = 0;
      for (int64_t lo = p_begin; lo < p_end; lo += prange) {
        const int64_t hi = std::min<int64_t>(p_end, lo + prange);
        a.p_lo = (int32_t)lo;
        a.p_hi = (int32_t)hi;
        a.row_lo = prow(lo);
        a.row_hi = prow(hi);
        a.chunk0 = chunk0;
        a.n_chunks = ceil_div(a.row_hi - a.row_lo, chunk_rows);
        hip_check(launch_level(a, root, s), "k_level");
        chunk0 += a.n_chunks;
      }
      // the recount and the write see the whole piece (segment starts are
      // pattern changes at the launch boundaries anyway)
      a.row_lo = prow_lo;
      a.row_hi = prow_hi;
      a.chunk0 = 0;
      a.n_chunks = total_chunks;
      hip_check(launch_level_fix(a, root, s), "k_level_fix");
      if (time_stage) hip_check(hipEventRecord(w.ev[1], s), "hipEventRecord");
      hip_check(launch_reduce(a, root, s), "k_reduce");
      if (canon_dev) {
        CanonArgs c;
        std::memset(&c, 0, sizeof(c));
        c.freq = a.freq;
        c.freq_count = a.freq_count;
        c.freq_cap = fcap;
        c.meta = a.meta;
        c.p_base = (int32_t)p_begin;
        c.pcode = reinterpret_cast<const uint32_t *>(w.row(cur).upl.as<char>() + code_off);
        c.pvlab = reinterpret_cast<const uint8_t *>(w.row(cur).upl.as<char>() + vlab_off);
        c.len = L.len;
        c.K = K;
        c.S = S;
        c.NEL = NEL;
        c.NVL = NVL;
        // one wave per record; the grid loops over however many there are
        const int64_t grid = std::min<int64_t>(fcap, std::max<int64_t>(256, std::min<int64_t>(8192, 2 * (int64_t)last_fcount)));
        hip_check(launch_canon_level(c, (int)grid, s), "k_canon_level");
      }
      // counters and (most often all of) the frequent keys in one copy
      guess = std::min<uint32_t>(fcap, std::max<uint32_t>(1024, 2 * last_fcount));
      const size_t head = sizeof(Counters) + sizeof(Freq) * guess;
      hbuf = static_cast<const char *>(w.down.ensure(head));
      stage_down(w, w.down, dctr, head, s);
      hip_check(hipStreamSynchronize(s), "k_level");
      std::memcpy(&cnt, hbuf, sizeof(cnt));
      if (cnt.err) throw Error(GM_E_DEVICE, "internal: level step error " + std::to_string(cnt.err));
      ecand = max_fill = pruned = failed = 0;
      unsigned long long written = 0, t_start = 0, t_end = 0;
      for (int i = 0; i < kSpread; ++i) {
        const unsigned long long *line = cnt.spread + i * kSpreadStride;
        ecand += line[kSpreadCand];
        pruned += line[kSpreadPruned];
        max_fill = std::max(max_fill, line[kSpreadFill]);
        failed += line[kSpreadFailed];
        written += line[kSpreadWritten];
        t_start = std::max(t_start, line[kSpreadStart]);
        t_end = std::max(t_end, line[kSpreadEnd]);
      }
      if (written != written_expect)
        throw Error(GM_E_DEVICE, "child rows written (" + std::to_string(written) +
                                     ") differ from the embedding counts (" + std::to_string(written_expect) + ")");
      t_start = ~t_start;
      ms = (t_end > t_start && t_start != ~0ull) ? (double)(t_end - t_start) / wall_khz : 0.0;
      if (time_stage) {
        float ev_ms = 0.f;
        hip_check(hipEventElapsedTime(&ev_ms, w.ev[0], w.ev[1]), "hipEventElapsedTime");
        ev_stage_ms += ev_ms;
      }
      if (cnt.fcount <= fcap) break;
      // the list was too short: the stage is run again with room for every
      // frequent key, from cleared counters (k_reduce kept the frequent
      // keys' replicas, listed or not)
      if (attempt > 0) throw Error(GM_E_DEVICE, "internal: frequent-key count changed between runs");
      fcap = cnt.fcount;
      hip_check(launch_fill8(w.acc.p, w.acc.cap, s), "clear counters");
      hip_check(launch_fill8(w.accb.p, w.accb.cap, s), "clear counters");
    }
    if (!root) {
      ext_ms += ms;
      ext_launches++;
    }
    // a level with recounted chunks wants a larger table next time
    if (failed * 64 > (unsigned long long)total_chunks && table_bits < 11 && opt.table_bits == 0) ++table_bits;
    last_fcount = cnt.fcount;
    const int64_t F = cnt.fcount;
    if (F) {
      const Freq *src = reinterpret_cast<const Freq *>(hbuf + sizeof(Counters));
      if (F > (int64_t)guess) {  // the rest of the list, after the part downloaded
        frall.resize(F);
        std::memcpy(frall.data(), src, sizeof(Freq) * guess);
        const size_t rest = sizeof(Freq) * (F - guess);
        char *h2 = static_cast<char *>(w.down.ensure(rest));
        stage_down(w, w.down, reinterpret_cast<const char *>(a.freq) + sizeof(Freq) * guess, rest, s);
        hip_check(hipStreamSynchronize(s), "download frequent keys");
        std::memcpy(frall.data() + guess, h2, rest);
        src = frall.data();
      }
      // in key order already (k_reduce_list)
      fr.resize(F);
      std::memcpy(fr.data(), src, sizeof(Freq) * F);
      for (int64_t i = 1; i < F; ++i)
        if (fr[i].key <= fr[i - 1].key) throw Error(GM_E_DEVICE, "internal: frequent keys out of order");
    }
    dev_sec += tdev.sec();

    // Under a budget: the children of the piece's keys not written yet
    // (key_lo on) must fit beside the levels held.  Else the piece is cut at
    // the last parent pattern whose children fit and its stage run again
    // over that range (results of a run are per pattern, so the shorter run
    // finds the same keys); a single pattern whose children do not fit is
    // written in key ranges, each with at least one child.
    // (piece keys: the pattern's next key when it is split over pieces)
    const uint64_t key_lo = Fr.key_next;
    uint64_t key_cut = ~0ull;  // keys [key_lo, key_cut) become children
    if (budget > 0 && !past_max && F > 0) {
      const int CSb = tuple_stride(root ? 2 : std::min(L.max_nv + 1, kMaxPatV));
      auto kbytes = [&](int64_t i) { return rows_bytes((int64_t)fr[i].proj, CSb) + (int64_t)fr[i].items * 8; };
      int64_t total = 0;
      for (int64_t i = 0; i < F; ++i)
        if (fr[i].key >= key_lo) total += kbytes(i);
      if (total > avail) {
        // the last parent boundary whose children fit
        int64_t acc_b = 0, p_cut = p_begin;
        for (int64_t i = 0; i < F;) {
          const int64_t p = parent_of(fr[i].key);
          int64_t grp = 0, j = i;
          for (; j < F && parent_of(fr[j].key) == p; ++j)
            if (fr[j].key >= key_lo) grp += kbytes(j);
          if (acc_b + grp > avail) {
            p_cut = p;
            break;
          }
          acc_b += grp;
          i = j;
        }
        const bool measured_again = p_end_measured > 0;
        if ((p_cut > p_begin && p_cut < p_end) || (p_cut == p_begin && p_end > p_begin + 1)) {
          // run the stage again over the patterns whose children fit (or
          // the first pattern alone), from cleared counters
          Fr.p_end_next = std::max<int64_t>(p_cut, p_begin + 1);
          if (measured_again && Fr.p_end_next >= p_end) throw Error(GM_E_DEVICE, "internal: piece did not shrink");
          hip_check(launch_fill8(w.acc.p, w.acc.cap, s), "clear counters");
          hip_check(launch_fill8(w.accb.p, w.accb.cap, s), "clear counters");
          continue;
        }
        if (p_end == p_begin + 1) {
          // one pattern: its keys in order while they fit, at least one
          int64_t acc_k = 0;
          bool first = true;
          for (int64_t i = 0; i < F; ++i) {
            if (fr[i].key < key_lo) continue;
            if (!first && acc_k + kbytes(i) > avail) {
              key_cut = fr[i].key;
              break;
            }
            acc_k += kbytes(i);
            first = false;
          }
        }
      }
    }
    // a key of this piece that becomes a child now
    auto in_piece = [&](int64_t i) { return fr[i].key >= key_lo && fr[i].key < key_cut; };
    // e_cand and the stage statistics count each pattern's enumeration once
    // (a piece is either all new patterns or one pattern seen before)
    if (p_begin < Fr.p_counted && p_end > Fr.p_counted)
      throw Error(GM_E_DEVICE, "internal: piece mixes counted and new patterns");
    if (p_begin >= Fr.p_counted) {
      if (!root) st.e_cand += ecand;
      st.split_passes += failed;
      st.pruned += pruned;
      st.max_table_fill = std::max<uint64_t>(st.max_table_fill, max_fill);
      Fr.p_counted = p_end;
    }
    Timer t_host;

    // decode the frequent keys into child tuples
    Timer tcan;
    const int new_len = root ? 1 : L.len + 1;
    kidx.resize(F);
    ktup.resize(F);
    knv.resize(F);
    knode.resize(F);
    kalive.resize(F);
    auto decode = [&](int64_t i) {
      const uint64_t key = fr[i].key, q = fdK.div(key);
      const uint64_t x = key - q * (uint64_t)K;  // the key within its pattern's
      DFS t;
      int cnv;
      if (root) {
        const uint64_t x1 = fdNVL.div(x), fl = fdNEL.div(x1);
        t = DFS{0, 1, (int)fl, (int)(x1 - fl * NEL), (int)(x - x1 * NVL)};
        cnv = 2;
      } else {
        const int64_t p = (int64_t)q + p_begin;
        const uint8_t *pv = L.vl(p);
        const int nv = L.nv[p], maxtoc = nv - 1;
        if (x < (uint64_t)S * NEL) {
          const uint64_t to = fdNEL.div(x);
          t = DFS{maxtoc, (int)to, pv[maxtoc], (int)(x - to * NEL), pv[to]};
          cnv = nv;
        } else {
          const uint64_t y = x - (uint64_t)S * NEL, y1 = fdNVL.div(y), f = fdNEL.div(y1);
          t = DFS{(int)f, nv, pv[f], (int)(y1 - f * NEL), (int)(y - y1 * NVL)};
          cnv = nv + 1;
        }
      }
      ktup[i] = t;
      knv[i] = (uint8_t)std::min(cnv, 255);
      knode[i] = 0;
    };
    // the canonical check of child i (its parent's code plus its tuple)
    auto is_min_child = [&](int64_t i) {
      const int64_t p = parent_of(fr[i].key);
      thread_local std::vector<DFS> code;
      code.assign(L.c(p), L.c(p) + L.len);
      code.push_back(ktup[i]);
      return dfs_code_is_min(code.data(), new_len);
    };
    // Speculation: at later levels every frequent key becomes a child at
    // once (ids in key order), and the device writes the children while the
    // host checks which codes are minimal; the others are marked dead in
    // their metadata and the next level skips their rows.  Not at the root
    // (every 1-edge code is minimal), in query mode (the trie selects), past
    // max_edges, or when a child would have more than kMaxPatV vertices.
    // adaptive: speculate while the last level's non-minimal children held
    // at most 30 % of its rows (each costs a written, unused row; molecule-
    // like databases, rich in symmetric patterns, hold about half)
    const bool want_spec = spec_wanted && !canon_dev && !root && !opt.queries && !past_max;
    // A speculative level decodes its keys while the device writes; the
    // children's tuple stride needs only their largest vertex count: the
    // parent's plus one for a forward key, from the keys in order (the
    // parent's key range advances with them, no division).  Otherwise the
    // keys are decoded now, in parallel when there are many.
    int max_nv_exact = 0;
    if (want_spec) {
      int64_t p = p_begin;
      uint64_t p_end_key = (uint64_t)K;  // keys of pattern p are below this
      const uint64_t bwd_keys = (uint64_t)S * NEL;
      for (int64_t i = 0; i < F; ++i) {
        const uint64_t key = fr[i].key;
        while (key >= p_end_key) {
          ++p;
          p_end_key += (uint64_t)K;
        }
        const uint64_t x = key - (p_end_key - (uint64_t)K);
        max_nv_exact = std::max(max_nv_exact, (int)L.nv[p] + (x >= bwd_keys ? 1 : 0));
      }
    }
    const bool spec = want_spec && max_nv_exact <= kMaxPatV;
    const bool late_decode = spec;
    if (!spec) Pool::get().run(F, threads, decode, 4096);
    // the device's answer for frequent key i (written by k_canon_level as
    // answer + 1 in bits 32-39 of the support word); undecided keys are
    // checked here
    std::atomic<uint64_t> host_keys(0);
    auto dev_min = [&](int64_t i) {
      const unsigned code = (unsigned)((fr[i].sup >> 32) & 0xFF);
      if (code == 0 || code > kCanonHost + 1)
        throw Error(GM_E_DEVICE, "internal: frequent key without a minimum-code answer");
      if (code - 1 == kCanonHost) {
        host_keys.fetch_add(1, std::memory_order_relaxed);
        return is_min_child(i);
      }
      return code - 1 == kCanonMin;
    };
    if (!spec) {
      std::atomic<bool> too_big(false);
      Pool::get().run(F, threads, [&](int64_t i) {
        const uint32_t sup = (uint32_t)fr[i].sup;
        const int cnv = knv[i];
        bool keep = false;
        if (!in_piece(i)) {  // a later piece's
          keep = false;
        } else if (opt.queries) {  // follow the query codes only
          const int32_t parent = root ? 0 : L.node[parent_of(fr[i].key)];
          const int32_t node = opt.queries->child(parent, ktup[i]);
          knode[i] = node;
          keep = node >= 0 && query_reached(node, sup) && cnv <= kMaxPatV;
        } else if (root) {
          keep = true;
        } else if (!past_max && (canon_dev ? dev_min(i) : is_min_child(i))) {
          if (cnv > kMaxPatV)
            too_big = true;
          else
            keep = true;
        }
        kidx[i] = keep ? 0 : -1;
        kalive[i] = keep ? 1 : 0;
      });
      if (too_big) throw Error(GM_E_UNSUPPORTED, "pattern with more than 64 vertices is frequent");
    } else {
      for (int64_t i = 0; i < F; ++i) kidx[i] = in_piece(i) ? 0 : -1;  // decoded while the device writes
    }
    canon_sec += tcan.sec();
    double us_canon = tcan.sec() * 1e6;
    st.canon_host_keys += host_keys.load();
    st.n_candidates += root ? 0 : F;

    // children in key order; child c's rows are [row0[c], row0[c + 1])
    int64_t C = 0, child_rows = 0;
    int max_nv = 0;
    for (int64_t i = 0; i < F; ++i) {
      if (kidx[i] < 0) continue;
      kidx[i] = C++;
      child_rows += (int64_t)fr[i].proj;
      if (!late_decode) max_nv = std::max(max_nv, (int)knv[i]);
    }
    if (late_decode) max_nv = max_nv_exact;
    const int CS = C > 0 ? tuple_stride(max_nv) : 4;  // the children's tuple stride
    const bool wide = CS > kNarrowPatV;                 // and metadata layout
    N.resize(C, new_len);
    nproj.resize(C);
    double us_host = 0, us_overlap = 0;
    int64_t alive = 0, alive_rows = 0;  // minimal children and their embeddings
    int child_buf = 0, child_mask_stride = 0;  // the children's row set and upload layout
    size_t child_mask_off = 0, child_code_off = 0, child_vlab_off = 0;
    uint64_t child_items = 0;  // the children's list entries
    // Upload (mapped staging, read by k_cursors and copied by it): the
    // children's metadata, then per frequent key its child id (-1:
    // dropped), per child its first row and first list entry, then the
    // children's path masks.  Child c's rows follow the earlier children's
    // (embedding counts), its list entries likewise (item sums).
    const size_t meta_all = (size_t)meta_bytes(CS) * C;
    const size_t fc_off = (meta_all + 7) & ~size_t(7);
    const size_t rb_off = (fc_off + sizeof(int32_t) * F + 7) & ~size_t(7);
    const size_t lb_off = rb_off + sizeof(unsigned long long) * C;
    // path masks of the children with >= 2 edges: a forward and a
    // backward word per rightmost-path vertex (at most max_nv of them)
    // (a 2-edge level written in pieces leaves no complete path table)
    if (new_len == 2 && !(p_begin == 0 && p_end == P && key_lo == 0 && key_cut == ~0ull)) t2_partial = true;
    const bool masks = prune && !t2_partial && new_len >= 2 && C > 0;
    const int mstride = masks ? max_nv : 0;
    const size_t masks_off = lb_off + sizeof(unsigned long long) * C;
    // then per child its packed code (from | to << 8 | elabel << 16 |
    // tolabel << 24 per tuple) and its vertex labels, for the next level's
    // minimum-code check on the device
    const size_t codes_off = masks_off + sizeof(uint64_t) * 2 * mstride * C;
    const size_t cvl_off = (codes_off + (want_codes ? sizeof(uint32_t) * new_len * C : 0) + 7) & ~size_t(7);
    const size_t up_bytes = (cvl_off + (want_codes ? (size_t)CS * C : 0) + 7) & ~size_t(7);  // copied in 8-byte words
    char *hu = F > 0 ? static_cast<char *>(w.up.ensure(up_bytes)) : nullptr;
    if (F > 0) {
      int32_t *hfc = reinterpret_cast<int32_t *>(hu + fc_off);
      unsigned long long *hrb = reinterpret_cast<unsigned long long *>(hu + rb_off);
      unsigned long long *hlb = reinterpret_cast<unsigned long long *>(hu + lb_off);
      int64_t run = 0;
      for (int64_t i = 0; i < F; ++i) {
        const int64_t c = kidx[i];
        hfc[i] = (int32_t)c;
        if (c < 0) continue;
        hrb[c] = (unsigned long long)run;
        hlb[c] = (unsigned long long)child_items;
        N.row0[c] = run;
        run += (int64_t)fr[i].proj;
        child_items += fr[i].items;
        // k_write reserves rows << 32 | list entries in one 64-bit atomic
        if (fr[i].proj >= (1ull << 32) || fr[i].items >= (1ull << 32))
          throw Error(GM_E_UNSUPPORTED, "a pattern with 2^32 or more embeddings or candidate-list entries");
      }
      N.row0[C] = run;
    }
    if (F > 0) stage_up(w, w.up, fc_off, masks_off - fc_off, s);  // child ids, bases
    if (F > 0 && C == 0) {  // nothing to write: the frequent keys' replicas are cleared
      const char *su = static_cast<const char *>(w.up.dev);
      hip_check(launch_cursors(a, nullptr, nullptr, 0, reinterpret_cast<const int32_t *>(su + fc_off), nullptr, nullptr,
                               F, s),
                "k_cursors");
    }
    if (C > 0) {
      // The embedding counts size the child level exactly; k_write reserves
      // each child's rows per chunk from the cursors of the chunk's replica.
      // The chosen keys point at their child ids and the cursors are set
      // first (read from the mapped staging buffer) and the write pass
      // starts; the children's metadata (the next level's) follows in one
      // upload.
      Timer tdev2;
      // a row set no level on the stack holds (this one's is read by the write)
      auto held_by_stack = [&](int b) {
        for (const Frame &f : frames)
          if (f.buf == b) return true;
        return false;
      };
      int nxt = 0;
      while (held_by_stack(nxt)) ++nxt;
      if (w.arena && nxt > 1) throw Error(GM_E_DEVICE, "internal: row set beyond the static workspace's two");
      child_buf = nxt;
      PatMeta *hm = reinterpret_cast<PatMeta *>(hu);  // narrow layout (else hmw)
      PatMetaWide *hmw = reinterpret_cast<PatMetaWide *>(hu);
      char *du = static_cast<char *>(grow(w.row(nxt).upl, up_bytes, s));
      const char *su = static_cast<const char *>(w.up.dev);
      grow(w.row(nxt).pid, sizeof(int32_t) * std::max<int64_t>(child_rows, 1), s);
      grow(w.row(nxt).rgraph, sizeof(int32_t) * std::max<int64_t>(child_rows, 1), s);
      grow(w.row(nxt).vt, (size_t)CS * std::max<int64_t>(child_rows, 1), s);
      grow(w.row(nxt).rbound, sizeof(uint32_t) * std::max<int64_t>(child_rows, 1), s);
      grow(w.row(nxt).rlist, sizeof(unsigned long long) * std::max<int64_t>(child_rows, 1), s);
      grow(w.cur_rows, sizeof(unsigned long long) * C * rep, s);
      grow(w.cur_list, sizeof(unsigned long long) * C * rep, s);
      grow(w.cursor, sizeof(unsigned long long) * C * rep, s);
      a.cur_rows = w.cur_rows.as<unsigned long long>();
      a.cur_list = w.cur_list.as<unsigned long long>();
      a.cursor = w.cursor.as<unsigned long long>();
      a.ccap = child_rows;
      a.cpid = w.row(nxt).pid.as<int32_t>();
      a.crgraph = w.row(nxt).rgraph.as<int32_t>();
      a.cvt = w.row(nxt).vt.as<uint8_t>();
      a.cbound = w.row(nxt).rbound.as<uint32_t>();
      a.crlist = w.row(nxt).rlist.as<unsigned long long>();
      a.CS = CS;
      a.chunk0 = 0;
      a.n_chunks = total_chunks;
      // child ids and cursors, then the write pass (launched back to back:
      // the device does not wait on the host between)
      hip_check(launch_cursors(a, du, su, 0, reinterpret_cast<const int32_t *>(su + fc_off),
                               reinterpret_cast<const unsigned long long *>(su + rb_off),
                               reinterpret_cast<const unsigned long long *>(su + lb_off), F, s),
                "k_cursors");
      hip_check(launch_write(a, root, s), "k_write");
      written_expect += (unsigned long long)child_rows;
      us_host = t_host.sec() * 1e6;

      // while the device writes the children: the canonical check of the
      // speculative children, their codes and metadata, the path masks
      Timer t_over;
      uint64_t *hmask = reinterpret_cast<uint64_t *>(hu + masks_off);
      auto build_child = [&](int64_t i) {
        const int64_t c = kidx[i];
        if (c < 0) return;
        const int64_t p = parent_of(fr[i].key);
        DFS *cc = N.c(c);
        uint8_t *cv = N.vl(c);
        const DFS &t = ktup[i];
        const int cnv = knv[i];
        if (root) {
          cc[0] = t;
          cv[0] = (uint8_t)t.fromlabel;
          cv[1] = (uint8_t)t.tolabel;
          N.job[c] = (int32_t)p;
        } else {
          std::memcpy(cc, L.c(p), sizeof(DFS) * L.len);
          cc[L.len] = t;
          std::memcpy(cv, L.vl(p), L.nv[p]);
          if (cnv > L.nv[p]) cv[cnv - 1] = (uint8_t)t.tolabel;
          N.job[c] = L.job[p];
        }
        N.nv[c] = (uint8_t)cnv;
        N.sup[c] = (uint32_t)fr[i].sup;
        N.node[c] = knode[i];
        N.alive[c] = kalive[i];
        nproj[c] = fr[i].proj;
        if (wide) {
          hmw[c] = make_meta<kMaxPatV>(cc, new_len, cv, cnv, N.job[c]);
          if (!kalive[i]) hmw[c].flags |= kPatDead;
        } else {
          hm[c] = make_meta<kNarrowPatV>(cc, new_len, cv, cnv, N.job[c]);
          if (!kalive[i]) hm[c].flags |= kPatDead;
        }
        if (want_codes) {
          uint32_t *pc = reinterpret_cast<uint32_t *>(hu + codes_off) + (size_t)c * new_len;
          for (int k = 0; k < new_len; ++k)
            pc[k] = (uint32_t)cc[k].from | (uint32_t)cc[k].to << 8 | (uint32_t)cc[k].elabel << 16 |
                    (uint32_t)cc[k].tolabel << 24;
          std::memcpy(hu + cvl_off + (size_t)c * CS, cv, CS);
        }
      };
      auto mask_child = [&](int64_t c) {
        if (!N.alive[c]) return;  // dead patterns' rows are never extended
        uint64_t *fw = hmask + (size_t)c * 2 * mstride, *bw = fw + mstride;
        const DFS *cc = N.c(c);
        const uint8_t *cv = N.vl(c);
        // nv, nrm and the rightmost path's from vertices, either layout
        const int m_nv = N.nv[c], m_nrm = wide ? hmw[c].nrm : hm[c].nrm;
        const uint32_t *m_rm = wide ? hmw[c].rm : hm[c].rm;
        const uint64_t *tj = &t2[(size_t)N.job[c] * NVL * AR];
        // per vertex: the arms that form a frequent path with every
        // pattern edge at it (one pass over the edges)
        uint64_t vm[kMaxPatV];
        for (int v = 0; v < m_nv; ++v) vm[v] = ~0ull;
        for (int k = 0; k < new_len; ++k) {
          const int u = cc[k].from, v = cc[k].to, e = cc[k].elabel * NVL;
          vm[u] &= tj[(size_t)cv[u] * AR + e + cv[v]];
          vm[v] &= tj[(size_t)cv[v] * AR + e + cv[u]];
        }
        const int rv = m_nv - 1;
        fw[0] = vm[rv];
        bw[0] = 0;
        for (int i = 0; i < m_nrm; ++i) {
          const int t = m_rm[i] & 0xFF;
          fw[1 + i] = vm[t];
          uint64_t b = 0;  // backward rv -> t: a new arm at both ends
          if (i > 0)
            for (int el = 0; el < NEL; ++el)
              b |= ((vm[rv] >> (el * NVL + cv[t])) & (vm[t] >> (el * NVL + cv[rv])) & 1ull) << el;
          bw[1 + i] = b;
        }
      };
      Timer tc2;
      if (!masks || (t2_ready && new_len > 2)) {
        // one pass per child: check, metadata, masks (the path table is known)
        Pool::get().run(F, threads, [&](int64_t i) {
          if (kidx[i] < 0) return;
          if (spec) {
            if (late_decode) decode(i);
            kalive[i] = is_min_child(i) ? 1 : 0;
          }
          build_child(i);
          if (masks) mask_child(kidx[i]);
        });
      } else {
        // the 2-edge level: its minimal children are the path table
        if (spec)
          Pool::get().run(F, threads, [&](int64_t i) {
            if (late_decode) decode(i);
            kalive[i] = is_min_child(i) ? 1 : 0;
          });
        Pool::get().run(F, threads, build_child, 4096);
        if (new_len == 2) {
          // the job's frequent 2-edge patterns: (0,1,a,e1,b) then
          // (1,2,b,e2,c) (centre b) or (0,2,a,e2,c) (centre a)
          t2.assign((size_t)db.n_jobs * NVL * AR, 0ull);
          for (int64_t c = 0; c < C; ++c) {
            if (!N.alive[c]) continue;
            const DFS *cc = N.c(c);
            const bool at1 = cc[1].from == 1;
            const int centre = at1 ? cc[0].tolabel : cc[0].fromlabel;
            const int arm1 = cc[0].elabel * NVL + (at1 ? cc[0].fromlabel : cc[0].tolabel);
            const int arm2 = cc[1].elabel * NVL + cc[1].tolabel;
            uint64_t *row = &t2[((size_t)N.job[c] * NVL + centre) * AR];
            row[arm1] |= 1ull << arm2;
            row[arm2] |= 1ull << arm1;
          }
          t2_ready = true;
        }
        if (t2_ready) Pool::get().run(C, threads, mask_child, 4096);
      }
      if (spec) {  // the check with the metadata it overlaps
        canon_sec += tc2.sec();
        us_canon += tc2.sec() * 1e6;
      }
      int64_t dead_rows = 0;
      N.max_nv = 0;
      for (int64_t c = 0; c < C; ++c) {
        N.max_nv = std::max(N.max_nv, (int)N.nv[c]);
        if (!N.alive[c]) {
          dead_rows += (int64_t)nproj[c];
          continue;
        }
        ++alive;
        alive_rows += (int64_t)nproj[c];
      }
      // non-minimal children's share of the rows: measured when the check
      // ran before the write too (the children the write would have made)
      int64_t nonmin_rows = dead_rows;
      if (!spec && !root && !opt.queries)
        for (int64_t i = 0; i < F; ++i)
          if (kidx[i] < 0 && in_piece(i)) nonmin_rows += (int64_t)fr[i].proj;
      dead_share = (double)nonmin_rows / (double)std::max<int64_t>(1, alive_rows + nonmin_rows);
      if (w.arena) {
        hip_check(hipMemcpyAsync(du, hu, up_bytes, hipMemcpyHostToDevice, s), "upload children");
        hip_check(hipStreamSynchronize(s), "upload children");
      } else if (up_bytes < kSideUploadBytes) {
        hip_check(launch_copy8(du, su, up_bytes, s), "upload children");
      } else {
        // a DMA copy from the pinned staging buffer on the side stream, so
        // that it overlaps the write pass; the next level's launches wait
        // for it (and the host rewrites the staging buffer only after the
        // next level's synchronisation, which covers it).  Small uploads
        // stay on the stream: handles mining side by side (serving slots)
        // would otherwise double their streams on the few hardware queues.
        if (!w.side) hip_check(hipStreamCreateWithFlags(&w.side, hipStreamNonBlocking), "hipStreamCreate");
        if (!w.up_done) hip_check(hipEventCreateWithFlags(&w.up_done, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipMemcpyAsync(du, hu, up_bytes, hipMemcpyHostToDevice, w.side), "upload children");
        hip_check(hipEventRecord(w.up_done, w.side), "upload children");
        hip_check(hipStreamWaitEvent(s, w.up_done, 0), "upload children");
      }
      us_overlap = t_over.sec() * 1e6;
      dev_sec += tdev2.sec();
      size_t live = db.bytes + w.sup.cap + w.tiles.cap + w.acc.cap + w.accb.cap + w.freq.cap + w.rccnt.cap + w.cand.cap +
                    w.chunks.cap + w.fail.cap + w.cur_rows.cap + w.cur_list.cap + w.cursor.cap + w.chunk_keys.cap;
      for (const RowSet &rs : w.rows) live += rs.bytes();
      peak = std::max(peak, live);
      child_mask_off = masks_off;
      child_code_off = codes_off;
      child_vlab_off = cvl_off;
      child_mask_stride = masks && t2_ready ? mstride : 0;
    }
    if (root)
      st.e_root += alive_rows;
    else
      st.e_child += alive_rows;
    if (!opt.queries) {
      if (!report) {
        st.n_patterns += alive;
      } else {
        orig.resize(new_len);
        for (int64_t c = 0; c < C && !stopped; ++c) {
          if (!N.alive[c]) continue;
          st.n_patterns++;
          const DFS *cc = N.c(c);
          for (int k = 0; k < new_len; ++k)
            orig[k] = DFS{cc[k].from, cc[k].to, db.vl_values[cc[k].fromlabel], db.el_values[cc[k].elabel],
                          db.vl_values[cc[k].tolabel]};
          if (!report(N.job[c], orig.data(), new_len, N.sup[c], opt.collect_proj ? nproj[c] : 0)) stopped = true;
        }
      }
    }
    if (trace_levels)
      fprintf(stderr,
              "[gm level %d] piece %lld-%lld of %lld patterns, rows %lld of %lld K %lld chunks %lld table_bits %d stage_us %.1f "
              "candidates %llu pruned %llu frequent %u children %lld child_rows %lld minimal %lld spec %d "
              "failed %llu max_fill %llu list_cap %llu canon_dev %d host_us canon %.1f host %.1f overlap %.1f\n",
              level, (long long)p_begin, (long long)p_end, (long long)P, (long long)(prow_hi - prow_lo), (long long)rows_now,
              (long long)K, (long long)total_chunks, table_bits, ms * 1e3,
              ecand, pruned, cnt.fcount, (long long)C, (long long)child_rows, (long long)alive, spec ? 1 : 0, failed,
              max_fill, (unsigned long long)list_cap, canon_dev ? 1 : 0, us_canon, us_host, us_overlap);
    st.n_levels++;
    if (stopped) break;
    // the piece is written: the level continues after it, or is done (its
    // rows are still read by the queued write only); the children, if any,
    // are the next level
    if (key_cut != ~0ull) {
      Fr.key_next = key_cut;  // a key of pattern p_next, relative to it (its pieces start there)
    } else {
      Fr.p_next = p_end;
      Fr.key_next = 0;
    }
    if (budget > 0 && key_cut == ~0ull && prow_hi > prow_lo) {
      const int64_t cb = rows_bytes(child_rows, CS) + (int64_t)child_items * 8;
      piece_bytes_per_row = std::max(1.0, (double)cb / (double)(prow_hi - prow_lo));
    }
    if (Fr.p_next >= P) frames.pop_back();
    if (C > 0) {
      frames.emplace_back();
      Frame &cf = frames.back();
      std::swap(cf.L, N);
      cf.level = level + 1;
      cf.buf = child_buf;
      cf.S = CS;
      cf.n_rows = child_rows;
      cf.list_cap = std::max<uint64_t>(child_items, 1);  // the children's lists: exactly their item counts
      cf.mask_off = child_mask_off;
      cf.mask_stride = child_mask_stride;
      cf.code_off = child_code_off;
      cf.vlab_off = child_vlab_off;
      cf.codes_ready = want_codes;
    }
  }
  // the last write's row count
  if (written_expect) {
    Counters c;
    hip_check(hipMemcpyAsync(&c, w.freq.p, sizeof(c), hipMemcpyDeviceToHost, s), "download counters");
    hip_check(hipStreamSynchronize(s), "mining");
    unsigned long long written = 0;
    for (int i = 0; i < kSpread; ++i) written += c.spread[i * kSpreadStride + kSpreadWritten];
    if (written != written_expect && !stopped)
      throw Error(GM_E_DEVICE, "child rows written (" + std::to_string(written) +
                                   ") differ from the embedding counts (" + std::to_string(written_expect) + ")");
  }
  hip_check(hipStreamSynchronize(s), "mining");
  w.acc_zero = true;
  st.seconds_total = total.sec();
  st.seconds_device = dev_sec;
  st.seconds_canon = canon_sec;
  st.seconds_ext_kernel = ext_ms * 1e-3;
  st.seconds_stage_events = ev_stage_ms * 1e-3;
  st.ext_kernel_launches = ext_launches;
  st.peak_device_bytes = w.arena ? w.arena->peak : peak;
  if (stats) *stats = st;
  return !stopped;
}

}  // namespace gm
