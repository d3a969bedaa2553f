// miner.cpp -- device database upload and the level-synchronous driver.
//
// The reference grows the pattern tree depth-first on the host and launches
// a handful of kernels per pattern, each followed by a device sync
// (src/gspan_cuda/gspan_cuda_no_sort.cpp:275-489; kernel_execution.hpp:61-71).
// This driver grows it level by level instead: one level step extends the
// embeddings of *every* pattern of the current size in the same launches,
// so each step costs a fixed number of kernels and two small host round
// trips (the frequent-key list down, the chosen children up) no matter how
// many patterns the level holds.  Patterns are independent subtrees, so the
// set of reported (code, support) pairs is the one the depth-first oracle
// reports (src/gspan/gspan.cpp:172-274); only the emission order differs.
//
// Level step for patterns P_1..P_n with embeddings in graph-major rows:
//   1. k_support   : distinct-graph support of every (pattern, extension key)
//                    and the list of keys whose support reached min support;
//   2. host        : decode each frequent key into a child code, keep the
//                    minimal ones (canonical.cpp), number the children;
//   3. k_write     : child rows of the chosen keys, each graph's rows
//                    reserved from a level cursor (sized exactly by the
//                    embedding counts k_support gathered).
#include "miner.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <map>
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

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

void *DevBuf::ensure(size_t bytes) {
  if (bytes <= cap && p) return p;
  release();
  size_t want = std::max<size_t>(bytes, 256);
  hip_check(hipMalloc(&p, want), "hipMalloc");
  cap = want;
  return p;
}

namespace {

template <class T>
void upload(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
  b.ensure(sizeof(T) * std::max<size_t>(v.size(), 1));
  if (!v.empty()) hip_check(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s), "upload");
}

template <class T>
void download(std::vector<T> &v, const void *src, size_t n, hipStream_t s) {
  v.resize(n);
  if (n) hip_check(hipMemcpyAsync(v.data(), src, sizeof(T) * n, hipMemcpyDeviceToHost, s), "download");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

int64_t round4(int64_t x) { return (x + 3) & ~int64_t(3); }

}  // namespace

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

  std::vector<int32_t> &gvbase = pk.gvbase, &vadj = pk.vadj, &gjob = pk.gjob;
  std::vector<uint32_t> &adj = pk.adj;
  std::vector<uint8_t> &vlab = pk.vlab;
  pk.max_graph_v = pk.max_graph_adj = 0;
  gvbase.clear();
  vadj.clear();
  gjob.clear();
  adj.clear();
  vlab.clear();
  gvbase.reserve(ng + 1);
  vadj.reserve(nv + 1);
  adj.reserve(2 * ne);
  vlab.reserve(nv);
  gjob.reserve(ng);
  out.job_first_graph.assign(1, 0);
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
      for (int32_t e = e0; e < e1; ++e) {
        int32_t a = db->efrom[e], b = db->eto[e];
        uint32_t el = (uint32_t)emap(db->elabel[e]);
        adj[abase + fill[a]++] = (uint32_t)b | (el << 8) | ((uint32_t)dl[b] << 16);
        adj[abase + fill[b]++] = (uint32_t)a | (el << 8) | ((uint32_t)dl[a] << 16);
      }
    }
    ++job;
    out.job_first_graph.push_back((int32_t)gvbase.size());
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
  hip_check(hipStreamSynchronize(s), "upload database");
  out.bytes = 4 * pk.gvbase.size() + 4 * pk.vadj.size() + 4 * pk.adj.size() + pk.vlab.size() + 4 * pk.gjob.size();
}

namespace {

// Host view of one pattern of the current level.
struct Pat {
  std::vector<DFS> code;      // dense labels, fully labelled
  std::vector<uint8_t> vlab;  // dense label of each pattern vertex
  int32_t job = 0;
  uint32_t sup = 0;
};

PatMeta make_meta(const Pat &p) {
  PatMeta m;
  std::memset(&m, 0, sizeof(m));
  m.job = p.job;
  const int nv = (int)p.vlab.size();
  m.nv = (uint8_t)nv;
  m.minlabel = p.vlab[0];
  m.maxlabel = p.vlab[nv - 1];
  // rightmost path (src/globals/dfs_code.cpp:75-90)
  int old_from = -1, nrm = 0;
  for (int i = (int)p.code.size() - 1; i >= 0; --i) {
    const DFS &d = p.code[i];
    if (d.from < d.to && (nrm == 0 || old_from == d.to)) {
      m.rm[nrm++] = (uint32_t)d.from | ((uint32_t)d.elabel << 8) | ((uint32_t)p.vlab[d.to] << 16);
      old_from = d.from;
    }
  }
  m.nrm = (uint8_t)nrm;
  for (const DFS &d : p.code)
    if (d.from == nv - 1 && d.to < d.from) m.bwd_mask |= 1u << d.to;
  return m;
}

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
  // Run fn(i) for i in [0, n) on up to `threads` threads (the caller helps).
  void run(int64_t n, int threads, const std::function<void(int64_t)> &fn) {
    if (n <= 0) return;
    if (threads <= 1 || n < 64) {
      for (int64_t i = 0; i < n; ++i) fn(i);
      return;
    }
    std::unique_lock<std::mutex> lk(m_);
    while ((int)workers_.size() < threads - 1) workers_.emplace_back([this]() { loop(); });
    fn_ = &fn;
    n_ = n;
    next_ = 0;
    err_ = nullptr;
    active_ = (int)workers_.size();
    ++gen_;
    lk.unlock();
    cv_.notify_all();
    work();
    lk.lock();
    done_.wait(lk, [this]() { return active_ == 0; });
    fn_ = nullptr;
    if (err_) std::rethrow_exception(err_);
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
  void work() {
    for (;;) {
      const int64_t i = next_.fetch_add(32);
      if (i >= n_) break;
      const int64_t e = std::min(n_, i + 32);
      try {
        for (int64_t j = i; j < e; ++j) (*fn_)(j);
      } catch (...) {
        std::lock_guard<std::mutex> lk(err_m_);
        if (!err_) err_ = std::current_exception();
        next_ = n_;
      }
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(m_);
      cv_.wait(lk, [&]() { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      lk.unlock();
      work();
      lk.lock();
      if (--active_ == 0) done_.notify_all();
    }
  }
  std::mutex m_, err_m_;
  std::condition_variable cv_, done_;
  std::vector<std::thread> workers_;
  const std::function<void(int64_t)> *fn_ = nullptr;
  std::atomic<int64_t> next_{0};
  int64_t n_ = 0;
  int active_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
  std::exception_ptr err_;
};

}  // namespace

bool mine(DeviceDB &db, const MineOptions &opt, const ReportFn &report, gm_stats *stats, hipStream_t s) {
  Timer total;
  const int32_t G = db.n_graphs;
  const int32_t NEL = db.NEL, NVL = db.NVL;
  int threads = opt.canon_threads > 0 ? opt.canon_threads
                                      : (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
  gm_stats st;
  std::memset(&st, 0, sizeof(st));

  // device state
  DevBuf minsup_d, sup, sup_rep, proj_rep, freq, counters, meta_d, gs_a, gs_b, gc_a, gc_b, pid_a, pid_b, vt_a, vt_b,
      rowcnt, idx_d, val_d;
  upload(minsup_d, opt.minsup, s);
  counters.ensure(64);
  for (DevBuf *b : {&gs_a, &gs_b, &gc_a, &gc_b}) b->ensure(sizeof(int64_t) * std::max(G, 1));
  uint32_t *freq_count = reinterpret_cast<uint32_t *>(counters.as<char>());
  unsigned long long *ecand = reinterpret_cast<unsigned long long *>(counters.as<char>() + 8);
  int32_t *error_flag = reinterpret_cast<int32_t *>(counters.as<char>() + 16);
  unsigned long long *split_passes = reinterpret_cast<unsigned long long *>(counters.as<char>() + 24);
  unsigned long long *cursor = reinterpret_cast<unsigned long long *>(counters.as<char>() + 32);

  hipEvent_t ev0, ev1;
  hip_check(hipEventCreate(&ev0), "hipEventCreate");
  hip_check(hipEventCreate(&ev1), "hipEventCreate");
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } evg{ev0, ev1};
  double ext_ms = 0, dev_sec = 0, canon_sec = 0;
  uint64_t ext_launches = 0;
  size_t peak = db.bytes;

  LevelArgs a;
  std::memset(&a, 0, sizeof(a));
  a.gvbase = db.gvbase.as<int32_t>();
  a.vadj = db.vadj.as<int32_t>();
  a.adj = db.adj.as<uint32_t>();
  a.vlab = db.vlab.as<uint8_t>();
  a.gjob = db.gjob.as<int32_t>();
  a.NEL = NEL;
  a.NVL = NVL;
  a.minsup_job = minsup_d.as<int32_t>();
  a.freq_count = freq_count;
  a.ecand = ecand;
  a.error_flag = error_flag;
  a.split_passes = split_passes;
  a.cursor = cursor;
  // LDS: dedup table of 2^11 keys; stage the largest graph's adjacency when
  // it fits 16 KB (larger graphs are read from HBM)
  a.table_bits = 11;
  a.stage_v = db.max_graph_v + 1;
  a.stage_a = std::min<int32_t>(db.max_graph_adj, 4096);

  std::vector<Pat> pats;      // patterns of the current level
  bool stopped = false;
  int level = 0;              // 0 = root stage
  int S = 0;                  // tuple stride of the current level
  int64_t n_rows = 0;
  DevBuf *gs_cur = &gs_a, *gs_nxt = &gs_b, *gc_cur = &gc_a, *gc_nxt = &gc_b;
  DevBuf *pid_cur = &pid_a, *pid_nxt = &pid_b, *vt_cur = &vt_a, *vt_nxt = &vt_b;

  auto to_original = [&](const std::vector<DFS> &c) {
    std::vector<DFS> o(c);
    for (DFS &d : o) {
      d.fromlabel = db.vl_values[d.fromlabel];
      d.tolabel = db.vl_values[d.tolabel];
      d.elabel = db.el_values[d.elabel];
    }
    return o;
  };

  for (;;) {
    const bool root = level == 0;
    if (!root && pats.empty()) break;
    Timer tdev;
    const int64_t P = root ? db.n_jobs : (int64_t)pats.size();
    const int64_t K = root ? (int64_t)NVL * NEL * NVL : (int64_t)S * NEL * (NVL + 1);
    const int64_t nkeys = P * K;
    std::vector<PatMeta> metas;
    if (!root) {
      metas.resize(P);
      for (int64_t i = 0; i < P; ++i) metas[i] = make_meta(pats[i]);
      upload(meta_d, metas, s);
    }
    // Support and embedding counters: `rep` replicas (summed by k_reduce)
    // spread the atomics of small key spaces, where every graph hits the
    // same keys.  Replicas stay within 256 KB of support counters so the
    // atomics stay in L2.
    int rep = 1;
    while (rep < 64 && (int64_t)rep * 2 * nkeys * 4 <= (int64_t)(256 << 10)) rep *= 2;
    sup.ensure(sizeof(int32_t) * nkeys);
    sup_rep.ensure(sizeof(int32_t) * nkeys * rep);
    proj_rep.ensure(sizeof(unsigned long long) * nkeys * rep);
    hip_check(hipMemsetAsync(sup_rep.p, 0, sizeof(int32_t) * nkeys * rep, s), "memset supports");
    hip_check(hipMemsetAsync(proj_rep.p, 0, sizeof(unsigned long long) * nkeys * rep, s), "memset embeddings");
    hip_check(hipMemsetAsync(counters.p, 0, 64, s), "memset counters");
    uint32_t fcap = (uint32_t)std::min<int64_t>(nkeys, std::max<int64_t>(1 << 20, P * 64));
    freq.ensure(3 * sizeof(unsigned long long) * fcap);
    a.gstart = root ? nullptr : gs_cur->as<int64_t>();
    a.gcount = root ? nullptr : gc_cur->as<int64_t>();
    a.pid = root ? nullptr : pid_cur->as<int32_t>();
    a.vt = root ? nullptr : vt_cur->as<uint8_t>();
    a.S = S;
    a.meta = root ? nullptr : meta_d.as<PatMeta>();
    a.K = K;
    a.sup = sup.as<int32_t>();
    a.nkeys = nkeys;
    a.sup_rep = sup_rep.as<int32_t>();
    a.proj_rep = proj_rep.as<unsigned long long>();
    a.rep = rep;
    a.freq = freq.as<unsigned long long>();
    a.freq_cap = fcap;
    // workgroup size: enough lanes for the average graph's rows (or vertices)
    {
      const double per_graph =
          root ? (double)db.n_vertices / std::max(1, G) : (double)n_rows / (double)std::max(1, G);
      a.block = per_graph > 128 ? 256 : per_graph > 64 ? 128 : 64;
    }

    // 32-bit dedup keys: launch over pattern ranges with range * K < 2^32
    const int64_t chunk = std::max<int64_t>(1, (int64_t)0xFFFFFFFEll / K);
    hip_check(hipEventRecord(ev0, s), "hipEventRecord");
    for (int64_t lo = 0; lo < P; lo += chunk) {
      a.p_lo = (int32_t)lo;
      a.p_hi = (int32_t)std::min<int64_t>(P, lo + chunk);
      hip_check(launch_support(a, G, root, s), "k_support");
    }
    hip_check(hipEventRecord(ev1, s), "hipEventRecord");
    hip_check(launch_reduce(a, root, s), "k_reduce");
    struct {
      uint32_t fcount, pad;
      unsigned long long ecand;
      int32_t err, pad2;
      unsigned long long splits;
      unsigned long long cursor;
    } cnt;
    hip_check(hipMemcpyAsync(&cnt, counters.p, sizeof(cnt), hipMemcpyDeviceToHost, s), "download counters");
    hip_check(hipStreamSynchronize(s), "k_support");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, ev0, ev1), "hipEventElapsedTime");
    if (!root) {
      ext_ms += ms;
      ext_launches++;
      st.e_cand += cnt.ecand;
    }
    st.split_passes += cnt.splits;
    if (cnt.err) throw Error(GM_E_UNSUPPORTED, "per-graph deduplication table overflow");
    if (cnt.fcount > fcap)
      throw Error(GM_E_OOM, "frequent-key list overflow (" + std::to_string(cnt.fcount) + " keys)");
    struct Freq {
      unsigned long long key, sup, proj;
    };
    std::vector<Freq> fr(cnt.fcount);
    if (cnt.fcount) {
      hip_check(hipMemcpyAsync(fr.data(), freq.p, sizeof(Freq) * cnt.fcount, hipMemcpyDeviceToHost, s),
                "download frequent keys");
      hip_check(hipStreamSynchronize(s), "download frequent keys");
    }
    std::sort(fr.begin(), fr.end(), [](const Freq &x, const Freq &y) { return x.key < y.key; });
    dev_sec += tdev.sec();

    // decode frequent keys into child codes; keep the minimal ones
    Timer tcan;
    const int64_t F = (int64_t)fr.size();
    std::vector<Pat> kids(F);
    std::vector<char> keep(F, 0);
    bool too_big = false;
    Pool::get().run(F, threads, [&](int64_t i) {
      const uint64_t key = fr[i].key;
      Pat &c = kids[i];
      c.sup = (uint32_t)fr[i].sup;
      if (root) {
        const int64_t j = key / K, x = key % K;
        const int fl = (int)(x / (NEL * NVL)), el = (int)((x / NVL) % NEL), tl = (int)(x % NVL);
        c.job = (int32_t)j;
        c.code.push_back(DFS{0, 1, fl, el, tl});
        c.vlab = {(uint8_t)fl, (uint8_t)tl};
        keep[i] = 1;
        return;
      }
      const int64_t p = key / K, x = key % K;
      const Pat &par = pats[p];
      c.job = par.job;
      c.code = par.code;
      c.vlab = par.vlab;
      const int nv = (int)par.vlab.size(), maxtoc = nv - 1;
      if (x < (int64_t)S * NEL) {
        const int t = (int)(x / NEL), el = (int)(x % NEL);
        c.code.push_back(DFS{maxtoc, t, par.vlab[maxtoc], el, par.vlab[t]});
      } else {
        const int64_t y = x - (int64_t)S * NEL;
        const int nl = (int)(y % NVL), el = (int)((y / NVL) % NEL), f = (int)(y / NVL / NEL);
        c.code.push_back(DFS{f, nv, par.vlab[f], el, nl});
        c.vlab.push_back((uint8_t)nl);
      }
      if (opt.max_edges > 0 && (int)c.code.size() > opt.max_edges) return;
      if (!dfs_code_is_min(c.code.data(), (int)c.code.size())) return;
      if ((int)c.vlab.size() > kMaxPatV) {
        too_big = true;
        return;
      }
      keep[i] = 1;
    });
    if (too_big) throw Error(GM_E_UNSUPPORTED, "pattern with more than 32 vertices is frequent");
    canon_sec += tcan.sec();
    st.n_candidates += root ? 0 : F;

    std::vector<Pat> next;
    std::vector<unsigned long long> sel_keys;
    std::vector<int32_t> sel_vals;
    std::vector<uint64_t> nproj;
    int64_t child_rows = 0;
    for (int64_t i = 0; i < F; ++i) {
      if (!keep[i]) continue;
      sel_keys.push_back(fr[i].key);
      sel_vals.push_back(-((int32_t)next.size() + 1));
      nproj.push_back(fr[i].proj);
      child_rows += (int64_t)fr[i].proj;
      next.push_back(std::move(kids[i]));
    }
    const int64_t C = (int64_t)next.size();
    int CS = 4;
    if (C > 0) {
      // The embedding counts size the child level exactly: one write kernel
      // reserves each graph's rows from a cursor and fills them.
      Timer tdev2;
      int max_nv = 0;
      for (const Pat &c : next) max_nv = std::max(max_nv, (int)c.vlab.size());
      CS = (int)round4(max_nv);
      upload(idx_d, sel_keys, s);
      upload(val_d, sel_vals, s);
      hip_check(launch_scatter_i32(sup.as<int32_t>(), idx_d.as<uint64_t>(), val_d.as<int32_t>(), C, s), "scatter");
      pid_nxt->ensure(sizeof(int32_t) * std::max<int64_t>(child_rows, 1));
      vt_nxt->ensure((size_t)CS * std::max<int64_t>(child_rows, 1));
      rowcnt.ensure(sizeof(uint16_t) * std::max<int64_t>(root ? db.n_vertices : n_rows, 1));
      a.rowcnt = rowcnt.as<uint16_t>();
      a.ccap = child_rows;
      a.cgstart = gs_nxt->as<int64_t>();
      a.cgcount = gc_nxt->as<int64_t>();
      a.cpid = pid_nxt->as<int32_t>();
      a.cvt = vt_nxt->as<uint8_t>();
      a.CS = CS;
      hip_check(launch_write(a, G, root, s), "k_write");
      hip_check(hipMemcpyAsync(&cnt, counters.p, sizeof(cnt), hipMemcpyDeviceToHost, s), "download counters");
      hip_check(hipStreamSynchronize(s), "k_write");
      if (cnt.err || cnt.cursor != (unsigned long long)child_rows)
        throw Error(GM_E_DEVICE, "child rows written (" + std::to_string(cnt.cursor) +
                                     ") differ from the embedding count (" + std::to_string(child_rows) + ")");
      dev_sec += tdev2.sec();
      size_t live = db.bytes + sup.cap + sup_rep.cap + proj_rep.cap + pid_a.cap + pid_b.cap + vt_a.cap + vt_b.cap +
                    freq.cap + meta_d.cap + rowcnt.cap;
      peak = std::max(peak, live);
    }
    if (!opt.collect_proj) std::fill(nproj.begin(), nproj.end(), 0);
    if (root)
      st.e_root += child_rows;
    else
      st.e_child += child_rows;
    for (int64_t i = 0; i < C && !stopped; ++i) {
      st.n_patterns++;
      if (!report(next[i].job, to_original(next[i].code), next[i].sup, nproj[i])) stopped = true;
    }
    st.n_levels++;
    if (stopped) break;
    pats.swap(next);
    std::swap(gs_cur, gs_nxt);
    std::swap(gc_cur, gc_nxt);
    std::swap(pid_cur, pid_nxt);
    std::swap(vt_cur, vt_nxt);
    S = CS;
    n_rows = child_rows;
    ++level;
  }
  (void)n_rows;
  st.seconds_total = total.sec();
  st.seconds_device = dev_sec;
  st.seconds_canon = canon_sec;
  st.seconds_ext_kernel = ext_ms * 1e-3;
  st.ext_kernel_launches = ext_launches;
  st.peak_device_bytes = peak;
  if (stats) *stats = st;
  return !stopped;
}

}  // namespace gm
