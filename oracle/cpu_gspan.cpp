// cpu_gspan.cpp -- CPU restatement of the reference's sequential gSpan.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it may be built,
// linked or run only by tests/, by __graft_entry__.smoke() and by bench.py's
// cpu_baseline leg, and only as the checker.  The product library
// (gpu_graph_mining_amd/) never links or calls it.
//
// It restates, in code of its own, the algorithm of
//   src/gspan/gspan.cpp:88-102   gSpan::support        (runs of equal graph ids)
//   src/gspan/gspan.cpp:172-274  gSpan::project        (recursive growth)
//   src/gspan/gspan.cpp:281-347  gSpan::run_intern     (1-edge roots)
//   src/gspan/ismin.cpp:36-140   gSpan::is_min / project_is_min
//   src/globals/graph_types.cpp:34-60    Graph::buildEdge (edge ids)
//   src/globals/graph_types.cpp:66-173   Graph::read / read_fsg
//   src/globals/graph_types.cpp:258-277  History::build
//   src/globals/graph_types.cpp:305-413  get_forward_rmpath / _pure / _root, get_backward
//   src/globals/dfs_code.cpp:75-90       DFSCode::buildRMPath
//   src/globals/dfs_code.cpp:245-265     DFSCode::toGraph
//   src/globals/graph_output.cpp:90-95   pattern line = DFSCode::to_string(false)
// including the std::map iteration orders (graph_types.hpp:183-189) that fix
// the emission order, so its output file is byte-identical to the reference
// CLI's (src/gspan/main.cpp:21-96).
//
// Instrumentation (not in the reference): per reported pattern |Proj(P)|,
// E_cand (every Projected::push inside project), and the algorithmic byte
// count B_alg of the GPU extension/support/child stages:
//   B_ext(P)   = sum_rows [ 8*k + sum_{v in RM}(8 + 12*deg(v)) ] + 12*E_cand(P)
//   B_sup(P)   = 12*E_cand(P)
//   B_child(P) = sum over reported children c of 20*|Proj(c)|
// where k = vertex count of P and RM = the rightmost-path vertices of P.
//
// Parity pins (tests/): byte-identical output on testdata/Chemical_340 against
// the reference build in oracle/_ref (and the md5 values measured from it),
// the reference's is-min known-answer codes (src/globals/tests/
// dfs_ismin_tests.cpp:48-61), the 18-graph fixtures of
// src/globals/test_support.cpp:10-153, and brute-force subgraph counting on
// tiny graphs.

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace orc {

struct Edge {
  int from = 0, to = 0, elabel = 0;
  unsigned id = 0;
};

struct Vertex {
  int label = 0;
  std::vector<Edge> edge;
  void push(int from, int to, int elabel) {
    Edge e;
    e.from = from;
    e.to = to;
    e.elabel = elabel;
    edge.push_back(e);
  }
};

struct Graph : public std::vector<Vertex> {
  unsigned edge_size_ = 0;
  // Edge ids: one id per distinct undirected (min,max,elabel) triple, in
  // first-seen order while walking vertices and their adjacency lists
  // (graph_types.cpp:34-60).
  void build_edge() {
    std::map<std::tuple<int, int, int>, unsigned> ids;
    unsigned next = 0;
    for (int from = 0; from < (int)size(); ++from) {
      for (Edge &e : (*this)[from].edge) {
        std::tuple<int, int, int> key =
            from <= e.to ? std::make_tuple(from, e.to, e.elabel) : std::make_tuple(e.to, from, e.elabel);
        auto it = ids.find(key);
        if (it == ids.end()) {
          e.id = next;
          ids[key] = next++;
        } else {
          e.id = it->second;
        }
      }
    }
    edge_size_ = next;
  }
};

struct DFS {
  int from, to, fromlabel, elabel, tolabel;
  bool operator==(const DFS &o) const {
    return from == o.from && to == o.to && fromlabel == o.fromlabel && elabel == o.elabel && tolabel == o.tolabel;
  }
  bool operator!=(const DFS &o) const { return !(*this == o); }
};

struct DFSCode : public std::vector<DFS> {
  std::vector<int> rmpath;
  void push(int from, int to, int fl, int el, int tl) { push_back(DFS{from, to, fl, el, tl}); }
  void pop() { pop_back(); }
  // dfs_code.cpp:75-90: indexes of the forward edges on the rightmost path,
  // last edge first.
  const std::vector<int> &build_rmpath() {
    rmpath.clear();
    int old_from = -1;
    for (int i = (int)size() - 1; i >= 0; --i) {
      const DFS &d = (*this)[i];
      if (d.from < d.to && (rmpath.empty() || old_from == d.to)) {
        rmpath.push_back(i);
        old_from = d.from;
      }
    }
    return rmpath;
  }
  // dfs_code.cpp:245-265 (undirected).
  void to_graph(Graph &g) const {
    g.clear();
    for (const DFS &d : *this) {
      g.resize(std::max(d.from, d.to) + 1);
      if (d.fromlabel != -1) g[d.from].label = d.fromlabel;
      if (d.tolabel != -1) g[d.to].label = d.tolabel;
      g[d.from].push(d.from, d.to, d.elabel);
      g[d.to].push(d.to, d.from, d.elabel);
    }
    g.build_edge();
  }
  std::string to_string() const {
    std::ostringstream ss;
    for (size_t i = 0; i < size(); ++i) {
      const DFS &d = (*this)[i];
      if (i) ss << ';';
      ss << '(' << d.from << ' ' << d.to << ' ' << d.fromlabel << ' ' << d.elabel << ' ' << d.tolabel << ')';
    }
    return ss.str();
  }
};

struct PDFS {
  unsigned id = 0;
  Edge *edge = nullptr;
  PDFS *prev = nullptr;
};

struct Projected : public std::vector<PDFS> {
  void push(unsigned id, Edge *e, PDFS *prev) {
    PDFS d;
    d.id = id;
    d.edge = e;
    d.prev = prev;
    push_back(d);
  }
};

typedef std::map<int, std::map<int, std::map<int, Projected>>> ProjMap3;
typedef std::map<int, std::map<int, Projected>> ProjMap2;
typedef std::map<int, Projected> ProjMap1;
typedef std::vector<Edge *> EdgeList;

// graph_types.cpp:258-277: the edges of one embedding in code order plus
// used-edge / used-vertex flags.
struct History : public std::vector<Edge *> {
  std::vector<char> edge_used, vertex_used;
  void build(const Graph &g, PDFS *e) {
    clear();
    edge_used.assign(g.edge_size_, 0);
    vertex_used.assign(g.size(), 0);
    for (PDFS *p = e; p; p = p->prev) {
      push_back(p->edge);
      edge_used[p->edge->id] = vertex_used[p->edge->from] = vertex_used[p->edge->to] = 1;
    }
    std::reverse(begin(), end());
  }
  bool has_edge(unsigned id) const { return edge_used[id] != 0; }
  bool has_vertex(int v) const { return vertex_used[v] != 0; }
};

// graph_types.cpp:305-323
static bool get_forward_rmpath(const Graph &g, Edge *e, int minlabel, History &h, EdgeList &out) {
  out.clear();
  int tolabel = g[e->to].label;
  for (const Edge &it : g[e->from].edge) {
    int tolabel2 = g[it.to].label;
    if (e->to == it.to || minlabel > tolabel2 || h.has_vertex(it.to)) continue;
    if (e->elabel < it.elabel || (e->elabel == it.elabel && tolabel <= tolabel2))
      out.push_back(const_cast<Edge *>(&it));
  }
  return !out.empty();
}

// graph_types.cpp:337-355
static bool get_forward_pure(const Graph &g, Edge *e, int minlabel, History &h, EdgeList &out) {
  out.clear();
  for (const Edge &it : g[e->to].edge) {
    if (minlabel > g[it.to].label || h.has_vertex(it.to)) continue;
    out.push_back(const_cast<Edge *>(&it));
  }
  return !out.empty();
}

// graph_types.cpp:365-375
static bool get_forward_root(const Graph &g, const Vertex &v, EdgeList &out) {
  out.clear();
  for (const Edge &it : v.edge)
    if (v.label <= g[it.to].label) out.push_back(const_cast<Edge *>(&it));
  return !out.empty();
}

// graph_types.cpp:389-413
static Edge *get_backward(const Graph &g, Edge *e1, Edge *e2, History &h) {
  if (e1 == e2) return nullptr;
  for (const Edge &it : g[e2->to].edge) {
    if (h.has_edge(it.id)) continue;
    if (it.to == e1->from &&
        (e1->elabel < it.elabel || (e1->elabel == it.elabel && g[e1->to].label <= g[e2->to].label)))
      return const_cast<Edge *>(&it);
  }
  return nullptr;
}

struct Stats {
  uint64_t n_patterns = 0;   // reported codes (all sizes)
  uint64_t e_child = 0;      // sum |Proj(P)| over reported P with >= 2 edges
  uint64_t e_cand = 0;       // Projected pushes inside project()
  uint64_t e_root = 0;       // sum |Proj(P)| over reported 1-edge P
  uint64_t b_alg = 0;        // algorithmic bytes, formula in the header
  uint64_t b_ext = 0;        // extension part of b_alg
  uint64_t b_sup = 0;        // support part of b_alg
  uint64_t b_child = 0;      // child part of b_alg
  uint64_t n_expanded = 0;   // patterns whose extensions were enumerated
  uint64_t max_edges = 0;
};

typedef void (*pattern_cb)(void *user, const int32_t *dfs5, int32_t n_edges, uint32_t support, uint64_t n_proj);

class Miner {
 public:
  std::vector<Graph> db;
  unsigned minsup = 1;
  unsigned max_edges = 0;  // 0: unbounded (the reference's maxpat bound, gspan.cpp:191-198, is commented out)
  pattern_cb cb = nullptr;
  void *user = nullptr;
  Stats st;
  bool want_balg = true;
  // Sampling for a bounded CPU timing: only the 1-edge roots whose ordinal
  // (in the order below) is a multiple of root_stride are grown (1 = all).
  unsigned root_stride = 1;

  void run() {
    // gspan.cpp:321-332: every directed edge with fromlabel <= tolabel.
    EdgeList edges;
    ProjMap3 root;
    for (unsigned id = 0; id < db.size(); ++id) {
      Graph &g = db[id];
      for (unsigned from = 0; from < g.size(); ++from)
        if (get_forward_root(g, g[from], edges))
          for (Edge *e : edges) root[g[from].label][e->elabel][g[e->to].label].push(id, e, nullptr);
    }
    unsigned ordinal = 0;
    for (auto &fl : root)
      for (auto &el : fl.second)
        for (auto &tl : el.second) {
          if (ordinal++ % root_stride != 0) continue;
          code_.push(0, 1, fl.first, el.first, tl.first);
          project(tl.second);
          code_.pop();
        }
  }

  // Standalone canonical check of an arbitrary code (ismin.cpp:36-140).
  bool is_min_of(const DFSCode &c) {
    code_ = c;
    bool r = is_min();
    code_.clear();
    return r;
  }

 private:
  DFSCode code_, min_code_;
  Graph min_graph_;

  // gspan.cpp:88-102: number of runs of equal consecutive graph ids.
  static unsigned support(const Projected &p) {
    unsigned oid = 0xffffffffu, size = 0;
    for (const PDFS &d : p) {
      if (oid != d.id) ++size;
      oid = d.id;
    }
    return size;
  }

  void report(const Projected &p, unsigned sup) {
    st.n_patterns++;
    if (code_.size() >= 2)
      st.e_child += p.size();
    else
      st.e_root += p.size();
    st.max_edges = std::max<uint64_t>(st.max_edges, code_.size());
    if (cb) {
      std::vector<int32_t> flat(code_.size() * 5);
      for (size_t i = 0; i < code_.size(); ++i) {
        flat[i * 5 + 0] = code_[i].from;
        flat[i * 5 + 1] = code_[i].to;
        flat[i * 5 + 2] = code_[i].fromlabel;
        flat[i * 5 + 3] = code_[i].elabel;
        flat[i * 5 + 4] = code_[i].tolabel;
      }
      cb(user, flat.data(), (int32_t)code_.size(), sup, (uint64_t)p.size());
    }
  }

  // gspan.cpp:172-274
  void project(Projected &projected) {
    if (max_edges && code_.size() > max_edges) return;
    unsigned sup = support(projected);
    if (sup < minsup) return;
    if (!is_min()) return;
    report(projected, sup);

    const std::vector<int> rmpath = code_.build_rmpath();
    int minlabel = code_[0].fromlabel;
    int maxtoc = code_[rmpath[0]].to;

    ProjMap3 new_fwd_root;
    ProjMap2 new_bck_root;
    EdgeList edges;
    History history;
    uint64_t cand = 0;
    uint64_t row_bytes = 0;
    // vertex count of the pattern = maxtoc + 1 (forward edges number vertices)
    int k = 0;
    for (const DFS &d : code_) k = std::max(k, std::max(d.from, d.to) + 1);

    for (unsigned n = 0; n < projected.size(); ++n) {
      unsigned id = projected[n].id;
      PDFS *cur = &projected[n];
      Graph &g = db[id];
      history.build(g, cur);

      if (want_balg) {
        row_bytes += 8ull * k;
        // RM vertices: the from of every rightmost-path edge plus maxtoc.
        for (size_t i = 0; i < rmpath.size(); ++i)
          row_bytes += 8 + 12ull * g[history[rmpath[i]]->from].edge.size();
        row_bytes += 8 + 12ull * g[history[rmpath[0]]->to].edge.size();
      }

      // backward (gspan.cpp:220-225)
      for (int i = (int)rmpath.size() - 1; i >= 1; --i) {
        Edge *e = get_backward(g, history[rmpath[i]], history[rmpath[0]], history);
        if (e) {
          new_bck_root[code_[rmpath[i]].from][e->elabel].push(id, e, cur);
          ++cand;
        }
      }
      // pure forward (gspan.cpp:233-237)
      if (get_forward_pure(g, history[rmpath[0]], minlabel, history, edges))
        for (Edge *e : edges) {
          new_fwd_root[maxtoc][e->elabel][g[e->to].label].push(id, e, cur);
          ++cand;
        }
      // rightmost-path forward (gspan.cpp:240-246)
      for (int i = 0; i < (int)rmpath.size(); ++i)
        if (get_forward_rmpath(g, history[rmpath[i]], minlabel, history, edges))
          for (Edge *e : edges) {
            new_fwd_root[code_[rmpath[i]].from][e->elabel][g[e->to].label].push(id, e, cur);
            ++cand;
          }
    }
    st.e_cand += cand;
    st.n_expanded++;
    if (want_balg) {
      st.b_ext += row_bytes + 12ull * cand;  // rows + candidate writes
      st.b_sup += 12ull * cand;              // candidate reads
      st.b_alg += row_bytes + 24ull * cand;
    }

    // backward children: to ascending, elabel ascending (gspan.cpp:251-257)
    for (auto &to : new_bck_root)
      for (auto &el : to.second) {
        code_.push(maxtoc, to.first, -1, el.first, -1);
        uint64_t before = st.n_patterns;
        project(el.second);
        if (want_balg && st.n_patterns > before) charge_child(el.second);
        code_.pop();
      }
    // forward children: from descending, elabel asc, tolabel asc (260-271)
    for (auto from = new_fwd_root.rbegin(); from != new_fwd_root.rend(); ++from)
      for (auto &el : from->second)
        for (auto &tl : el.second) {
          code_.push(from->first, maxtoc + 1, -1, el.first, tl.first);
          uint64_t before = st.n_patterns;
          project(tl.second);
          if (want_balg && st.n_patterns > before) charge_child(tl.second);
          code_.pop();
        }
  }

  // A child is materialised exactly when it is reported (frequent and
  // minimal); charge 20 bytes per child row (read + write of the row).
  void charge_child(const Projected &p) {
    st.b_alg += 20ull * p.size();
    st.b_child += 20ull * p.size();
  }

  // ismin.cpp:36-63
  bool is_min() {
    if (code_.size() == 1) return true;
    code_.to_graph(min_graph_);
    min_code_.clear();
    ProjMap3 root;
    EdgeList edges;
    for (unsigned from = 0; from < min_graph_.size(); ++from)
      if (get_forward_root(min_graph_, min_graph_[from], edges))
        for (Edge *e : edges)
          root[min_graph_[from].label][e->elabel][min_graph_[e->to].label].push(0, e, nullptr);
    auto fl = root.begin();
    auto el = fl->second.begin();
    auto tl = el->second.begin();
    min_code_.push(0, 1, fl->first, el->first, tl->first);
    return project_is_min(tl->second);
  }

  // ismin.cpp:65-140
  bool project_is_min(Projected &projected) {
    const std::vector<int> rmpath = min_code_.build_rmpath();
    int minlabel = min_code_[0].fromlabel;
    int maxtoc = min_code_[rmpath[0]].to;
    History history;
    {
      ProjMap1 root;
      bool flg = false;
      int newto = 0;
      for (int i = (int)rmpath.size() - 1; !flg && i >= 1; --i) {
        for (unsigned n = 0; n < projected.size(); ++n) {
          PDFS *cur = &projected[n];
          history.build(min_graph_, cur);
          Edge *e = get_backward(min_graph_, history[rmpath[i]], history[rmpath[0]], history);
          if (e) {
            root[e->elabel].push(0, e, cur);
            newto = min_code_[rmpath[i]].from;
            flg = true;
          }
        }
      }
      if (flg) {
        auto el = root.begin();
        min_code_.push(maxtoc, newto, -1, el->first, -1);
        if (code_[min_code_.size() - 1] != min_code_[min_code_.size() - 1]) return false;
        return project_is_min(el->second);
      }
    }
    {
      bool flg = false;
      int newfrom = 0;
      ProjMap2 root;
      EdgeList edges;
      for (unsigned n = 0; n < projected.size(); ++n) {
        PDFS *cur = &projected[n];
        history.build(min_graph_, cur);
        if (get_forward_pure(min_graph_, history[rmpath[0]], minlabel, history, edges)) {
          flg = true;
          newfrom = maxtoc;
          for (Edge *e : edges) root[e->elabel][min_graph_[e->to].label].push(0, e, cur);
        }
      }
      for (int i = 0; !flg && i < (int)rmpath.size(); ++i) {
        for (unsigned n = 0; n < projected.size(); ++n) {
          PDFS *cur = &projected[n];
          history.build(min_graph_, cur);
          if (get_forward_rmpath(min_graph_, history[rmpath[i]], minlabel, history, edges)) {
            flg = true;
            newfrom = min_code_[rmpath[i]].from;
            for (Edge *e : edges) root[e->elabel][min_graph_[e->to].label].push(0, e, cur);
          }
        }
      }
      if (flg) {
        auto el = root.begin();
        auto tl = el->second.begin();
        min_code_.push(newfrom, maxtoc + 1, -1, el->first, tl->first);
        if (code_[min_code_.size() - 1] != min_code_[min_code_.size() - 1]) return false;
        return project_is_min(tl->second);
      }
    }
    return true;
  }
};

// Line splitting on single spaces, as utils::split with its default
// delimiter (src/globals/utils.hpp:33-35, utils.cpp:192-210).
static void split_spaces(const std::string &s, std::vector<std::string> &out) {
  out.clear();
  size_t last = s.find_first_not_of(' ', 0);
  size_t pos = s.find_first_of(' ', last);
  while (pos != std::string::npos || last != std::string::npos) {
    out.push_back(s.substr(last, pos - last));
    last = s.find_first_not_of(' ', pos);
    pos = s.find_first_of(' ', last);
  }
}

// graph_types.cpp:66-113 (txt) and 115-173 (fsg: labels are strings mapped
// to dense ids in first-seen order, per graph; edges on 'u' lines).
static void read_db(const std::string &path, bool fsg, std::vector<Graph> &db) {
  std::ifstream in(path.c_str());
  if (!in) throw std::runtime_error("cannot open " + path);
  std::string line;
  std::vector<std::string> tok;
  Graph g;
  std::map<std::string, int> vl, el;
  auto flush = [&]() {
    if (!g.empty()) {
      g.build_edge();
      db.push_back(g);
    }
    g.clear();
    vl.clear();
    el.clear();
  };
  while (std::getline(in, line)) {
    split_spaces(line, tok);
    if (tok.empty()) continue;
    if (tok[0] == "t") {
      flush();
    } else if (tok[0] == "v" && tok.size() >= 3) {
      unsigned id = (unsigned)atoi(tok[1].c_str());
      g.resize(id + 1);
      if (fsg) {
        if (!vl.count(tok[2])) {
          int n = (int)vl.size();
          vl[tok[2]] = n;
        }
        g[id].label = vl[tok[2]];
      } else {
        g[id].label = atoi(tok[2].c_str());
      }
    } else if (tok[0] == (fsg ? "u" : "e") && tok.size() >= 4) {
      int from = atoi(tok[1].c_str()), to = atoi(tok[2].c_str());
      int lab;
      if (fsg) {
        if (!el.count(tok[3])) {
          int n = (int)el.size();
          el[tok[3]] = n;
        }
        lab = el[tok[3]];
      } else {
        lab = atoi(tok[3].c_str());
      }
      if ((int)g.size() <= from || (int)g.size() <= to)
        throw std::runtime_error("Format Error:  define vertex lists before edges");
      g[from].push(from, to, lab);
      g[to].push(to, from, lab);
    }
  }
  flush();
}

static void build_db_from_arrays(int32_t n_graphs, const int32_t *graph_vbegin, const int32_t *vlabels,
                                 const int32_t *graph_ebegin, const int32_t *efrom, const int32_t *eto,
                                 const int32_t *elabel, std::vector<Graph> &db) {
  db.clear();
  db.resize(n_graphs);
  for (int32_t gi = 0; gi < n_graphs; ++gi) {
    Graph &g = db[gi];
    int32_t nv = graph_vbegin[gi + 1] - graph_vbegin[gi];
    g.resize(nv);
    for (int32_t v = 0; v < nv; ++v) g[v].label = vlabels[graph_vbegin[gi] + v];
    for (int32_t e = graph_ebegin[gi]; e < graph_ebegin[gi + 1]; ++e) {
      g[efrom[e]].push(efrom[e], eto[e], elabel[e]);
      g[eto[e]].push(eto[e], efrom[e], elabel[e]);
    }
    g.build_edge();
  }
}

}  // namespace orc

extern "C" {

struct orc_stats {
  uint64_t n_patterns, e_child, e_cand, e_root, b_alg, n_expanded, max_edges;
  double seconds;
  uint64_t b_ext, b_sup, b_child;
};

static void fill_stats(const orc::Stats &s, double secs, orc_stats *out) {
  if (!out) return;
  out->n_patterns = s.n_patterns;
  out->e_child = s.e_child;
  out->e_cand = s.e_cand;
  out->e_root = s.e_root;
  out->b_alg = s.b_alg;
  out->n_expanded = s.n_expanded;
  out->max_edges = s.max_edges;
  out->seconds = secs;
  out->b_ext = s.b_ext;
  out->b_sup = s.b_sup;
  out->b_child = s.b_child;
}

// Mine a database given as per-graph vertex and edge lists (edges in input
// order, local vertex ids).  Returns 0 on success, -1 on bad arguments.
int orc_mine(int32_t n_graphs, const int32_t *graph_vbegin, const int32_t *vlabels, const int32_t *graph_ebegin,
             const int32_t *efrom, const int32_t *eto, const int32_t *elabel, int32_t minsup,
             orc::pattern_cb cb, void *user, orc_stats *stats) {
  if (n_graphs < 0 || minsup < 1) return -1;
  orc::Miner m;
  orc::build_db_from_arrays(n_graphs, graph_vbegin, vlabels, graph_ebegin, efrom, eto, elabel, m.db);
  m.minsup = (unsigned)minsup;
  m.cb = cb;
  m.user = user;
  auto t0 = std::chrono::steady_clock::now();
  m.run();
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(m.st, secs, stats);
  return 0;
}

// orc_mine with patterns capped at max_edges edges (0 = unbounded).
int orc_mine_capped(int32_t n_graphs, const int32_t *graph_vbegin, const int32_t *vlabels, const int32_t *graph_ebegin,
                    const int32_t *efrom, const int32_t *eto, const int32_t *elabel, int32_t minsup, int32_t max_edges,
                    orc::pattern_cb cb, void *user, orc_stats *stats) {
  if (n_graphs < 0 || minsup < 1 || max_edges < 0) return -1;
  orc::Miner m;
  orc::build_db_from_arrays(n_graphs, graph_vbegin, vlabels, graph_ebegin, efrom, eto, elabel, m.db);
  m.minsup = (unsigned)minsup;
  m.max_edges = (unsigned)max_edges;
  m.cb = cb;
  m.user = user;
  auto t0 = std::chrono::steady_clock::now();
  m.run();
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(m.st, secs, stats);
  return 0;
}

// orc_mine over a sample of the pattern tree: only every root_stride-th
// 1-edge root (in the miner's root order) is grown, each subtree in full.
// Used to time the CPU baseline on a bounded part of a large workload.
int orc_mine_sampled(int32_t n_graphs, const int32_t *graph_vbegin, const int32_t *vlabels,
                     const int32_t *graph_ebegin, const int32_t *efrom, const int32_t *eto, const int32_t *elabel,
                     int32_t minsup, int32_t root_stride, orc::pattern_cb cb, void *user, orc_stats *stats) {
  if (n_graphs < 0 || minsup < 1 || root_stride < 1) return -1;
  orc::Miner m;
  orc::build_db_from_arrays(n_graphs, graph_vbegin, vlabels, graph_ebegin, efrom, eto, elabel, m.db);
  m.minsup = (unsigned)minsup;
  m.root_stride = (unsigned)root_stride;
  m.cb = cb;
  m.user = user;
  auto t0 = std::chrono::steady_clock::now();
  m.run();
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(m.st, secs, stats);
  return 0;
}

// Mine a txt ("txt") or fsg ("fsg") file.  Returns 0, -1 bad args, -2 I/O.
int orc_mine_file(const char *path, const char *ftype, int32_t minsup, orc::pattern_cb cb, void *user,
                  orc_stats *stats) {
  if (!path || !ftype || minsup < 1) return -1;
  orc::Miner m;
  try {
    std::string ft(ftype);
    if (!ft.empty() && ft[0] == '-') ft = ft.substr(1);
    if (ft != "txt" && ft != "fsg") return -1;
    orc::read_db(path, ft == "fsg", m.db);
  } catch (const std::exception &) {
    return -2;
  }
  m.minsup = (unsigned)minsup;
  m.cb = cb;
  m.user = user;
  auto t0 = std::chrono::steady_clock::now();
  m.run();
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(m.st, secs, stats);
  return 0;
}

// Canonical (minimum DFS code) check of a code given as n rows of
// (from, to, fromlabel, elabel, tolabel), in the reference's -1 convention
// (fromlabel -1 on edges after the first, tolabel -1 on backward edges) or
// fully labelled.  Returns 1 if minimal, 0 if not, -1 on bad input.
int orc_is_min(const int32_t *dfs5, int32_t n) {
  if (!dfs5 || n < 1) return -1;
  orc::DFSCode c;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t *d = dfs5 + 5 * i;
    // The check compares against codes it builds in the -1 convention.
    int fl = i == 0 ? d[2] : -1;
    int tl = d[0] < d[1] ? d[4] : -1;
    c.push(d[0], d[1], fl, d[3], tl);
  }
  orc::Miner m;
  return m.is_min_of(c) ? 1 : 0;
}

// Distinct-graph support of extension records: records r with key[r] and
// graph[r]; out[k] = number of distinct graph ids among records of key k
// (k in [0, n_keys)).  This is the support semantics of
// src/gspan_cuda/compute_support.cpp:885-959 restated on the CPU.
int orc_support_from_records(const int32_t *key, const int32_t *graph, int64_t n, int32_t n_keys, int32_t *out) {
  if (n < 0 || n_keys < 0) return -1;
  std::vector<std::vector<int32_t>> per(n_keys);
  for (int64_t r = 0; r < n; ++r) {
    if (key[r] < 0 || key[r] >= n_keys) return -1;
    per[key[r]].push_back(graph[r]);
  }
  for (int32_t k = 0; k < n_keys; ++k) {
    std::sort(per[k].begin(), per[k].end());
    out[k] = (int32_t)(std::unique(per[k].begin(), per[k].end()) - per[k].begin());
  }
  return 0;
}

}  // extern "C"

#ifdef ORC_MAIN
// CLI: cpu_gspan <txt|fsg> <file> <support-int-abs> [<out>|-]
// Writes one pattern per line, "<code to_string(false)>\t<support>", or only
// the code when GM_ORC_CODE_ONLY=1 (the reference file sink's format).
static FILE *g_out = nullptr;
static bool g_code_only = false;
static void cli_cb(void *, const int32_t *d, int32_t n, uint32_t sup, uint64_t) {
  if (!g_out) return;
  for (int32_t i = 0; i < n; ++i)
    fprintf(g_out, "%s(%d %d %d %d %d)", i ? ";" : "", d[5 * i], d[5 * i + 1], d[5 * i + 2], d[5 * i + 3],
            d[5 * i + 4]);
  if (g_code_only)
    fprintf(g_out, "\n");
  else
    fprintf(g_out, "\t%u\n", sup);
}
int main(int argc, char **argv) {
  if (argc != 4 && argc != 5) {
    fprintf(stderr, "usage: %s <filetype> <filename> <support-int-abs> [<output-filename>]\n", argv[0]);
    return 1;
  }
  int minsup = atoi(argv[3]);
  if (minsup < 1) {
    fprintf(stderr, "error: absolute_min_support < 1\n");
    return 3;
  }
  const char *co = getenv("GM_ORC_CODE_ONLY");
  g_code_only = co && co[0] == '1';
  if (argc == 5 && strcmp(argv[4], "-") != 0) g_out = fopen(argv[4], "w");
  orc_stats st;
  int rc = orc_mine_file(argv[2], argv[1], minsup, cli_cb, nullptr, &st);
  if (g_out) fclose(g_out);
  if (rc) {
    fprintf(stderr, "error %d\n", rc);
    return 1;
  }
  printf("[VAL] [total_time=%.6f]\n[VAL] [total_frequent_graphs=%llu]\n", st.seconds,
         (unsigned long long)st.n_patterns);
  printf("[VAL] [e_child=%llu] [e_cand=%llu] [b_alg=%llu]\n", (unsigned long long)st.e_child,
         (unsigned long long)st.e_cand, (unsigned long long)st.b_alg);
  return 0;
}
#endif
