// canonical.cpp -- minimum DFS code check for one candidate pattern.
//
// Decides, for a fully labelled DFS code C of n edges, whether C equals the
// code that the sequential gSpan's is_min grows greedily on C's own pattern
// graph (src/gspan/ismin.cpp:36-140, also src/globals/dfs_ismin.cpp:23-160):
//   1. the first tuple is the smallest 1-edge (fromlabel, elabel, tolabel)
//      over directed pattern edges with fromlabel <= tolabel (it is not
//      compared with C[0], as in ismin.cpp:56-62);
//   2. at each later position, with the embeddings of the prefix grown so
//      far, the next tuple is
//        a backward edge from the rightmost vertex to the rightmost-path
//        vertex nearest the root that any embedding can reach (smallest
//        elabel), subject to the order rule of get_backward
//        (src/globals/graph_types.cpp:389-413); else
//        a forward edge from the rightmost vertex (smallest (elabel,
//        tolabel), tolabel >= minlabel; graph_types.cpp:337-355); else
//        a forward edge from the deepest rightmost-path vertex that has
//        one, under the rule of get_forward_rmpath (graph_types.cpp:305-323);
//   3. the first position where that tuple differs from C answers "not
//      minimal"; reaching the end answers "minimal".
// Design: an embedding is a vertex map (min-code vertex -> pattern vertex)
// plus a bitmask of used pattern edges, both in flat arrays; no per-step
// maps or history rebuilds.  Patterns are capped at 128 edges / 64 vertices.
#include <algorithm>
#include <cstring>
#include <vector>

#include "gm_internal.hpp"
#include "gspan_amd.h"

namespace gm {
namespace {

constexpr int kMaxV = 64;
constexpr int kMaxE = 128;

struct PGraph {
  int nv = 0;
  int label[kMaxV];
  int deg[kMaxV];
  // adjacency in code order: (neighbour, elabel, edge id)
  int nbr[kMaxV][kMaxV];
  int el[kMaxV][kMaxV];
  int eid[kMaxV][kMaxV];
};

struct Emb {
  int8_t map[kMaxV];  // min-code vertex -> pattern vertex
  uint64_t used[kMaxE / 64];  // pattern edge ids already matched
  uint64_t vused;             // pattern vertices already matched
  bool has(int eid) const { return used[eid >> 6] >> (eid & 63) & 1; }
  void add(int eid) { used[eid >> 6] |= 1ull << (eid & 63); }
};

struct MinCode {
  int n = 0;
  DFS t[kMaxE];
  int rm[kMaxE];  // rightmost path, last forward edge first (dfs_code.cpp:75-90)
  int nrm = 0;
  void build_rm() {
    nrm = 0;
    int old_from = -1;
    for (int i = n - 1; i >= 0; --i)
      if (t[i].from < t[i].to && (nrm == 0 || old_from == t[i].to)) {
        rm[nrm++] = i;
        old_from = t[i].from;
      }
  }
};

// Tuple comparison as the sequential gSpan does it: fromlabel is only
// present on the first tuple, tolabel only on forward tuples.
inline bool same_tuple(const DFS &a, const DFS &b) {
  if (a.from != b.from || a.to != b.to || a.elabel != b.elabel) return false;
  if (a.from < a.to && a.tolabel != b.tolabel) return false;
  return true;
}

// The pattern graph of a DFS code (DFSCode::toGraph,
// src/globals/dfs_code.cpp:245-265): vertex labels from the tuples that
// carry them (-1 carries none), edges in code order.
void code_graph(const DFS *code, int n, PGraph &G) {
  if (n > kMaxE) throw Error(GM_E_UNSUPPORTED, "pattern larger than 128 edges");
  G.nv = 0;
  for (int i = 0; i < n; ++i) G.nv = std::max(G.nv, std::max(code[i].from, code[i].to) + 1);
  if (G.nv > kMaxV) throw Error(GM_E_UNSUPPORTED, "pattern larger than 64 vertices");
  std::fill(G.deg, G.deg + G.nv, 0);
  std::fill(G.label, G.label + G.nv, 0);
  for (int i = 0; i < n; ++i) {
    const DFS &d = code[i];
    if (d.from < 0 || d.to < 0 || d.from == d.to) throw Error(GM_E_ARG, "malformed DFS code");
    // labels as DFSCode::toGraph takes them (src/globals/dfs_code.cpp:245-265)
    if (d.fromlabel != -1) G.label[d.from] = d.fromlabel;
    if (d.tolabel != -1) G.label[d.to] = d.tolabel;
    int a = d.from, b = d.to;
    G.nbr[a][G.deg[a]] = b;
    G.el[a][G.deg[a]] = d.elabel;
    G.eid[a][G.deg[a]++] = i;
    G.nbr[b][G.deg[b]] = a;
    G.el[b][G.deg[b]] = d.elabel;
    G.eid[b][G.deg[b]++] = i;
  }
}

// Grow the minimum DFS code of G (n edges) greedily.  With want != nullptr
// stop at the first tuple that differs from want (returns false) or when
// the code is complete (true); with want == nullptr write the whole minimum
// code to out (returns true).
bool grow_min(const PGraph &G, int n, const DFS *want_code, DFS *out) {

  // Step 1: smallest root tuple over directed edges with fromlabel <= tolabel.
  int bfl = 1 << 30, bel = 1 << 30, btl = 1 << 30;
  for (int u = 0; u < G.nv; ++u)
    for (int k = 0; k < G.deg[u]; ++k) {
      int v = G.nbr[u][k];
      if (G.label[u] > G.label[v]) continue;
      int fl = G.label[u], e = G.el[u][k], tl = G.label[v];
      if (fl < bfl || (fl == bfl && (e < bel || (e == bel && tl < btl)))) {
        bfl = fl;
        bel = e;
        btl = tl;
      }
    }
  // embedding lists, reused by later calls on this thread (no allocation
  // once warm; grow_min does not recurse)
  thread_local std::vector<Emb> cur_store, nxt_store;
  std::vector<Emb> &cur = cur_store, &nxt = nxt_store;
  cur.clear();
  nxt.clear();
  for (int u = 0; u < G.nv; ++u)
    for (int k = 0; k < G.deg[u]; ++k) {
      int v = G.nbr[u][k];
      if (G.label[u] == bfl && G.el[u][k] == bel && G.label[v] == btl) {
        Emb e;
        e.map[0] = (int8_t)u;
        e.map[1] = (int8_t)v;
        std::fill(e.used, e.used + kMaxE / 64, 0ull);
        e.add(G.eid[u][k]);
        e.vused = (1ull << u) | (1ull << v);
        cur.push_back(e);
      }
    }
  MinCode mc;
  mc.t[0] = DFS{0, 1, bfl, bel, btl};
  mc.n = 1;

  while (mc.n < n) {
    mc.build_rm();
    const int minlabel = mc.t[0].fromlabel;
    const int maxtoc = mc.t[mc.rm[0]].to;
    nxt.clear();

    // Backward: rightmost-path edges rm[i], i from the root end down to 1.
    bool found = false;
    for (int i = mc.nrm - 1; i >= 1 && !found; --i) {
      const DFS &e1 = mc.t[mc.rm[i]];
      const DFS &e2 = mc.t[mc.rm[0]];
      int best = 1 << 30;
      for (const Emb &e : cur) {
        int r = e.map[e2.to], target = e.map[e1.from];
        int e1_tolabel = G.label[e.map[e1.to]], e2_tolabel = G.label[r];
        for (int k = 0; k < G.deg[r]; ++k) {
          if (e.has(G.eid[r][k])) continue;
          int lab = G.el[r][k];
          if (G.nbr[r][k] == target && (e1.elabel < lab || (e1.elabel == lab && e1_tolabel <= e2_tolabel))) {
            best = std::min(best, lab);
            found = true;
            break;  // the first qualifying edge, as get_backward returns
          }
        }
      }
      if (found) {
        DFS t{maxtoc, e1.from, -1, best, -1};
        if (want_code && !same_tuple(want_code[mc.n], t)) return false;
        for (const Emb &e : cur) {
          int r = e.map[e2.to], target = e.map[e1.from];
          int e1_tolabel = G.label[e.map[e1.to]], e2_tolabel = G.label[r];
          for (int k = 0; k < G.deg[r]; ++k) {
            if (e.has(G.eid[r][k])) continue;
            int lab = G.el[r][k];
            if (G.nbr[r][k] == target && (e1.elabel < lab || (e1.elabel == lab && e1_tolabel <= e2_tolabel))) {
              if (lab == best) {
                Emb x = e;
                x.add(G.eid[r][k]);
                nxt.push_back(x);
              }
              break;
            }
          }
        }
        mc.t[mc.n++] = t;
      }
    }
    if (found) {
      cur.swap(nxt);
      continue;
    }

    // Forward from the rightmost vertex (pure forward).
    int bel2 = 1 << 30, btl2 = 1 << 30, newfrom = -1;
    for (const Emb &e : cur) {
      int r = e.map[maxtoc];
      for (int k = 0; k < G.deg[r]; ++k) {
        int w = G.nbr[r][k];
        if (minlabel > G.label[w] || (e.vused >> w & 1)) continue;
        int lab = G.el[r][k], tl = G.label[w];
        if (lab < bel2 || (lab == bel2 && tl < btl2)) {
          bel2 = lab;
          btl2 = tl;
        }
        newfrom = maxtoc;
      }
    }
    // Forward from the rightmost path, deepest vertex first.
    for (int i = 0; newfrom < 0 && i < mc.nrm; ++i) {
      const DFS &ed = mc.t[mc.rm[i]];
      for (const Emb &e : cur) {
        int u = e.map[ed.from], x = e.map[ed.to];
        int tolabel = G.label[x];
        for (int k = 0; k < G.deg[u]; ++k) {
          int w = G.nbr[u][k];
          int tl = G.label[w];
          if (w == x || minlabel > tl || (e.vused >> w & 1)) continue;
          int lab = G.el[u][k];
          if (ed.elabel < lab || (ed.elabel == lab && tolabel <= tl)) {
            if (lab < bel2 || (lab == bel2 && tl < btl2)) {
              bel2 = lab;
              btl2 = tl;
            }
            newfrom = ed.from;
          }
        }
      }
    }
    if (newfrom < 0) {
      if (want_code) return true;  // no extension: the whole pattern is covered
      throw Error(GM_E_ARG, "pattern is not connected");
    }
    DFS t{newfrom, maxtoc + 1, -1, bel2, btl2};
    if (want_code && !same_tuple(want_code[mc.n], t)) return false;
    for (const Emb &e : cur) {
      int u = e.map[newfrom];
      int x = newfrom == maxtoc ? -1 : -2;
      int tolabel = 0;
      const DFS *ed = nullptr;
      if (newfrom != maxtoc) {
        for (int i = 0; i < mc.nrm; ++i)
          if (mc.t[mc.rm[i]].from == newfrom) ed = &mc.t[mc.rm[i]];
        x = e.map[ed->to];
        tolabel = G.label[x];
      }
      for (int k = 0; k < G.deg[u]; ++k) {
        int w = G.nbr[u][k];
        int tl = G.label[w];
        if (w == x || minlabel > tl || (e.vused >> w & 1)) continue;
        int lab = G.el[u][k];
        if (ed && !(ed->elabel < lab || (ed->elabel == lab && tolabel <= tl))) continue;
        if (lab == bel2 && tl == btl2) {
          Emb y = e;
          y.map[maxtoc + 1] = (int8_t)w;
          y.add(G.eid[u][k]);
          y.vused |= 1ull << w;
          nxt.push_back(y);
        }
      }
    }
    mc.t[mc.n++] = t;
    cur.swap(nxt);
  }
  if (out) {
    // full labels: vertex labels of the minimum code's own numbering
    int vl[kMaxV];
    vl[0] = mc.t[0].fromlabel;
    vl[1] = mc.t[0].tolabel;
    for (int i = 1; i < mc.n; ++i)
      if (mc.t[i].from < mc.t[i].to) vl[mc.t[i].to] = mc.t[i].tolabel;
    for (int i = 0; i < mc.n; ++i) {
      out[i] = mc.t[i];
      out[i].fromlabel = vl[mc.t[i].from];
      out[i].tolabel = vl[mc.t[i].to];
    }
  }
  return true;
}

}  // namespace

bool dfs_code_is_min(const DFS *code, int n) {
  if (n <= 1) return true;
  PGraph G;
  code_graph(code, n, G);
  return grow_min(G, n, code, nullptr);
}

void min_dfs_code(const DFS *code, int n, std::vector<DFS> &out) {
  out.clear();
  if (n <= 0) return;
  for (int i = 0; i < n; ++i)
    if (code[i].from < code[i].to && (code[i].tolabel == -1 || (i == 0 && code[i].fromlabel == -1)))
      throw Error(GM_E_ARG, "query code must carry every vertex label");
  PGraph G;
  code_graph(code, n, G);
  out.resize(n);
  grow_min(G, n, nullptr, out.data());
}

}  // namespace gm
