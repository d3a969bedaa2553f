// generators.cpp -- seeded synthetic graph databases.
//
// Every generator draws from one SplitMix64 stream (Steele, Lea & Flood,
// OOPSLA 2014) seeded with the caller's seed, so a (generator, parameters,
// seed) triple always yields the same database on every machine.
// Integer and real variates:
//   uniform int in [0,n)  : 64-bit multiply-shift of a raw draw (Lemire);
//   uniform real in [0,1) : top 53 bits of a raw draw;
//   Poisson(mu)           : Knuth's product-of-uniforms for mu <= 30, else
//                           the normal approximation rounded to an integer;
//   Exp(1)                : -log(1 - U).
// No generator in the reference is used: its FSG generator ships only as a
// prebuilt binary, and its GSL wrapper needs a library this image lacks.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <set>
#include <utility>

#include "gm_internal.hpp"
#include "gspan_amd.h"

namespace gm {
namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [0, n)
  int32_t below(int32_t n) { return (int32_t)(((unsigned __int128)next() * (uint64_t)n) >> 64); }
  int32_t range(int32_t lo, int32_t hi) { return lo + below(hi - lo + 1); }  // inclusive
  double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  int32_t poisson(double mu) {
    if (mu <= 0) return 0;
    if (mu <= 30.0) {
      double L = std::exp(-mu), p = 1.0;
      int32_t k = 0;
      do {
        ++k;
        p *= unit();
      } while (p > L);
      return k - 1;
    }
    // Box-Muller normal approximation for large means.
    double u1 = unit(), u2 = unit();
    double z = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
    double v = std::floor(mu + std::sqrt(mu) * z + 0.5);
    return v < 0 ? 0 : (int32_t)v;
  }
  double exp1() { return -std::log(1.0 - unit()); }
  int32_t pick(const std::vector<double> &cdf) {
    double u = unit() * cdf.back();
    return (int32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
  }
};

typedef std::pair<int32_t, int32_t> PairKey;
inline PairKey pk(int32_t a, int32_t b) { return a < b ? PairKey(a, b) : PairKey(b, a); }

struct Builder {
  std::vector<int32_t> labels, ef, et, el;
  std::set<PairKey> present;
  std::vector<int32_t> deg;
  int32_t add_vertex(int32_t label) {
    labels.push_back(label);
    deg.push_back(0);
    return (int32_t)labels.size() - 1;
  }
  bool has(int32_t a, int32_t b) const { return present.count(pk(a, b)) != 0; }
  bool add_edge(int32_t a, int32_t b, int32_t lab) {
    if (a == b || has(a, b)) return false;
    present.insert(pk(a, b));
    ef.push_back(a);
    et.push_back(b);
    el.push_back(lab);
    deg[a]++;
    deg[b]++;
    return true;
  }
  void emit(HostDB &db) { db.add_graph(labels, ef, et, el); }
};

}  // namespace

void gen_random(uint64_t seed, int32_t n_graphs, int32_t vmin, int32_t vmax, double edge_factor, int32_t n_vl,
                int32_t n_el, HostDB &db) {
  if (n_graphs < 0 || vmin < 1 || vmax < vmin || n_vl < 1 || n_el < 1 || edge_factor < 0)
    throw Error(GM_E_ARG, "gen_random: bad parameters");
  Rng r(seed);
  db = HostDB();
  for (int32_t g = 0; g < n_graphs; ++g) {
    Builder b;
    int32_t nv = r.range(vmin, vmax);
    for (int32_t v = 0; v < nv; ++v) b.add_vertex(r.below(n_vl));
    // random spanning tree: vertex i attaches to a uniform earlier vertex
    for (int32_t v = 1; v < nv; ++v) b.add_edge(r.below(v), v, r.below(n_el));
    int64_t target = (int64_t)std::llround(edge_factor * nv);
    int64_t max_edges = (int64_t)nv * (nv - 1) / 2;
    target = std::min(target, max_edges);
    while ((int64_t)b.ef.size() < target) {
      int32_t a = r.below(nv), c = r.below(nv);
      if (a == c || b.has(a, c)) continue;
      b.add_edge(a, c, r.below(n_el));
    }
    b.emit(db);
  }
}

void gen_planted(uint64_t seed, int32_t n_graphs, double mean_edges, int32_t n_seed, double mean_pat_edges,
                 int32_t n_vl, int32_t n_el, double merge_prob, HostDB &db) {
  if (n_graphs < 0 || n_seed < 1 || n_vl < 1 || n_el < 1 || mean_edges < 0 || mean_pat_edges < 0)
    throw Error(GM_E_ARG, "gen_planted: bad parameters");
  Rng r(seed);
  db = HostDB();
  // Seed patterns: connected random graphs with max(1, Poisson(mean)) edges.
  // Each edge either hangs a new vertex off a uniform existing one
  // (probability 0.7, or always when the graph is complete) or closes a
  // cycle between two uniform non-adjacent existing vertices.
  struct Pat {
    std::vector<int32_t> labels, ef, et, el;
  };
  std::vector<Pat> pats(n_seed);
  for (auto &p : pats) {
    int32_t m = std::max(1, r.poisson(mean_pat_edges));
    Builder b;
    b.add_vertex(r.below(n_vl));
    while ((int32_t)b.ef.size() < m) {
      int32_t nv = (int32_t)b.labels.size();
      bool complete = (int64_t)b.ef.size() >= (int64_t)nv * (nv - 1) / 2;
      if (complete || r.unit() < 0.7) {
        int32_t a = r.below(nv);
        int32_t v = b.add_vertex(r.below(n_vl));
        b.add_edge(a, v, r.below(n_el));
      } else {
        int32_t a = r.below(nv), c = r.below(nv);
        if (a == c || b.has(a, c)) continue;
        b.add_edge(a, c, r.below(n_el));
      }
    }
    p.labels = b.labels;
    p.ef = b.ef;
    p.et = b.et;
    p.el = b.el;
  }
  std::vector<double> cdf(n_seed);
  double acc = 0;
  for (int32_t i = 0; i < n_seed; ++i) cdf[i] = (acc += r.exp1());

  for (int32_t g = 0; g < n_graphs; ++g) {
    Builder b;
    int32_t target = std::max(1, r.poisson(mean_edges));
    std::vector<std::vector<int32_t>> by_label(n_vl);
    int guard = 0;
    while ((int32_t)b.ef.size() < target && guard++ < 10000) {
      const Pat &p = pats[r.pick(cdf)];
      std::vector<int32_t> map(p.labels.size());
      std::set<int32_t> used;  // transaction vertices taken by this insertion
      for (size_t v = 0; v < p.labels.size(); ++v) {
        int32_t lab = p.labels[v];
        int32_t chosen = -1;
        if (!by_label[lab].empty() && r.unit() < merge_prob) {
          int32_t cand = by_label[lab][r.below((int32_t)by_label[lab].size())];
          if (!used.count(cand)) chosen = cand;
        }
        if (chosen < 0) {
          chosen = b.add_vertex(lab);
          by_label[lab].push_back(chosen);
        }
        used.insert(chosen);
        map[v] = chosen;
      }
      for (size_t e = 0; e < p.ef.size(); ++e) b.add_edge(map[p.ef[e]], map[p.et[e]], p.el[e]);
    }
    // Join connected components in order of their smallest vertex with
    // random labelled edges between uniform members.
    int32_t nv = (int32_t)b.labels.size();
    std::vector<int32_t> parent(nv);
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](int32_t x) {
      while (parent[x] != x) x = parent[x] = parent[parent[x]];
      return x;
    };
    for (size_t e = 0; e < b.ef.size(); ++e) parent[find(b.ef[e])] = find(b.et[e]);
    std::vector<std::vector<int32_t>> comps;
    std::vector<int32_t> comp_of(nv, -1);
    for (int32_t v = 0; v < nv; ++v) {
      int32_t root = find(v);
      if (comp_of[root] < 0) {
        comp_of[root] = (int32_t)comps.size();
        comps.emplace_back();
      }
      comps[comp_of[root]].push_back(v);
    }
    for (size_t c = 1; c < comps.size(); ++c) {
      int32_t a = comps[c - 1][r.below((int32_t)comps[c - 1].size())];
      int32_t d = comps[c][r.below((int32_t)comps[c].size())];
      b.add_edge(a, d, r.below(n_el));
    }
    b.emit(db);
  }
}

void gen_molecule(uint64_t seed, int32_t n_graphs, double mean_v, int32_t max_deg, const std::vector<double> &vp,
                  const std::vector<double> &ep, HostDB &db) {
  if (n_graphs < 0 || max_deg < 2 || vp.empty() || ep.empty()) throw Error(GM_E_ARG, "gen_molecule: bad parameters");
  Rng r(seed);
  db = HostDB();
  std::vector<double> vcdf(vp.size()), ecdf(ep.size());
  std::partial_sum(vp.begin(), vp.end(), vcdf.begin());
  std::partial_sum(ep.begin(), ep.end(), ecdf.begin());
  for (int32_t g = 0; g < n_graphs; ++g) {
    Builder b;
    int32_t nv = std::max(2, r.poisson(mean_v));
    b.add_vertex(r.pick(vcdf));
    for (int32_t v = 1; v < nv; ++v) {
      // attach to a uniform earlier vertex with spare valence
      int32_t a = -1;
      for (int tries = 0; tries < 64 && a < 0; ++tries) {
        int32_t c = r.below(v);
        if (b.deg[c] < max_deg) a = c;
      }
      if (a < 0)
        for (int32_t c = 0; c < v && a < 0; ++c)
          if (b.deg[c] < max_deg) a = c;
      int32_t nvx = b.add_vertex(r.pick(vcdf));
      b.add_edge(a, nvx, r.pick(ecdf));
    }
    // close 1-3 rings of size 5 or 6: join two vertices at tree distance 4
    // or 5 that both have spare valence.
    int32_t rings = r.range(1, 3);
    std::vector<std::vector<int32_t>> adj(nv);
    for (size_t e = 0; e < b.ef.size(); ++e) {
      adj[b.ef[e]].push_back(b.et[e]);
      adj[b.et[e]].push_back(b.ef[e]);
    }
    for (int32_t k = 0; k < rings; ++k) {
      int32_t want = r.range(5, 6) - 1;  // path length between the endpoints
      for (int tries = 0; tries < 32; ++tries) {
        int32_t s = r.below(nv);
        if (b.deg[s] >= max_deg) continue;
        std::vector<int32_t> dist(nv, -1);
        std::vector<int32_t> q{s};
        dist[s] = 0;
        for (size_t h = 0; h < q.size(); ++h)
          for (int32_t w : adj[q[h]])
            if (dist[w] < 0) {
              dist[w] = dist[q[h]] + 1;
              q.push_back(w);
            }
        std::vector<int32_t> cands;
        for (int32_t w = 0; w < nv; ++w)
          if (dist[w] == want && b.deg[w] < max_deg && !b.has(s, w)) cands.push_back(w);
        if (cands.empty()) continue;
        int32_t t = cands[r.below((int32_t)cands.size())];
        if (b.add_edge(s, t, r.pick(ecdf))) {
          adj[s].push_back(t);
          adj[t].push_back(s);
        }
        break;
      }
    }
    b.emit(db);
  }
}

}  // namespace gm
