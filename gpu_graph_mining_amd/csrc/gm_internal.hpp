// gm_internal.hpp -- host-side types shared by the library's C++ sources.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace gm {

// Error carrying one of the GM_E_* classes of include/gspan_amd.h.
struct Error : public std::runtime_error {
  int code;
  Error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

// A graph database in edge-list form: graph g owns vertices
// [vbegin[g], vbegin[g+1]) and edges [ebegin[g], ebegin[g+1]) whose
// endpoints are local vertex ids.  This is the host-side equivalent of
// types::graph_database_t (src/globals/types.hpp:31), kept as flat arrays.
struct HostDB {
  int32_t n_graphs = 0;
  std::vector<int32_t> vbegin{0}, vlabel, ebegin{0}, efrom, eto, elabel;

  int64_t n_vertices() const { return (int64_t)vlabel.size(); }
  int64_t n_edges() const { return (int64_t)efrom.size(); }
  int32_t max_vertices() const {
    int32_t m = 0;
    for (int32_t g = 0; g < n_graphs; ++g) m = std::max(m, vbegin[g + 1] - vbegin[g]);
    return m;
  }
  void add_graph(const std::vector<int32_t> &labels, const std::vector<int32_t> &ef,
                 const std::vector<int32_t> &et, const std::vector<int32_t> &el) {
    vlabel.insert(vlabel.end(), labels.begin(), labels.end());
    efrom.insert(efrom.end(), ef.begin(), ef.end());
    eto.insert(eto.end(), et.begin(), et.end());
    elabel.insert(elabel.end(), el.begin(), el.end());
    vbegin.push_back((int32_t)vlabel.size());
    ebegin.push_back((int32_t)efrom.size());
    ++n_graphs;
  }
  // Throws Error(GM_E_ARG) when arrays are inconsistent.
  void check_consistent() const;
};

// Readers / writers (host_db.cpp).
void read_txt(const std::string &path, bool fsg, HostDB &db);
void read_bin(const std::string &path, HostDB &db);
void read_bin_part(const std::string &path, int32_t parts, int32_t part, HostDB &db);
void read_cudabin(const std::string &path, HostDB &db);
void write_txt(const HostDB &db, const std::string &path);
void write_bin(const HostDB &db, const std::string &path);
void write_cudabin(const HostDB &db, const std::string &path);

// Synthetic generators (generators.cpp).
void gen_random(uint64_t seed, int32_t n_graphs, int32_t vmin, int32_t vmax, double edge_factor,
                int32_t n_vl, int32_t n_el, HostDB &db);
void gen_planted(uint64_t seed, int32_t n_graphs, double mean_edges, int32_t n_seed, double mean_pat_edges,
                 int32_t n_vl, int32_t n_el, double merge_prob, HostDB &db);
void gen_molecule(uint64_t seed, int32_t n_graphs, double mean_v, int32_t max_deg, const std::vector<double> &vp,
                  const std::vector<double> &ep, HostDB &db);

// One DFS tuple.  POD: no virtual methods (the reference's DFS has some,
// src/globals/dfs_code.hpp:60-67, which this layout deliberately drops).
struct DFS {
  int32_t from, to, fromlabel, elabel, tolabel;
  bool forward() const { return from < to; }
};

// Canonical (minimum DFS code) check (canonical.cpp).  The code must be
// fully labelled.  Returns true when the code is its own minimum in the
// sense of the sequential gSpan's is_min (src/gspan/ismin.cpp:36-140).
bool dfs_code_is_min(const DFS *code, int n);

// The minimum DFS code of the pattern a code describes (any valid DFS code
// of a connected pattern; vertex labels carried by its forward tuples),
// fully labelled, in canonical.cpp.  Throws Error(GM_E_ARG) for codes that
// do not describe a connected pattern.
void min_dfs_code(const DFS *code, int n, std::vector<DFS> &out);

}  // namespace gm
