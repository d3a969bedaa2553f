// host_db.cpp -- graph database readers and writers (host side).
//
// Formats follow the reference's dbio (src/globals/dbio.cpp) and
// Graph::read / read_fsg (src/globals/graph_types.cpp:66-173):
//   txt      "t # id" / "v id label" / "e from to label", tokens split on
//            single spaces (src/globals/utils.cpp:192-210), numbers read with
//            atoi so CRLF line ends are tolerated;
//   fsg      as txt with string labels on "v" lines and "u from to label"
//            edges; label strings map to ids in first-seen order *per graph*
//            (graph_types.cpp:122-157 keeps the maps local to one graph);
//   bin      dbio BIN: size_t graph count, size_t offsets, then per graph a
//            size_t length and the Graph::serialize bytes
//            (dbio.cpp:209-268, graph_types.cpp:446-614);
//   cudabin  graph_database_cuda arrays (dbio.cpp:346-404): padded vertex
//            slots g*max_vertices+v, -1 labels/offsets on padding.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

#include "gm_internal.hpp"
#include "gspan_amd.h"

namespace gm {

void HostDB::check_consistent() const {
  if (n_graphs < 0 || (int32_t)vbegin.size() != n_graphs + 1 || (int32_t)ebegin.size() != n_graphs + 1)
    throw Error(GM_E_ARG, "database offsets do not match the graph count");
  if (vbegin[0] != 0 || ebegin[0] != 0) throw Error(GM_E_ARG, "database offsets must start at 0");
  for (int32_t g = 0; g < n_graphs; ++g) {
    if (vbegin[g + 1] < vbegin[g] || ebegin[g + 1] < ebegin[g])
      throw Error(GM_E_ARG, "database offsets must be non-decreasing");
  }
  if ((int64_t)vbegin[n_graphs] != (int64_t)vlabel.size() || (int64_t)ebegin[n_graphs] != (int64_t)efrom.size() ||
      efrom.size() != eto.size() || efrom.size() != elabel.size())
    throw Error(GM_E_ARG, "database array lengths do not match the offsets");
  for (int32_t g = 0; g < n_graphs; ++g) {
    int32_t nv = vbegin[g + 1] - vbegin[g];
    for (int32_t e = ebegin[g]; e < ebegin[g + 1]; ++e)
      if (efrom[e] < 0 || efrom[e] >= nv || eto[e] < 0 || eto[e] >= nv)
        throw Error(GM_E_ARG, "edge endpoint outside its graph (graph " + std::to_string(g) + ")");
  }
}

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

void read_txt(const std::string &path, bool fsg, HostDB &db) {
  std::ifstream in(path.c_str());
  if (!in) throw Error(GM_E_IO, "cannot open " + path);
  db = HostDB();
  std::string line;
  std::vector<std::string> tok;
  std::vector<int32_t> labels, ef, et, el;
  std::map<std::string, int32_t> vmap, emap;
  auto flush = [&]() {
    if (!labels.empty()) db.add_graph(labels, ef, et, el);
    labels.clear();
    ef.clear();
    et.clear();
    el.clear();
    vmap.clear();
    emap.clear();
  };
  while (std::getline(in, line)) {
    split_spaces(line, tok);
    if (tok.empty()) continue;
    if (tok[0] == "t") {
      flush();
    } else if (tok[0] == "v" && tok.size() >= 3) {
      int32_t id = atoi(tok[1].c_str());
      if (id < 0) throw Error(GM_E_ARG, "negative vertex id in " + path);
      labels.resize((size_t)id + 1, 0);
      if (fsg) {
        auto it = vmap.find(tok[2]);
        if (it == vmap.end()) it = vmap.emplace(tok[2], (int32_t)vmap.size()).first;
        labels[id] = it->second;
      } else {
        labels[id] = atoi(tok[2].c_str());
      }
    } else if (tok[0] == (fsg ? "u" : "e") && tok.size() >= 4) {
      int32_t a = atoi(tok[1].c_str()), b = atoi(tok[2].c_str());
      if (a < 0 || b < 0 || a >= (int32_t)labels.size() || b >= (int32_t)labels.size())
        throw Error(GM_E_ARG, "Format Error: define vertex lists before edges (" + path + ")");
      int32_t lab;
      if (fsg) {
        auto it = emap.find(tok[3]);
        if (it == emap.end()) it = emap.emplace(tok[3], (int32_t)emap.size()).first;
        lab = it->second;
      } else {
        lab = atoi(tok[3].c_str());
      }
      ef.push_back(a);
      et.push_back(b);
      el.push_back(lab);
    }
  }
  flush();
}

void write_txt(const HostDB &db, const std::string &path) {
  FILE *f = fopen(path.c_str(), "w");
  if (!f) throw Error(GM_E_IO, "cannot write " + path);
  for (int32_t g = 0; g < db.n_graphs; ++g) {
    fprintf(f, "t # %d\n", g);
    for (int32_t v = db.vbegin[g]; v < db.vbegin[g + 1]; ++v) fprintf(f, "v %d %d\n", v - db.vbegin[g], db.vlabel[v]);
    for (int32_t e = db.ebegin[g]; e < db.ebegin[g + 1]; ++e)
      fprintf(f, "e %d %d %d\n", db.efrom[e], db.eto[e], db.elabel[e]);
  }
  if (fclose(f) != 0) throw Error(GM_E_IO, "cannot write " + path);
}

namespace {
struct Cursor {
  const std::vector<char> &buf;
  size_t pos;
  template <class T>
  T get() {
    if (pos + sizeof(T) > buf.size()) throw Error(GM_E_ARG, "truncated database file");
    T v;
    std::memcpy(&v, buf.data() + pos, sizeof(T));
    pos += sizeof(T);
    return v;
  }
};

std::vector<char> slurp(const std::string &path) {
  std::ifstream in(path.c_str(), std::ios::binary);
  if (!in) throw Error(GM_E_IO, "cannot open " + path);
  return std::vector<char>((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}
}  // namespace

void read_bin(const std::string &path, HostDB &db) {
  std::vector<char> buf = slurp(path);
  db = HostDB();
  Cursor c{buf, 0};
  uint64_t n = c.get<uint64_t>();
  if (n > (1ull << 31)) throw Error(GM_E_ARG, "implausible graph count in " + path);
  std::vector<uint64_t> offs(n);
  for (uint64_t i = 0; i < n; ++i) offs[i] = c.get<uint64_t>();
  std::vector<int32_t> labels, ef, et, el;
  for (uint64_t i = 0; i < n; ++i) {
    if (c.pos != offs[i]) throw Error(GM_E_ARG, "graph stored at an unexpected offset in " + path);
    uint64_t stored = c.get<uint64_t>();
    size_t start = c.pos;
    int32_t total = c.get<int32_t>();
    (void)c.get<int32_t>();  // edge_size_
    (void)c.get<uint8_t>();  // directed flag (bool)
    int32_t nv = c.get<int32_t>();
    if ((uint64_t)total != stored || nv < 0) throw Error(GM_E_ARG, "corrupt graph record in " + path);
    labels.assign(nv, 0);
    ef.clear();
    et.clear();
    el.clear();
    for (int32_t v = 0; v < nv; ++v) {
      (void)c.get<int32_t>();  // serialized vertex size
      labels[v] = c.get<int32_t>();
      int32_t ne = c.get<int32_t>();
      for (int32_t k = 0; k < ne; ++k) {
        int32_t from = c.get<int32_t>(), to = c.get<int32_t>(), lab = c.get<int32_t>();
        (void)c.get<int32_t>();  // edge id
        if (from != v || to < 0 || to >= nv) throw Error(GM_E_ARG, "corrupt adjacency in " + path);
        if (from < to) {  // undirected: keep one direction of each pair
          ef.push_back(from);
          et.push_back(to);
          el.push_back(lab);
        }
      }
    }
    if (c.pos - start != stored) throw Error(GM_E_ARG, "graph record length mismatch in " + path);
    db.add_graph(labels, ef, et, el);
  }
}

void read_cudabin(const std::string &path, HostDB &db) {
  std::vector<char> buf = slurp(path);
  db = HostDB();
  Cursor c{buf, 0};
  int32_t n = c.get<int32_t>(), ne = c.get<int32_t>(), mvc = c.get<int32_t>();
  (void)c.get<int32_t>();  // vertex_count
  if (n < 0 || ne < 0 || mvc < 0) throw Error(GM_E_ARG, "corrupt cudabin header in " + path);
  std::vector<int32_t> edges(ne), elabs(ne), vlabs((size_t)n * mvc), voffs((size_t)n * mvc + 1);
  for (auto &x : edges) x = c.get<int32_t>();
  for (auto &x : elabs) x = c.get<int32_t>();
  for (auto &x : vlabs) x = c.get<int32_t>();
  for (auto &x : voffs) x = c.get<int32_t>();
  std::vector<int32_t> labels, ef, et, el;
  int64_t slots = (int64_t)n * mvc;
  for (int32_t g = 0; g < n; ++g) {
    labels.clear();
    ef.clear();
    et.clear();
    el.clear();
    int64_t base = (int64_t)g * mvc;
    for (int32_t v = 0; v < mvc && vlabs[base + v] != -1; ++v) labels.push_back(vlabs[base + v]);
    int32_t nv = (int32_t)labels.size();
    for (int32_t v = 0; v < nv; ++v) {
      int64_t s = voffs[base + v];
      int64_t nx = base + v + 1;  // next valid vertex slot (cuda_graph_types.hpp:98-113)
      while (nx < slots && voffs[nx] == -1) ++nx;
      int64_t e = voffs[nx];
      if (s < 0 || e < s || e > ne) throw Error(GM_E_ARG, "corrupt cudabin offsets in " + path);
      for (int64_t k = s; k < e; ++k) {
        int32_t to = (int32_t)(edges[k] - base);
        if (to < 0 || to >= nv) throw Error(GM_E_ARG, "corrupt cudabin adjacency in " + path);
        if (v < to) {
          ef.push_back(v);
          et.push_back(to);
          el.push_back(elabs[k]);
        }
      }
    }
    db.add_graph(labels, ef, et, el);
  }
}

}  // namespace gm
