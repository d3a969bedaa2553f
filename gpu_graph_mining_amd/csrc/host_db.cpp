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
#include <tuple>

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

// One graph record of a BIN file (Graph::serialize, graph_types.cpp:566-614)
// at c.pos, appended to db.  Each adjacency entry appears once per
// direction; an edge is kept from its lower endpoint, a self-loop (both of
// its entries sit at the vertex) once per pair of entries.
static void read_graph_record(Cursor &c, const std::string &path, HostDB &db, std::vector<int32_t> &labels,
                              std::vector<int32_t> &ef, std::vector<int32_t> &et, std::vector<int32_t> &el) {
  const uint64_t stored = c.get<uint64_t>();
  const size_t start = c.pos;
  const int32_t total = c.get<int32_t>();
  (void)c.get<int32_t>();  // edge_size_
  (void)c.get<uint8_t>();  // directed flag (bool)
  const int32_t nv = c.get<int32_t>();
  if ((uint64_t)total != stored || nv < 0) throw Error(GM_E_ARG, "corrupt graph record in " + path);
  labels.assign(nv, 0);
  ef.clear();
  et.clear();
  el.clear();
  for (int32_t v = 0; v < nv; ++v) {
    (void)c.get<int32_t>();  // serialized vertex size
    labels[v] = c.get<int32_t>();
    const int32_t ne = c.get<int32_t>();
    bool loop_open = false;
    for (int32_t k = 0; k < ne; ++k) {
      const int32_t from = c.get<int32_t>(), to = c.get<int32_t>(), lab = c.get<int32_t>();
      (void)c.get<int32_t>();  // edge id
      if (from != v || to < 0 || to >= nv) throw Error(GM_E_ARG, "corrupt adjacency in " + path);
      if (from < to || (from == to && (loop_open = !loop_open))) {
        ef.push_back(from);
        et.push_back(to);
        el.push_back(lab);
      }
    }
  }
  if (c.pos - start != stored) throw Error(GM_E_ARG, "graph record length mismatch in " + path);
  db.add_graph(labels, ef, et, el);
}

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
    read_graph_record(c, path, db, labels, ef, et, el);
  }
  if (c.pos != buf.size()) throw Error(GM_E_ARG, "trailing bytes after the last graph in " + path);
}

// Part `part` of `parts` of a BIN file: transactions [n / parts * part,
// n / parts * (part + 1)), the last part up to n, read from their stored
// offset (dbio.cpp:275-329).
void read_bin_part(const std::string &path, int32_t parts, int32_t part, HostDB &db) {
  if (parts < 1 || part < 0 || part >= parts)
    throw Error(GM_E_ARG, "incorrect parameter value: part should be from [0, total_parts-1]");
  std::vector<char> buf = slurp(path);
  db = HostDB();
  Cursor c{buf, 0};
  const uint64_t n = c.get<uint64_t>();
  if (n > (1ull << 31)) throw Error(GM_E_ARG, "implausible graph count in " + path);
  const int64_t first = (int64_t)n / parts * part;
  const int64_t last = part == parts - 1 ? (int64_t)n : (int64_t)n / parts * (part + 1);
  if (first >= last) return;
  c.pos = sizeof(uint64_t) * (1 + (size_t)first);
  c.pos = (size_t)c.get<uint64_t>();
  std::vector<int32_t> labels, ef, et, el;
  for (int64_t i = first; i < last; ++i) read_graph_record(c, path, db, labels, ef, et, el);
}

namespace {
// Adjacency of graph g as Graph::read builds it (graph_types.cpp:66-113):
// every edge in input order pushed at its from vertex, then at its to
// vertex; entry = (to, elabel, id) with ids as Graph::buildEdge assigns
// them (graph_types.cpp:34-60): in order of first appearance scanning the
// vertices, one id per (low, high, elabel).
struct AdjEntry {
  int32_t to, elabel, id;
};
int32_t build_adjacency(const HostDB &db, int32_t g, std::vector<std::vector<AdjEntry>> &adj) {
  const int32_t nv = db.vbegin[g + 1] - db.vbegin[g];
  adj.assign(nv, {});
  for (int32_t e = db.ebegin[g]; e < db.ebegin[g + 1]; ++e) {
    adj[db.efrom[e]].push_back(AdjEntry{db.eto[e], db.elabel[e], -1});
    adj[db.eto[e]].push_back(AdjEntry{db.efrom[e], db.elabel[e], -1});
  }
  std::map<std::tuple<int32_t, int32_t, int32_t>, int32_t> ids;
  for (int32_t v = 0; v < nv; ++v)
    for (AdjEntry &a : adj[v]) {
      const auto key = std::make_tuple(std::min(v, a.to), std::max(v, a.to), a.elabel);
      auto it = ids.find(key);
      if (it == ids.end()) it = ids.emplace(key, (int32_t)ids.size()).first;
      a.id = it->second;
    }
  return (int32_t)ids.size();
}

template <class T>
void put(std::vector<char> &out, T v) {
  const char *p = reinterpret_cast<const char *>(&v);
  out.insert(out.end(), p, p + sizeof(T));
}

void write_file(const std::string &path, const std::vector<char> &bytes) {
  std::ofstream os(path.c_str(), std::ios::binary | std::ios::trunc);
  if (!os) throw Error(GM_E_IO, "cannot write " + path);
  os.write(bytes.data(), (std::streamsize)bytes.size());
  if (!os) throw Error(GM_E_IO, "cannot write " + path);
}
}  // namespace

// BIN (dbio.cpp:209-268): size_t graph count, size_t offset of each graph
// record, then per graph a size_t length and the Graph::serialize bytes --
// int total size, int edge count, bool directed (false), int vertex count,
// per vertex int size, int label, int entry count and (from, to, elabel,
// id) per adjacency entry (graph_types.cpp:446-614).
void write_bin(const HostDB &db, const std::string &path) {
  if (db.n_graphs == 0) throw Error(GM_E_ARG, "Cannot save empty database.");
  std::vector<char> body;
  std::vector<uint64_t> offs;
  std::vector<std::vector<AdjEntry>> adj;
  const uint64_t head = sizeof(uint64_t) * (1 + (uint64_t)db.n_graphs);
  for (int32_t g = 0; g < db.n_graphs; ++g) {
    offs.push_back(head + body.size());
    const int32_t edges = build_adjacency(db, g, adj);
    const int32_t nv = (int32_t)adj.size();
    int32_t size = 3 * (int32_t)sizeof(int32_t) + (int32_t)sizeof(bool);
    for (int32_t v = 0; v < nv; ++v) size += 3 * (int32_t)sizeof(int32_t) + 16 * (int32_t)adj[v].size();
    put<uint64_t>(body, (uint64_t)size);
    put<int32_t>(body, size);
    put<int32_t>(body, edges);
    put<uint8_t>(body, 0);  // directed = false
    put<int32_t>(body, nv);
    for (int32_t v = 0; v < nv; ++v) {
      put<int32_t>(body, 3 * (int32_t)sizeof(int32_t) + 16 * (int32_t)adj[v].size());
      put<int32_t>(body, db.vlabel[db.vbegin[g] + v]);
      put<int32_t>(body, (int32_t)adj[v].size());
      for (const AdjEntry &a : adj[v]) {
        put<int32_t>(body, v);
        put<int32_t>(body, a.to);
        put<int32_t>(body, a.elabel);
        put<int32_t>(body, a.id);
      }
    }
  }
  std::vector<char> bytes;
  put<uint64_t>(bytes, (uint64_t)db.n_graphs);
  for (uint64_t o : offs) put<uint64_t>(bytes, o);
  bytes.insert(bytes.end(), body.begin(), body.end());
  write_file(path, bytes);
}

// cudabin (dbio.cpp:346-370): int graph count, int adjacency entries, int
// largest vertex count, int vertex count, then the padded-CSR arrays of
// graph_database_cuda::create_from_host_representation
// (cuda_graph_types.cpp:65-122): entries (neighbour slot g * m + to), their
// edge labels, vertex labels per slot (-1 on padding), vertex offsets per
// slot (-1 on padding) and the total at slot n * m.
void write_cudabin(const HostDB &db, const std::string &path) {
  const int32_t m = db.max_vertices();
  std::vector<std::vector<AdjEntry>> adj;
  std::vector<int32_t> edges, elabs, vlabs((size_t)db.n_graphs * m, -1), voffs((size_t)db.n_graphs * m + 1, -1);
  for (int32_t g = 0; g < db.n_graphs; ++g) {
    build_adjacency(db, g, adj);
    const int64_t base = (int64_t)g * m;
    for (int32_t v = 0; v < (int32_t)adj.size(); ++v) {
      voffs[base + v] = (int32_t)edges.size();
      vlabs[base + v] = db.vlabel[db.vbegin[g] + v];
      for (const AdjEntry &a : adj[v]) {
        edges.push_back((int32_t)(a.to + base));
        elabs.push_back(a.elabel);
      }
    }
  }
  voffs[(size_t)db.n_graphs * m] = (int32_t)edges.size();
  std::vector<char> bytes;
  put<int32_t>(bytes, db.n_graphs);
  put<int32_t>(bytes, (int32_t)edges.size());
  put<int32_t>(bytes, m);
  put<int32_t>(bytes, (int32_t)db.n_vertices());
  for (int32_t x : edges) put<int32_t>(bytes, x);
  for (int32_t x : elabs) put<int32_t>(bytes, x);
  for (int32_t x : vlabs) put<int32_t>(bytes, x);
  for (int32_t x : voffs) put<int32_t>(bytes, x);
  write_file(path, bytes);
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
      bool loop_open = false;
      for (int64_t k = s; k < e; ++k) {
        int32_t to = (int32_t)(edges[k] - base);
        if (to < 0 || to >= nv) throw Error(GM_E_ARG, "corrupt cudabin adjacency in " + path);
        if (v < to || (v == to && (loop_open = !loop_open))) {
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
