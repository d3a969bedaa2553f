/*
 * gspan_amd.hpp -- C++ driver surface of the MI355X-native gSpan miner,
 * header-only over the C ABI of gspan_amd.h.
 *
 * It keeps the method names and call order of the reference's mining
 * driver so that a reference-style driver (src/gspan_cuda/common_main.cpp:
 * 85-117) compiles against it with only its includes and namespaces
 * changed:
 *
 *   reference (file:line)                           here
 *   gspan_cuda::gspan_cuda (gspan_cuda.hpp:281-292)  gspan_amd::gspan_cuda
 *   get_gspan_cuda_instance() (common_main.cpp:28)   gspan_amd::get_gspan_cuda_instance()
 *   graph_output / graph_counting_output /
 *     graph_file_output (graph_output.hpp:9-52)      gspan_amd::graph_output / ...
 *   types::DFS / types::DFSCode to_string
 *     (dfs_code.cpp:109-118, 199-214)                gspan_amd::DFS / DFSCode
 *   dbio::read_database / ftype2str
 *     (dbio.hpp:8-37; dbio.cpp:33-41)                gspan_amd::graph_database::load
 *   cudapp::cuda_configurator
 *     (cuda_configurator.hpp)                        gspan_amd::cuda_configurator
 *
 * Nothing here talks to the GPU itself: every call goes through gm_*.  Errors
 * become std::runtime_error, as the reference's CUDA_EXEC checks throw
 * (src/cuda_pp/cuda_tools.hpp:10-18).
 */
#ifndef GSPAN_AMD_HPP
#define GSPAN_AMD_HPP

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <exception>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gspan_amd.h"

namespace gspan_amd {

inline void check(int rc, const char *what) {
  if (rc != GM_OK) throw std::runtime_error(std::string(what) + ": [" + std::to_string(rc) + "] " + gm_last_error());
}

// One DFS tuple (from, to, fromlabel, elabel, tolabel).
struct DFS {
  int from = 0, to = 0, fromlabel = 0, elabel = 0, tolabel = 0;
  bool is_forward() const { return from < to; }
  bool is_backward() const { return from > to; }
  // "F(0 1 2 0 3)" / "B(...)" with the edge type, "(0 1 2 0 3)" without.
  std::string to_string(bool print_edge_type = true) const {
    std::stringstream ss;
    if (print_edge_type) ss << (is_forward() ? "F" : "B");
    ss << "(" << from << " " << to << " " << fromlabel << " " << elabel << " " << tolabel << ")";
    return ss.str();
  }
};

// A DFS code; to_string joins the tuples with ';'.
struct DFSCode : public std::vector<DFS> {
  std::string to_string(bool print_edge_type = true) const {
    std::string s;
    for (size_t i = 0; i < size(); ++i) {
      if (i) s += ";";
      s += (*this)[i].to_string(print_edge_type);
    }
    return s;
  }
};

// Pattern sinks (graph_output.hpp:9-52).
class graph_output {
 public:
  virtual void output_graph(const DFSCode &code, size_t support = 0) = 0;
  virtual double get_count() = 0;
  virtual std::string to_string() = 0;
  virtual ~graph_output() {}
};

class graph_counting_output : public graph_output {
  double counter = 0;

 public:
  void output_graph(const DFSCode &, size_t = 0) override { counter += 1.0; }
  double get_count() override { return counter; }
  std::string to_string() override {
    std::stringstream ss;
    ss << counter;
    return ss.str();
  }
};

// The order in which the sequential gSpan reports patterns: depth first,
// the children of a pattern in the order of its extension maps
// (src/gspan/gspan.cpp:251-271): backward tuples by `to` then edge label
// ascending, before forward tuples by `from` descending, then edge label,
// then to-label ascending; roots by (fromlabel, elabel, tolabel).  A
// pattern precedes its extensions, so pre-order is this lexicographic order
// of the codes.
inline bool dfs_less(const DFSCode &a, const DFSCode &b) {
  const size_t n = a.size() < b.size() ? a.size() : b.size();
  for (size_t i = 0; i < n; ++i) {
    const DFS &x = a[i], &y = b[i];
    if (x.from == y.from && x.to == y.to && x.fromlabel == y.fromlabel && x.elabel == y.elabel &&
        x.tolabel == y.tolabel)
      continue;
    if (i == 0) {
      if (x.fromlabel != y.fromlabel) return x.fromlabel < y.fromlabel;
      if (x.elabel != y.elabel) return x.elabel < y.elabel;
      return x.tolabel < y.tolabel;
    }
    const bool xb = x.from > x.to, yb = y.from > y.to;
    if (xb != yb) return xb;  // backward first
    if (xb) return x.to != y.to ? x.to < y.to : x.elabel < y.elabel;
    if (x.from != y.from) return x.from > y.from;
    if (x.elabel != y.elabel) return x.elabel < y.elabel;
    return x.tolabel < y.tolabel;
  }
  return a.size() < b.size();
}

// One code.to_string(false) line per pattern (graph_output.cpp:90-95), as
// they come, or (dfs_order) all of them at close() / destruction in the
// sequential gSpan's order (dfs_less), which makes the file with wildcard
// labels byte-identical to that program's output file.
class graph_file_output : public graph_output {
  std::ofstream out;
  std::string filename;
  double count = 0;
  bool dfs_order;
  std::vector<DFSCode> held;

 public:
  explicit graph_file_output(const std::string &fname, bool dfs_order_ = false)
      : out(fname.c_str()), filename(fname), dfs_order(dfs_order_) {
    if (!out) throw std::runtime_error("cannot open " + fname);
  }
  ~graph_file_output() override {
    try {
      close();
    } catch (...) {
    }
  }
  void output_graph(const DFSCode &code, size_t = 0) override {
    if (dfs_order) {
      held.push_back(code);
    } else {
      out << code.to_string(false) << "\n";
      if (!out) throw std::runtime_error("cannot write " + filename);
    }
    count += 1.0;
  }
  // write what is held and close the file
  void close() {
    if (!out.is_open()) return;
    if (dfs_order) {
      std::stable_sort(held.begin(), held.end(), dfs_less);
      for (const DFSCode &c : held) out << c.to_string(false) << "\n";
      held.clear();
    }
    out.close();
    if (out.fail()) throw std::runtime_error("cannot write " + filename);
  }
  double get_count() override { return count; }
  std::string to_string() override { return "graph_file_output, filename: " + filename; }
};

// A graph database (types::graph_database_t with the dbio readers).
class graph_database {
  gm_db *db_ = nullptr;

 public:
  graph_database() = default;
  graph_database(const graph_database &) = delete;
  graph_database &operator=(const graph_database &) = delete;
  ~graph_database() { gm_db_destroy(db_); }
  // filetype "txt", "fsg", "bin" or "cudabin", optionally '-'-prefixed
  void load(const std::string &filetype, const std::string &filename) {
    gm_db *d = nullptr;
    check(gm_db_load(filename.c_str(), filetype.c_str(), &d), "read_database");
    gm_db_destroy(db_);
    db_ = d;
  }
  // from vertex and edge arrays (gm_db_create): the hand-over point for a
  // database already held in memory, e.g. a types::graph_database_t
  void create(const std::vector<int32_t> &graph_vbegin, const std::vector<int32_t> &vlabels,
              const std::vector<int32_t> &graph_ebegin, const std::vector<int32_t> &efrom,
              const std::vector<int32_t> &eto, const std::vector<int32_t> &elabel) {
    if (graph_vbegin.empty() || graph_ebegin.size() != graph_vbegin.size())
      throw std::runtime_error("create: graph_vbegin / graph_ebegin need one entry more than there are graphs");
    gm_db *d = nullptr;
    check(gm_db_create((int32_t)graph_vbegin.size() - 1, graph_vbegin.data(), vlabels.data(), graph_ebegin.data(),
                       efrom.data(), eto.data(), elabel.data(), &d),
          "create");
    gm_db_destroy(db_);
    db_ = d;
  }
  size_t size() const {
    if (!db_) return 0;
    int64_t c[6];
    check(gm_db_counts(db_, c), "gm_db_counts");
    return (size_t)c[0];
  }
  gm_db *handle() const { return db_; }
};

// Launch parameters by name, as cudapp::cuda_configurator holds the
// reference's per-kernel block sizes.  The names that mean something here
// are the gm_cfg fields (label_style, max_edges, canon_threads,
// path_pruning, speculate, chunk_rows, table_bits, canon_host,
// collect_proj); the
// reference's own kernel names are accepted and have no effect (their
// kernels do not exist here).
class cuda_configurator {
  std::map<std::string, int> p_;
  std::string file_, id_;

 public:
  cuda_configurator() = default;
  // a configuration file as the reference's reads it
  // (src/cuda_pp/cuda_configurator.cpp:26-76): header `id,<name>,...`, then
  // one row per configuration id; parse_config_file() takes row `config_id`
  // (gpu_graph_mining_amd/tune.py writes such files)
  cuda_configurator(const std::string &config_file, const std::string &config_id)
      : file_(config_file), id_(config_id) {}
  void parse_config_file() {
    std::ifstream is(file_.c_str());
    if (!is) throw std::runtime_error("cuda_configurator: could not open config file " + file_);
    auto split = [](const std::string &line) {
      std::vector<std::string> out;
      std::stringstream ss(line);
      std::string cell;
      while (std::getline(ss, cell, ',')) {
        const size_t a = cell.find_first_not_of(" \t\r"), b = cell.find_last_not_of(" \t\r");
        out.push_back(a == std::string::npos ? std::string() : cell.substr(a, b - a + 1));
      }
      return out;
    };
    std::string line;
    if (!std::getline(is, line)) throw std::runtime_error("cuda_configurator: empty config file " + file_);
    const std::vector<std::string> header = split(line);
    while (std::getline(is, line)) {
      const std::vector<std::string> row = split(line);
      if (row.empty() || row[0] != id_) continue;
      for (size_t i = 1; i < row.size() && i < header.size(); ++i) p_[header[i]] = std::atoi(row[i].c_str());
      return;
    }
    throw std::runtime_error("cuda_configurator: no configuration " + id_ + " in " + file_);
  }
  void set_parameter(const std::string &name, int value) { p_[name] = value; }
  int get_parameter(const std::string &name) const {
    auto it = p_.find(name);
    if (it == p_.end()) throw std::runtime_error("unknown parameter " + name);
    return it->second;
  }
  void fill(gm_cfg &cfg) const {
    gm_cfg_init(&cfg);
    auto get = [&](const char *k, int32_t &field) {
      auto it = p_.find(k);
      if (it != p_.end()) field = it->second;
    };
    get("label_style", cfg.label_style);
    get("max_edges", cfg.max_edges);
    get("canon_threads", cfg.canon_threads);
    get("path_pruning", cfg.path_pruning);
    get("speculate", cfg.speculate);
    get("chunk_rows", cfg.chunk_rows);
    get("table_bits", cfg.table_bits);
    get("canon_host", cfg.canon_host);
    get("collect_proj", cfg.collect_proj);
  }
};

// The mining driver: set_database, set_min_support, set_graph_output,
// set_exec_configurator, run, delete_database_from_device
// (gspan_cuda.hpp:281-292).  Patterns reach the graph_output fully
// labelled, as gspan_cuda reports them (gspan_cuda_no_sort.cpp:409-416).
class gspan_cuda {
 protected:
  graph_database *db_ = nullptr;
  int minsup_ = 0;
  graph_output *out_ = nullptr;
  cuda_configurator *cc_ = nullptr;

 public:
  gspan_cuda() = default;
  virtual ~gspan_cuda() {}
  // convert_to_cuda: copy the device layout now (else the first run does)
  void set_database(graph_database &graph_database, bool convert_to_cuda = true) {
    db_ = &graph_database;
    if (convert_to_cuda) check(gm_db_upload(db_->handle()), "set_database");
  }
  void delete_database_from_device() { db_ = nullptr; }
  void set_min_support(int minsup) { minsup_ = minsup; }
  void set_graph_output(graph_output *gout) { out_ = gout; }
  void set_exec_configurator(cuda_configurator *cc) { cc_ = cc; }
  virtual void run() {
    if (!cc_) throw std::runtime_error("exec configurator is not set");  // gspan_cuda.cpp:127
    if (!db_ || !db_->handle()) throw std::runtime_error("database is not set");
    gm_cfg cfg;
    cc_->fill(cfg);
    Sink sink{out_, nullptr};
    const int rc = gm_mine(db_->handle(), minsup_, &cfg, out_ ? &gspan_cuda::report : nullptr, &sink, nullptr);
    if (sink.error) std::rethrow_exception(sink.error);  // the sink's own exception
    check(rc, "run");
  }

 private:
  struct Sink {
    graph_output *out;
    std::exception_ptr error;
  };
  // no exception crosses the C ABI: a throwing sink stops mining and its
  // exception is rethrown by run()
  static int report(void *user, const int32_t *dfs5, int32_t n_edges, uint32_t support, uint64_t) {
    Sink *sink = static_cast<Sink *>(user);
    try {
      DFSCode code;
      code.resize((size_t)n_edges);
      for (int32_t i = 0; i < n_edges; ++i) {
        const int32_t *d = dfs5 + 5 * i;
        code[(size_t)i] = DFS{d[0], d[1], d[2], d[3], d[4]};
      }
      sink->out->output_graph(code, support);
      return 0;
    } catch (...) {
      sink->error = std::current_exception();
      return 1;
    }
  }
};

// The factory the reference CLI links against (common_main.cpp:28).
inline gspan_cuda *get_gspan_cuda_instance() { return new gspan_cuda(); }

}  // namespace gspan_amd

#endif /* GSPAN_AMD_HPP */
