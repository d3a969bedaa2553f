// gspan_amd_main.cpp -- command-line driver with the reference's contract
// (src/gspan_cuda/common_main.cpp:33-117), written against the C++ driver
// surface (include/gspan_amd.hpp) exactly as the reference's driver is
// written against gspan_cuda:
//
//   gspan_amd <filetype> <filename> <support-int-abs> [output-file]
//
// filetype: txt | bin | fsg | cudabin, optionally '-'-prefixed
// (dbio.cpp:33-41).  Exit codes: 1 usage, 2 support >= database size,
// 3 support < 1 (common_main.cpp:33-70); 4 a mining error (the reference
// logs it and exits 0).  Value lines use the reference logger's layout
// "%P [%t] [%l] [%T] [%n] [%f@%L] - [name=value]" with logger name VAL
// (logger.hpp:13, 79-80), so scripts/parse_log.awk:12-17 reads them:
// absolute_min_support, output, total_time (seconds, fixed),
// total_frequent_graphs.  The optional fourth argument writes one
// to_string(false) line per pattern, as the sequential gSpan's CLI does
// (src/gspan/main.cpp:23-33, 64-67), in that program's report order
// (fully labelled, as gspan_cuda reports codes).
#include <sys/time.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>

#include "gspan_amd.hpp"

namespace {

void log_value(const char *name, const std::string &value, int line) {
  char ts[32];
  const std::time_t now = std::time(nullptr);
  std::strftime(ts, sizeof(ts), "%Y-%m-%d %H:%M:%S", std::localtime(&now));
  std::cout << getpid() << " [0] [INFO] [" << ts << "] [VAL] [gspan_amd_main.cpp@" << line << "] - [" << name << "="
            << value << "]" << std::endl;
}

double seconds_between(const timeval &a, const timeval &b) {
  return (double)b.tv_sec + (double)b.tv_usec / 1e6 - ((double)a.tv_sec + (double)a.tv_usec / 1e6);
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 4 && argc != 5) {
    std::cerr << "usage: \n" << argv[0] << " <filetype> <filename> <support-int-abs> [output-file]" << std::endl;
    return 1;
  }
  const std::string filetype = argv[1], filename = argv[2];
  const int absolute_min_support = atoi(argv[3]);

  gspan_amd::graph_database database;
  try {
    database.load(filetype, filename);
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  if ((size_t)std::max(absolute_min_support, 0) >= database.size()) {
    std::cerr << "error: user specified support " << absolute_min_support << " is greater then the database size "
              << database.size() << std::endl;
    return 2;
  }
  if (absolute_min_support < 1) {
    std::cerr << "error: absolute_min_support < 1" << std::endl;
    return 3;
  }

  std::unique_ptr<gspan_amd::graph_output> out;
  gspan_amd::graph_file_output *file_out = nullptr;
  if (argc == 5) {
    file_out = new gspan_amd::graph_file_output(argv[4], true);
    out.reset(file_out);
  } else {
    out.reset(new gspan_amd::graph_counting_output());
  }

  // the reference's launch parameters are accepted (common_main.cpp:75-84);
  // none of its kernels exist here, so they change nothing
  gspan_amd::cuda_configurator exec_config;
  for (const char *k : {"create_first_embeddings.valid_vertex", "create_first_embeddings.store_map",
                        "create_first_embeddings.find_valid_edge", "create_first_embeddings.edge_store",
                        "compute_extensions.store_validity", "compute_extensions.exec_extension_op",
                        "compute_support.boundaries", "filter_backward_embeddings.store_col_indices",
                        "filter_backward_embeddings.copy_rows"})
    exec_config.set_parameter(k, 128);

  int rc = 0;
  try {
    gspan_amd::gspan_cuda *gspan = gspan_amd::get_gspan_cuda_instance();
    gspan->set_database(database, true);
    gspan->set_min_support(absolute_min_support);
    gspan->set_graph_output(out.get());
    gspan->set_exec_configurator(&exec_config);

    timeval start_time, stop_time;
    gettimeofday(&start_time, nullptr);
    log_value("absolute_min_support", std::to_string(absolute_min_support), __LINE__);
    log_value("output", out->to_string(), __LINE__);
    gspan->run();
    gettimeofday(&stop_time, nullptr);

    std::stringstream t;
    t << std::fixed << seconds_between(start_time, stop_time);
    log_value("total_time", t.str(), __LINE__);
    std::stringstream c;
    c << out->get_count();
    log_value("total_frequent_graphs", c.str(), __LINE__);

    gspan->delete_database_from_device();
    delete gspan;
    if (file_out) file_out->close();  // the patterns, in order
  } catch (const std::exception &e) {
    std::cerr << "caught exception: " << e.what() << std::endl;
    rc = 4;
  }
  return rc;
}
