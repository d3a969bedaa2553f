// ref_harness.cpp -- driver for a build of the reference's sequential gSpan
// (src/gspan + src/globals, compiled in place from /root/reference by
// oracle/Makefile.ref).  TEST INFRASTRUCTURE ONLY: it writes golden
// fixtures; nothing in the product links it.
//
// It mirrors the reference CLI (src/gspan/main.cpp:21-96) but plugs in a
// graph_output sink that records the support next to each code, which the
// reference's own file sink drops (src/globals/graph_output.cpp:145-150).
// Output: one line per reported pattern, "<code to_string(false)>\t<support>",
// in emission order.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <sys/time.h>

#include <dbio.hpp>
#include <graph_output.hpp>
#include <gspan.hpp>

class support_file_output : public graph_output {
  FILE *out_;
  double count_;

 public:
  explicit support_file_output(FILE *f) : out_(f), count_(0) {}
  virtual void output_graph(const types::Graph &, size_t) { abort(); }
  virtual void output_graph(const types::DFSCode &code, size_t support) {
    fprintf(out_, "%s\t%zu\n", code.to_string(false).c_str(), support);
    count_ += 1.0;
  }
  virtual double get_count() { return count_; }
  virtual std::string to_string() { return "support_file_output"; }
};

int main(int argc, char **argv) {
  if (argc != 4 && argc != 5) {
    fprintf(stderr, "usage: %s <txt|fsg> <filename> <support-int-abs> [<out>]\n", argv[0]);
    return 1;
  }
  types::graph_database_t database;
  dbio::read_database(dbio::ftype2str(std::string(argv[1])), argv[2], database);
  int minsup = atoi(argv[3]);
  if (minsup < 1) return 3;
  FILE *f = argc == 5 ? fopen(argv[4], "w") : stdout;
  support_file_output sink(f);
  timeval t0, t1;
  gettimeofday(&t0, 0);
  GSPAN::gSpan gspan;
  gspan.set_database(database);
  gspan.set_min_support(minsup);
  gspan.set_graph_output(&sink);
  gspan.run();
  gettimeofday(&t1, 0);
  double secs = (t1.tv_sec - t0.tv_sec) + 1e-6 * (t1.tv_usec - t0.tv_usec);
  fprintf(stderr, "[VAL] [total_time=%.6f]\n[VAL] [total_frequent_graphs=%.0f]\n", secs, sink.get_count());
  if (f != stdout) fclose(f);
  return 0;
}
