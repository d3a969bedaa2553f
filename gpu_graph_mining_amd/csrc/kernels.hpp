// kernels.hpp -- launch interface of the level-step kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace gm {

// One workgroup owns one graph; its size (64, 128 or 256 threads) follows
// the level's average rows per graph.  A row's children number at most
// 65535 (uint16 counts): a row extends by at most 32 * 256 edges.
constexpr int kMaxBlock = 256;

// Everything one level-step kernel reads or writes.  Plain device pointers.
struct LevelArgs {
  // database (device.hpp layout)
  const int32_t *gvbase;
  const int32_t *vadj;
  const uint32_t *adj;
  const uint8_t *vlab;
  const int32_t *gjob;
  // current level: rows of graph g are [gstart[g], gstart[g] + gcount[g])
  const int64_t *gstart;
  const int64_t *gcount;
  const int32_t *pid;
  const uint8_t *vt;
  int32_t S;  // tuple stride in bytes
  const PatMeta *meta;
  int64_t K;  // keys per pattern (root level: keys per job)
  // patterns (root level: jobs) [p_lo, p_hi) handled by this launch, so
  // that (p - p_lo) * K + key fits 32 bits
  int32_t p_lo, p_hi;
  int32_t NEL, NVL;
  // supports: >= 0 a distinct-graph count; < 0 encodes a selected child
  // as -(child + 1) once the host has chosen the children
  int32_t *sup;
  int64_t nkeys;      // P * K
  int32_t *sup_rep;   // rep replicas of nkeys distinct-graph counters
  unsigned long long *proj_rep;  // rep replicas of nkeys embedding counters
  int32_t rep;        // power of two
  const int32_t *minsup_job;
  // (key, support, embeddings) triples of the keys whose support reached the
  // job's min support
  unsigned long long *freq;
  uint32_t *freq_count;
  uint32_t freq_cap;
  unsigned long long *ecand;  // extension candidates enumerated
  int32_t *error_flag;        // set when a graph cannot be deduplicated
  // children: graph g's child rows are [cgstart[g], cgstart[g] + cgcount[g]),
  // reserved from *cursor; ccap rows are allocated
  uint16_t *rowcnt;  // children per current row (root level: per vertex)
  unsigned long long *cursor;
  int64_t ccap;
  int64_t *cgstart;
  int64_t *cgcount;
  int32_t *cpid;
  uint8_t *cvt;
  int32_t CS;  // child tuple stride
  // dynamic LDS: dedup table bits, staged vertex-offset and adjacency words
  int32_t table_bits, stage_v, stage_a;
  int32_t block;                       // threads per workgroup
  unsigned long long *split_passes;    // dedup passes abandoned for a split
};

size_t level_lds_bytes(const LevelArgs &a, bool with_table);
hipError_t launch_reduce(const LevelArgs &a, bool root, hipStream_t s);

hipError_t launch_support(const LevelArgs &a, int n_graphs, bool root, hipStream_t s);
hipError_t launch_write(const LevelArgs &a, int n_graphs, bool root, hipStream_t s);
hipError_t launch_scatter_i32(int32_t *dst, const uint64_t *idx, const int32_t *val, int64_t n, hipStream_t s);
hipError_t launch_fill_i32(int32_t *dst, int64_t n, int32_t v, hipStream_t s);

}  // namespace gm
