// kernels.hpp -- launch interface of the level-step kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace gm {

constexpr int kBlock = 256;      // threads per workgroup of the level kernels
// Support kernel: LDS dedup table of 2^kTableBits 32-bit keys; a pass whose
// table passes kTableFill distinct keys is split.
#ifndef GM_TABLE_BITS
#define GM_TABLE_BITS 12
#endif
constexpr int kTableBits = GM_TABLE_BITS;
constexpr int kTableCap = 1 << kTableBits;
constexpr int kTableFill = kTableCap * 3 / 4;

// Everything one level-step kernel reads or writes.  Plain device pointers.
struct LevelArgs {
  // database (device.hpp layout)
  const int32_t *gvbase;
  const int32_t *vadj;
  const uint32_t *adj;
  const uint8_t *vlab;
  const int32_t *gjob;
  // current level: rows of graph g are [grow[g], grow[g+1])
  const int64_t *grow;
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
  const int32_t *minsup_job;
  // (pattern, key) indexes whose support reached the job's min support
  unsigned long long *freq;
  uint32_t *freq_count;
  uint32_t freq_cap;
  unsigned long long *ecand;  // extension candidates enumerated
  int32_t *error_flag;        // set when a graph cannot be deduplicated
  // children
  int64_t *ccount;      // child rows per graph
  const int64_t *cgrow; // child row offsets per graph
  int32_t *cpid;
  uint8_t *cvt;
  int32_t CS;  // child tuple stride
};

hipError_t launch_support(const LevelArgs &a, int n_graphs, bool root, hipStream_t s);
hipError_t launch_children(const LevelArgs &a, int n_graphs, bool root, bool write, hipStream_t s);
// out[0..n] = exclusive scan of in[0..n), out[n] = total.  tmp holds
// scan_tmp_elems(n) values.
hipError_t launch_exclusive_scan(const int64_t *in, int64_t n, int64_t *out, int64_t *tmp, hipStream_t s);
int64_t scan_tmp_elems(int64_t n);
hipError_t launch_gather_i32(const int32_t *src, const uint64_t *idx, int64_t n, int32_t *dst, hipStream_t s);
hipError_t launch_scatter_i32(int32_t *dst, const uint64_t *idx, const int32_t *val, int64_t n, hipStream_t s);
hipError_t launch_fill_i32(int32_t *dst, int64_t n, int32_t v, hipStream_t s);
hipError_t launch_pattern_hist(const int32_t *pid, int64_t n, unsigned long long *hist, hipStream_t s);

}  // namespace gm
