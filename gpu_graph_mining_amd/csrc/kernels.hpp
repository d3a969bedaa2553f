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
constexpr int kScanTileGraphs = 1024;  // graphs per tile of the placement scan
constexpr int kScanSingleGraphs = 16 * kScanTileGraphs;  // up to this many graphs: one-workgroup placement scan
constexpr int kRowWaveRows = 64;       // rows per workgroup of the row-parallel children passes
constexpr int kTileStride = 16;        // 64-bit words between tile sums: one cache line each (atomics on a line serialise)

// Statistics counters that every workgroup bumps are spread over kSpread
// cache lines: atomics on one address serialise across the device.
constexpr int kSpread = 64;
constexpr int kSpreadStride = 16;  // 64-bit words per line
constexpr int kSpreadCand = 0;     // extension candidates enumerated
constexpr int kSpreadFill = 1;     // largest dedup table fill (atomicMax)
constexpr int kSpreadBound = 2;    // sum of the candidates' child bounds
constexpr int kSpreadPruned = 3;   // candidates dropped by the 2-edge path rule
constexpr int kSpreadStart = 4;    // ~(earliest start) of the extension + support stage (atomicMax)
constexpr int kSpreadEnd = 5;      // latest end of the stage (atomicMax); device wall clock

// Everything one level-step kernel reads or writes.  Plain device pointers.
struct LevelArgs {
  // database (device.hpp layout)
  const int32_t *gvbase;
  const int32_t *vadj;
  const uint32_t *adj;
  const uint8_t *vlab;
  const int32_t *gjob;
  const int32_t *vgraph;  // graph of each vertex
  // current level: rows of graph g are [gstart[g], gstart[g] + gcount[g])
  const int64_t *gstart;
  const int64_t *gcount;
  const int32_t *pid;
  const int32_t *rgraph;  // graph of each row
  int64_t n_rows;
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
  int64_t nkeys;                 // P * K
  // rep replicas of nkeys packed counters: distinct graphs << 40 | embeddings
  unsigned long long *acc_rep;
  int32_t rep;                   // power of two
  const int32_t *minsup_job;
  // (key, support, embeddings) of the keys whose support reached the job's
  // min support
  unsigned long long *freq;
  uint32_t *freq_count;
  uint32_t freq_cap;
  // statistics spread over kSpread cache lines (slot i of line j at
  // spread[j * kSpreadStride + i]): candidates enumerated, largest table
  // fill, sum of the candidates' child bounds
  unsigned long long *spread;
  int32_t *error_flag;           // set when a graph cannot be deduplicated
  unsigned long long *split_passes;  // dedup passes abandoned for a split
  // candidate lists of this level ({code, child bound} entries, one per
  // item, code kNoCand where the item extends nothing): row r's list is
  // cand[gcbase[g] + rcoff[r] ...] with rccnt[r] entries (root level:
  // cand[vadj[v] ...] for vertex v); gcbase[g] < 0: no lists
  int32_t use_lists;
  int32_t lists_ready;  // k_expand listed this level's candidates already
  uint2 *cand;
  const int64_t *gcbase;
  const uint32_t *rcoff;
  uint16_t *rccnt;
  // children: graph g's child rows are [cgstart[g], cgstart[g] + cgcount[g])
  // (counted by k_children<.., false>, placed by k_scan); ccap rows are
  // allocated
  uint16_t *rowcnt;   // enumeration mode: children per row (root: per vertex)
  uint32_t *rowbsum;  // enumeration mode: children's bound sum per row
  unsigned long long *gbsum;   // children's bound sum per graph
  unsigned long long *cursor;  // total child rows (written by k_scan)
  int64_t ccap;
  int64_t *cgstart;
  int64_t *cgcount;
  int32_t *cpid;
  int32_t *crgraph;
  uint8_t *cvt;
  int32_t CS;  // child tuple stride
  // next level's candidate lists: per-graph region base ngcbase (-1: none)
  // within ncand_cap entries, per-child offset nrcoff; *ncand_cursor gets
  // the entries all graphs asked for
  int32_t next_lists;
  int64_t ncand_cap;
  unsigned long long *ncand_cursor;
  int64_t *ngcbase;
  uint32_t *nrcoff;
  // dynamic LDS: dedup table bits, per-thread slot words (tuple_words of
  // tuple, then the rightmost path), staged vertex-offset and adjacency words
  int32_t table_bits, slot_words, tuple_words, stage_v, stage_a;
  int32_t rm_words, src_words;  // k_expand row records: rightmost path, item prefix
  // Path pruning masks (k_expand), 2 * mask_stride words per pattern:
  // fwd[0] new edges at the rightmost vertex, fwd[1 + i] new edges at
  // rightmost-path vertex rm[i].from (bit elabel * NVL + label of the new
  // neighbour); bwd[1 + i] backward edges to rm[i].from (bit elabel).  A
  // clear bit marks an extension one of whose 2-edge sub-paths is
  // infrequent.  mask_stride 0: no pruning.
  const unsigned long long *pmask;
  int32_t mask_stride;
  int32_t block;  // threads per workgroup
  // Graphs with more than big_rows rows at this level run in a second launch
  // of big_grid workgroups of big_block threads over the list big[0, *nbig)
  // that the previous level's k_scan built; k_scan writes the next level's
  // list to nbig_list / *nnbig.
  const int32_t *big;
  const unsigned long long *nbig;
  int32_t big_rows, big_block, big_grid;
  int32_t big_table_bits;  // dedup table of the large-graph launch (>= table_bits)
  int32_t wave_grid;  // workgroups of the persistent per-wave root kernel
  int32_t *nbig_list;
  unsigned long long *nnbig;
  // Row-parallel children passes (later levels whose candidates k_expand
  // listed): per wave of kRowWaveRows rows, the child rows and list entries
  // (wrows, wents); per tile of 64 waves their sums, then exclusive offsets
  // (trows, tents: tile i at [i * kTileStride]; n_wtiles tiles, zeroed by
  // k_reduce)
  int32_t rows_mode;
  int32_t n_wtiles;
  // speculative levels: the count pass runs before the child ids are
  // scattered, selecting keys by support >= the job's min support
  // (count_by_support), and launch_write then skips it (count_done)
  int32_t count_by_support, count_done;
  unsigned long long *wrows, *wents, *trows, *tents;
  // placement scan: per tile of 1024 graphs (row-parallel passes: waves),
  // sums then exclusive offsets
  long long *tile_rows;
  unsigned long long *tile_ents;
  long long *tile_big;
  long long *tile_max;
};

size_t level_lds_bytes(const LevelArgs &a, bool support);
hipError_t launch_reduce(const LevelArgs &a, bool root, hipStream_t s);

hipError_t launch_support(const LevelArgs &a, int n_graphs, bool root, hipStream_t s);
size_t expand_lds_bytes(const LevelArgs &a);
hipError_t launch_expand(const LevelArgs &a, hipStream_t s);
// Copy `bytes` (a multiple of 8) between device memory and mapped pinned
// host memory with a kernel on stream s.
hipError_t launch_copy8(void *dst, const void *src, size_t bytes, hipStream_t s);
// Write v to *flag (mapped pinned host memory) once the stream's earlier
// work is done and visible to the host.
hipError_t launch_signal(unsigned long long *flag, unsigned long long v, hipStream_t s);
// The row-parallel count pass (k_rows<*, false>) and its placement scan
// (k_rtop), ahead of the write.
hipError_t launch_count(const LevelArgs &a, bool root, hipStream_t s);
// Copy the children's upload (bytes, a multiple of 8) and set
// sup[keys[i]] = vals[i] for i < C, in one kernel.
hipError_t launch_upload_scatter(void *dst, const void *src, size_t bytes, int32_t *sup, const uint64_t *keys,
                                 const int32_t *vals, int64_t C, hipStream_t s);
// Zero the per-level counters (the write pass's scan also does it for the
// next level).
hipError_t launch_clear_counters(const LevelArgs &a, hipStream_t s);
hipError_t launch_write(const LevelArgs &a, int n_graphs, bool root, hipStream_t s);

}  // namespace gm
