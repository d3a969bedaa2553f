// kernels.hpp -- launch interface of the level-step kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace gm {

// Rows per chunk (nominal): one wave of k_level / k_write owns a chunk,
// whose bounds are moved forward to the next segment start.
constexpr int kChunkRows = 512;
constexpr int kMaxTupleWords = kMaxPatV / 4;  // 8

// Statistics counters that every wave bumps are spread over kSpread cache
// lines: atomics on one address serialise across the device.
constexpr int kSpread = 64;
constexpr int kSpreadStride = 16;  // 64-bit words per line
constexpr int kSpreadCand = 0;     // extension candidates enumerated (listed + pruned)
constexpr int kSpreadFill = 1;     // largest count-table fill of a chunk (atomicMax)
constexpr int kSpreadBound = 2;    // sum of the listed candidates' child bounds
constexpr int kSpreadPruned = 3;   // candidates dropped by the 2-edge path rule
constexpr int kSpreadStart = 4;    // ~(earliest start) of the extension + support stage (atomicMax)
constexpr int kSpreadEnd = 5;      // latest end of the stage (atomicMax); device wall clock
constexpr int kSpreadWritten = 6;  // child rows written by k_write
constexpr int kSpreadFailed = 7;   // chunks recounted by k_level_fix

// One chunk of a level's rows: [start, end) after alignment to segment
// starts (a segment is the rows of one pattern in one graph; at the root,
// the vertices of one graph), and the chunk's candidate-list region.
struct alignas(16) ChunkInfo {
  int64_t start, end;
  unsigned long long list_base;  // first list entry of the chunk's rows
  uint32_t flags;                // kChunkFailed
  uint32_t nkeys;                // (key, embeddings) records the chunk left for k_write
  int32_t p_lo;                  // their keys are (pattern - p_lo) * K + key
  uint32_t pad[3];
};
constexpr uint32_t kChunkFailed = 1;  // the LDS count table overflowed; k_level_fix counted it

// Everything the level-step kernels read or write.  Plain device pointers.
struct StepArgs {
  // database (device.hpp layout)
  const int32_t *gvbase;
  const int32_t *vadj;
  const uint32_t *adj;
  const uint8_t *vlab;
  const int32_t *gjob;
  const int32_t *vgraph;  // graph of each vertex
  // rows of this launch: [row_lo, row_hi); later levels: pattern-major rows
  // (pid, rgraph, vt) whose list regions hold rbound entries each; root
  // level: the database's vertices
  int64_t row_lo, row_hi;
  const int32_t *pid;
  const int32_t *rgraph;
  const uint8_t *vt;
  const uint32_t *rbound;
  int32_t S;  // tuple stride in bytes
  const PatMeta *meta;
  // path pruning masks (2 * mask_stride words per pattern; 0: none):
  // fwd[0] new arms at the rightmost vertex, fwd[1 + i] at rm[i].from
  // (bit elabel * NVL + label), bwd[1 + i] backward edges to rm[i].from (bit
  // elabel); a clear bit marks an extension with an infrequent 2-edge path
  const unsigned long long *pmask;
  int32_t mask_stride;
  int64_t K;              // keys per pattern (root: keys per job)
  int32_t p_lo, p_hi;     // patterns (root: jobs) of this launch; (p - p_lo) * K < 2^32
  int32_t NEL, NVL;
  // chunks of this launch: chunk i at chunks[chunk0 + i], n_chunks of them
  ChunkInfo *chunks;
  int64_t chunk0, n_chunks;
  // candidate lists: {code, child bound}; row r's list at its chunk's
  // list_base + the prefix of its chunk's rbound, rccnt[r] entries
  uint2 *cand;
  unsigned long long *list_cursor;
  unsigned long long list_cap;
  uint16_t *rccnt;
  // packed counters acc[p * K + key] = support << 40 | embeddings
  unsigned long long *acc;
  int32_t table_bits;  // per-wave LDS count table of k_level
  // per chunk, chunk_key_cap records {(pattern - p_lo) * K + key,
  // embeddings}: the chunk's candidates per key, so that k_write reserves
  // the children's rows without counting them again
  uint2 *chunk_keys;
  int32_t chunk_key_cap;
  uint32_t *fail_list, *fail_count;  // chunks whose count table overflowed
  unsigned long long *spread;
  int32_t *error_flag;
  // k_reduce: supports (>= 0), then child ids (-(child + 1)) once chosen
  int32_t *sup;
  int64_t nkeys;
  const int32_t *minsup_job;
  unsigned long long *freq;  // (key, support, embeddings) of frequent keys
  uint32_t *freq_count;
  uint32_t freq_cap;
  // k_write: child rows; cursor[c] is the next row of child c
  unsigned long long *cursor;
  int32_t *cpid;
  int32_t *crgraph;
  uint8_t *cvt;
  uint32_t *cbound;
  int32_t CS;  // child tuple stride
  int64_t ccap;
};

// Dynamic LDS of k_level for table_bits, and the per-chunk key records its
// table leaves at most.
size_t level_lds_bytes(int table_bits);
int level_chunk_keys(int table_bits);
hipError_t launch_level(const StepArgs &a, bool root, hipStream_t s);
// Recount the chunks k_level listed as failed (persistent grid; cheap when
// none failed).
hipError_t launch_level_fix(const StepArgs &a, bool root, hipStream_t s);
hipError_t launch_reduce(const StepArgs &a, bool root, hipStream_t s);
// Children of every chunk in chunks[0, n_chunks) (the whole level).
hipError_t launch_write(const StepArgs &a, bool root, hipStream_t s);
// Copy `bytes` (a multiple of 8) between device memory and mapped pinned
// host memory with a kernel on stream s.
hipError_t launch_copy8(void *dst, const void *src, size_t bytes, hipStream_t s);
// Copy the children's upload (bytes, a multiple of 8), set
// sup[keys[i]] = -(i + 1) and cursor[i] = bases[i] for i < C, in one kernel.
hipError_t launch_upload_scatter(void *dst, const void *src, size_t bytes, int32_t *sup, const uint64_t *keys,
                                 unsigned long long *cursor, const unsigned long long *bases, int64_t C,
                                 hipStream_t s);
// Zero the per-level counters (frequent-key count, failure list, spread
// statistics, list cursor).
hipError_t launch_clear_counters(const StepArgs &a, hipStream_t s);

}  // namespace gm
