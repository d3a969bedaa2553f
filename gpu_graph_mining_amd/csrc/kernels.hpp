// kernels.hpp -- launch interface of the level-step kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace gm {

// Rows per chunk (nominal): one wave of k_level / k_write owns a chunk,
// whose bounds are moved forward to the next segment start.  Large levels
// take kChunkRows; smaller ones shorter chunks (down to one block of 64
// rows), so that a level has enough waves to fill the device.
constexpr int kChunkRows = 256;
constexpr int kMinChunkRows = 64;
constexpr int64_t kLevelWaves = 16384;  // waves a level step aims for
// Counter keys of one level piece at most (16 bytes per key and replica): a
// level with more patterns x keys per pattern is mined in pieces.
constexpr int64_t kPieceKeys = 1ll << 26;

// Statistics counters that every wave bumps are spread over kSpread cache
// lines: atomics on one address serialise across the device.
constexpr int kSpread = 64;
constexpr int kSpreadStride = 16;  // 64-bit words per line
constexpr int kSpreadCand = 0;     // extension candidates enumerated (listed + pruned)
constexpr int kSpreadFill = 1;     // largest count-table fill of a chunk (atomicMax)
// (slot 2 unused)
constexpr int kSpreadPruned = 3;   // candidates dropped by the 2-edge path rule
constexpr int kSpreadStart = 4;    // ~(earliest start) of the extension + support stage (atomicMax)
constexpr int kSpreadEnd = 5;      // latest end of the stage (atomicMax); device wall clock
constexpr int kSpreadWritten = 6;  // child rows written by k_write
constexpr int kSpreadFailed = 7;   // chunks recounted by k_level_fix

// One chunk of a level's rows: [start, end) after alignment to segment
// starts (a segment is the rows of one pattern in one graph; at the root,
// the vertices of one graph).
struct alignas(16) ChunkInfo {
  int64_t start, end;
  uint32_t flags;  // kChunkFailed
  uint32_t nkeys;  // (key, embeddings, item sum) records the chunk left for k_write
  int32_t p_lo;    // their keys are (pattern - p_lo) * K + key
  uint32_t pad;
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
  // (pid, rgraph, vt) whose list regions hold rbound entries from rlist;
  // root level: the database's vertices (lists at their adjacency offsets)
  int64_t row_lo, row_hi;
  const int32_t *pid;
  const int32_t *rgraph;
  const uint8_t *vt;
  const uint32_t *rbound;
  const unsigned long long *rlist;
  int32_t S;  // tuple stride in bytes
  const PatMeta *meta;  // meta_bytes(S) bytes per pattern (the wide layout when S = 64)
  // path pruning masks (2 * mask_stride words per pattern; 0: none):
  // fwd[0] new arms at the rightmost vertex, fwd[1 + i] at rm[i].from
  // (bit elabel * NVL + label), bwd[1 + i] backward edges to rm[i].from (bit
  // elabel); a clear bit marks an extension with an infrequent 2-edge path
  const unsigned long long *pmask;
  int32_t mask_stride;
  int64_t K;              // keys per pattern (root: keys per job)
  int32_t p_lo, p_hi;     // patterns (root: jobs) of this launch; (p - p_lo) * K < 2^32
  int32_t p_base;         // first pattern of the level piece: counter keys are (p - p_base) * K + key
  int32_t NEL, NVL;
  // chunks of this launch: chunk i at chunks[chunk0 + i], n_chunks of them,
  // chunk_rows nominal rows each
  ChunkInfo *chunks;
  int64_t chunk0, n_chunks;
  int32_t chunk_rows;
  // candidate lists: {code, child item count}; row r's list at
  // row_list(r), rccnt[r] entries
  uint2 *cand;
  uint16_t *rccnt;
  // rep replicas (a power of two) of the packed counters acc[p * K + key] =
  // support << 40 | embeddings and of the item sums accb[p * K + key];
  // chunk i counts into replica i mod rep, replica r at offset r * nkeys
  unsigned long long *acc, *accb;
  int32_t rep;
  int32_t table_bits;  // per-wave LDS count table of k_level
  // per chunk, chunk_key_cap records {(pattern - p_lo) * K + key,
  // embeddings, item sum}: the chunk's candidates per key, so that k_write
  // reserves the children's rows without counting them again
  uint4 *chunk_keys;
  int32_t chunk_key_cap;
  uint32_t *fail_list, *fail_count;  // chunks whose count table overflowed
  unsigned long long *spread;
  int32_t *error_flag;
  // k_reduce: supports (>= 0), then child ids (-(child + 1)) once chosen
  int32_t *sup;
  int64_t nkeys;
  const int32_t *minsup_job;
  unsigned long long *freq;  // (key, support, embeddings, item sum) of frequent keys, in key order
  uint32_t *freq_count;
  uint32_t freq_cap;
  uint32_t *tile_count;  // k_reduce: frequent keys per tile of kReduceTile keys, then the tiles' first slots
  // k_write: child rows; cur_rows / cur_list[c * rep + r] are the first row
  // and the first list entry of child c for the chunks of replica r, and
  // cursor[c * rep + r] = rows << 32 | list entries reserved so far
  unsigned long long *cur_rows, *cur_list, *cursor;
  int32_t *cpid;
  int32_t *crgraph;
  uint8_t *cvt;
  uint32_t *cbound;
  unsigned long long *crlist;
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
// The level's supports and its frequent-key list in key order: a pass that
// counts the frequent keys of each tile of kReduceTile keys (and clears the
// other keys' replicas), a scan of the tile counts, and a pass that writes
// each tile's records from its first slot on.  tile_count holds
// reduce_tiles(nkeys) entries.
constexpr int kReduceTile = 256;
__host__ __device__ inline int64_t reduce_tiles(int64_t nkeys) { return (nkeys + kReduceTile - 1) / kReduceTile; }
hipError_t launch_reduce(const StepArgs &a, bool root, hipStream_t s);
// Children of every chunk in chunks[0, n_chunks) (the whole level).
hipError_t launch_write(const StepArgs &a, bool root, hipStream_t s);
// Copy `bytes` (a multiple of 8) between device memory and mapped pinned
// host memory with a kernel on stream s.
hipError_t launch_copy8(void *dst, const void *src, size_t bytes, hipStream_t s);
// Copy the children's upload (bytes, a multiple of 8) and, for the F
// frequent keys (the keys of the frequent list a.freq, in its order):
// sup[key] = -(fchild[i] + 1) for the chosen ones (fchild >= 0), their
// (child, replica) cursors from row_base / list_base[child] and the
// replicas' counts, and clear the keys' replicas.
hipError_t launch_cursors(const StepArgs &a, void *dst, const void *src, size_t bytes, const int32_t *fchild,
                          const unsigned long long *row_base, const unsigned long long *list_base, int64_t F,
                          hipStream_t s);
// Gather selected graphs of a resident packed database into a new one
// (k_gather_graphs): per selected graph t, its source graph, first vertex
// and first adjacency entry in the output, and its job.
struct GatherArgs {
  const int32_t *gvbase, *vadj;
  const uint32_t *adj;
  const uint8_t *vlab;
  const int32_t *sel, *dvb, *dab, *djob;
  int64_t n_sel;
  int32_t *out_gvbase, *out_vadj, *out_gjob, *out_vgraph;
  uint32_t *out_adj;
  uint8_t *out_vlab;
};
hipError_t launch_gather_graphs(const GatherArgs &g, hipStream_t s);
// Zero `bytes` (a multiple of 8) of device memory with a kernel.
hipError_t launch_fill8(void *dst, size_t bytes, hipStream_t s);
// The static workspace of the entry point that allocates nothing: a
// file-scope device array of kStaticScratchBytes in kernels.hip; *ptr gets
// its address on the current device (looked up once per device).
constexpr size_t kStaticScratchBytes = 64ull << 20;
hipError_t static_scratch(void **ptr, size_t *bytes);
// Minimum DFS code check on the device (canon.hip): one wave per code, the
// pattern graph and the embedding lists in LDS.  Codes beyond these sizes are
// left to the host check (kCanonHost).
constexpr int kCanonV = 40;     // pattern vertices (+1 for the grown code's next vertex)
constexpr int kCanonE = 64;     // pattern edges (64-bit used-edge masks)
constexpr int kCanonCap = 64;   // embeddings of the grown code held at once
constexpr int kCanonNotMin = 0, kCanonMin = 1, kCanonHost = 2, kCanonBad = 3;
struct CanonArgs {
  // a later level's frequent records (key, support, embeddings, item sum):
  // the answer + 1 for record i goes to bits 32-39 of its support word
  unsigned long long *freq;
  const uint32_t *freq_count;
  uint32_t freq_cap;
  // the level's patterns: nv from meta, len tuples each packed from | to << 8
  // | elabel << 16 | tolabel << 24 (dense labels), S vertex labels each
  // (S: the level's tuple stride; meta_bytes(S) bytes of metadata each)
  const PatMeta *meta;
  int32_t p_base;  // record keys are (pattern - p_base) * K + key
  const uint32_t *pcode;
  const uint8_t *pvlab;
  int32_t len;
  int64_t K;
  int32_t S, NEL, NVL;
  // explicit codes: n codes of len (from, to, fromlabel, elabel, tolabel)
  // tuples; out[i] gets the answer (kCanonBad: a malformed tuple)
  const int32_t *codes;
  int64_t n;
  uint8_t *out;
};
// grid: waves over the frequent records (any count covers them)
hipError_t launch_canon_level(const CanonArgs &c, int grid, hipStream_t s);
hipError_t launch_canon_list(const CanonArgs &c, hipStream_t s);
// Zero the per-level counters (frequent-key count, failure list, spread
// statistics but the written-row slots).
hipError_t launch_clear_counters(const StepArgs &a, hipStream_t s);

}  // namespace gm
