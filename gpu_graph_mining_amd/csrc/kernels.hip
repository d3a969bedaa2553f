// kernels.hip -- HIP kernels of the level step (gfx950, wave64).
//
// A level's rows (embeddings) are pattern-major: the rows of one pattern are
// contiguous and, within them, the rows of one database graph are
// contiguous (a *segment*).  At the root the rows are the database's
// vertices, the "pattern" of a vertex is its job and a segment is a graph.
// Support = number of segments of a pattern that have the extension, so
// every support question is answered inside a segment, and a run of
// consecutive segments (a chunk) touches only a few patterns' keys.
//
// Per level:
//   k_level      one wave per chunk of rows (chunk_rows of them, moved forward
//                to segment starts): every lane enumerates one row's
//                extension candidates, appends them to the row's list and
//                counts them in a per-wave LDS table keyed by (pattern,
//                key): embeddings always, support once per (segment, key)
//                through a 64-bit mask of the block's segments per entry.
//                The table is committed to the packed device counters once
//                per chunk -- one device atomic per (chunk, key) instead of
//                one per (graph, key) -- into replica (chunk mod R) of the
//                counters, so that keys every chunk hits (the root's) are
//                not one hot address.  A chunk whose table fills up is
//                listed and recounted by k_level_fix from its lists.
//   k_reduce     sums the replicas of every key, lists the frequent keys
//                (clearing the others);
//   k_cursors    (after the host chose the children) points the chosen
//                keys at their child ids and gives every (child, replica)
//                its row and list sub-range from the replica's counts,
//                clearing the frequent keys' replicas;
//   k_write      one wave per chunk: reserves each child's rows of the chunk
//                with one atomic on the (child, replica) cursors, and writes
//                them in list order, so the child's rows stay pattern-major
//                and graph-contiguous, each with its list offset.
//
// Extension rules (pattern side in PatMeta, graph side from the packed
// adjacency):
//   backward  rightmost vertex -> image of rightmost-path vertex f = from of
//             rightmost-path edge i >= 1, edge not yet in the pattern, and
//             el(i) < e || (el(i) == e && tolabel(i) <= label(rightmost))
//             (src/globals/graph_types.cpp:389-413, gspan.cpp:220-225);
//   pure fwd  rightmost vertex -> unvisited w, label(w) >= minlabel
//             (graph_types.cpp:337-355, gspan.cpp:233-237);
//   rm fwd    from of rightmost-path edge i -> unvisited w,
//             label(w) >= minlabel and
//             el(i) < e || (el(i) == e && tolabel(i) <= label(w))
//             (graph_types.cpp:305-323, gspan.cpp:240-246);
//   root      u -> w with label(u) <= label(w) (graph_types.cpp:365-375).
// Because the database graphs are simple, "edge not in the embedding" is a
// pattern-level fact (bwd_mask); "vertex not in the embedding" is a byte
// compare over the tuple held in registers.
//
// Items and bounds: a row's items are the adjacency entries of its
// rightmost vertex, then of each rightmost-path source (rm[0].from, ...,
// the root).  Every candidate is one item, so a row's item count bounds its
// candidates and sizes its list region.  The enumeration derives each
// child's item count from the parent's:
//   backward           the parent's (same path, same rightmost);
//   pure fwd to w      the parent's, plus deg(w);
//   fwd from source s  sources s..root of the parent, plus deg(w);
//   root edge (u, w)   deg(u) + deg(w).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "device.hpp"
#include "kernels.hpp"

namespace gm {

namespace {

constexpr uint32_t kEmpty32 = 0xFFFFFFFFu;
constexpr int kSupShift = 40;  // packed counters: support << 40 | embeddings
// k_write: children per chunk in one pass.  A chunk with more frequent
// children than that is counted again in passes split by child residue;
// the small table (1.75 KB with its row and list bases; 4 KB of LDS per
// wave in all) lets 8 waves share a SIMD, which the write's dependent
// gathers need more than they need one pass per chunk.
constexpr int kWriteTableBits = 6;
// The root level's chunks span many graphs and every label triple: there a
// chunk often holds more than 64 frequent children, and a 256-entry table
// (5 waves per SIMD) beats the residue passes.
constexpr int kWriteTableBitsRoot = 8;
constexpr int kFixTableBits = 11;   // k_level_fix: (segment, key) pairs per pass

// ----------------------------------------------------------- wave helpers
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T>
__device__ __forceinline__ T wave_exclusive_scan(T x, T &total) {
  const int lane = threadIdx.x & 63;
  T v = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const T y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  total = __shfl(v, 63, 64);
  return v - x;
}

template <class T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, 64));
  return x;
}

__device__ __forceinline__ uint64_t lanes_below(int lane) { return (1ull << lane) - 1ull; }

__device__ __forceinline__ uint32_t hash_slot32(uint32_t k, int bits) { return (k * 0x9E3779B1u) >> (32 - bits); }

// Wall-clock stamps of the extension + support stage: the first chunks
// record ~start (the max of complements is the earliest start), every
// chunk its end; the host reads them from the spread lines.
__device__ __forceinline__ void stamp_start(const StepArgs &a, int64_t ci) {
  if (ci < kSpread && threadIdx.x == 0)
    atomicMax(&a.spread[ci * kSpreadStride + kSpreadStart], ~(unsigned long long)__builtin_amdgcn_s_memrealtime());
}
__device__ __forceinline__ void stamp_end(const StepArgs &a, int64_t ci) {
  if (threadIdx.x == 0)
    atomicMax(&a.spread[(ci & (kSpread - 1)) * kSpreadStride + kSpreadEnd],
              (unsigned long long)__builtin_amdgcn_s_memrealtime());
}

// Candidate list entries: {code, child item count}.
//   root codes   key | w << 24          (key < 255^3 < 2^24)
//   later codes  key | (w + 1) << 23    (key < 64 * 255 * 256 < 2^23,
//                                        w = -1 for backward edges)
__device__ __forceinline__ uint32_t cand_code(bool root, int key, int w) {
  return root ? ((uint32_t)key | ((uint32_t)w << 24)) : ((uint32_t)key | ((uint32_t)(w + 1) << 23));
}
__device__ __forceinline__ void cand_decode(bool root, uint32_t c, int &key, int &w) {
  if (root) {
    key = (int)(c & 0xFFFFFFu);
    w = (int)(c >> 24);
  } else {
    key = (int)(c & 0x7FFFFFu);
    w = (int)(c >> 23) - 1;
  }
}

// Row r starts a segment: the first row of the launch, or its pattern or
// graph differs from the previous row's (root: its graph).
template <bool ROOT>
__device__ __forceinline__ bool seg_start(const StepArgs &a, int64_t r) {
  if (r <= a.row_lo) return true;
  if (ROOT) return a.vgraph[r] != a.vgraph[r - 1];
  return a.pid[r] != a.pid[r - 1] || a.rgraph[r] != a.rgraph[r - 1];
}

// First segment start at or after row x (row_hi if none); wave-uniform.
template <bool ROOT>
__device__ __forceinline__ int64_t seg_at_or_after(const StepArgs &a, int64_t x) {
  if (x <= a.row_lo) return a.row_lo;
  for (int64_t base = x; base < a.row_hi; base += 64) {
    const int64_t r = base + (threadIdx.x & 63);
    const bool f = r < a.row_hi && seg_start<ROOT>(a, r);
    const uint64_t m = __ballot(f);
    if (m) return base + __ffsll((long long)m) - 1;
  }
  return a.row_hi;
}

// Chunk ci of this launch: [start, end) moved forward to segment starts.
template <bool ROOT>
__device__ __forceinline__ void chunk_bounds(const StepArgs &a, int64_t ci, int64_t &start, int64_t &end) {
  const int64_t nom0 = a.row_lo + ci * a.chunk_rows;
  const int64_t nom1 = min(a.row_hi, nom0 + a.chunk_rows);
  start = seg_at_or_after<ROOT>(a, nom0);
  end = nom1 >= a.row_hi ? a.row_hi : seg_at_or_after<ROOT>(a, nom1);
}

// Items of row r (the size of its list region): root, the vertex's
// degree; later levels, the count the parent level recorded.
template <bool ROOT>
__device__ __forceinline__ uint32_t row_items(const StepArgs &a, int64_t r) {
  if (ROOT) return (uint32_t)(a.vadj[r + 1] - a.vadj[r]);
  return a.rbound[r];
}

// First list entry of row r: root, the vertex's adjacency offset; later
// levels, the offset the parent level's k_write gave the row.
template <bool ROOT>
__device__ __forceinline__ unsigned long long row_list(const StepArgs &a, int64_t r) {
  if (ROOT) return (unsigned long long)a.vadj[r];
  return a.rlist[r];
}

// Block of the chunk starting at row b (< end): up to 64 rows, cut at the
// start of the last segment when that segment continues past the window;
// a window that is one segment continuing past it is taken whole and its
// segment carried into the next block.  Returns the block's rows; sets the
// segment-start ballot and whether the last segment is carried.
template <bool ROOT>
__device__ __forceinline__ int block_rows(const StepArgs &a, int64_t b, int64_t end, uint64_t &starts,
                                          bool &carry) {
  const int lane = threadIdx.x & 63;
  const int ws = (int)min<int64_t>(64, end - b);
  const int64_t r = b + lane;
  starts = __ballot(lane < ws && seg_start<ROOT>(a, r));
  carry = false;
  const bool cont = b + ws < end && !seg_start<ROOT>(a, b + ws);
  if (!cont) return ws;
  const uint64_t later = starts & ~1ull;
  if (later) return 63 - __clzll((long long)later);
  carry = true;
  return ws;
}

// Lane t's segment head within the block: the last segment start at or
// before t (0 when the segment began in an earlier block).
__device__ __forceinline__ int head_lane(uint64_t starts, int lane) {
  const uint64_t upto = starts & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
  return upto ? 63 - __clzll((long long)upto) : 0;
}

// ----------------------------------------------------------- tuples
// A row's vertex tuple in registers: byte i of word i / 4 is the graph-local
// id of pattern vertex i.
template <int TW>
struct Tuple {
  uint32_t w[TW > 0 ? TW : 1];
  int nv;
  // the words only (nv set apart): they depend on the row alone
  __device__ __forceinline__ void load(const uint8_t *vt, int64_t r, int S) {
    const uint32_t *p = reinterpret_cast<const uint32_t *>(vt + r * S);
#pragma unroll
    for (int q = 0; q < TW; ++q) w[q] = p[q];
  }
  // 0x80 in each byte of word q below nv (computed, not kept: registers)
  __device__ __forceinline__ uint32_t live(int q) const {
    const int valid = nv - 4 * q;
    return valid >= 4 ? 0x80808080u : valid <= 0 ? 0u : (0x80808080u >> (8 * (4 - valid)));
  }
  // byte i: masks instead of selects, so that the compiler does not turn
  // the words into an indexed (scratch) array
  __device__ __forceinline__ int get(int i) const {
    uint32_t x = 0;
    const int wi = i >> 2;
#pragma unroll
    for (int q = 0; q < TW; ++q) x |= w[q] & (0u - (uint32_t)(wi == q));
    return (int)((x >> ((i & 3) * 8)) & 0xFFu);
  }
  // Pattern vertex mapped to graph vertex v, or -1.  A zero byte of
  // word ^ (v in every byte) marks a match; a borrow only reaches higher
  // bytes, and the padding bytes (at or past nv) are the highest.
  __device__ __forceinline__ int find(int v) const {
    const uint32_t pat = (uint32_t)v * 0x01010101u;
    int f = -1;
#pragma unroll
    for (int q = TW - 1; q >= 0; --q) {
      const uint32_t x = w[q] ^ pat;
      const uint32_t z = (x - 0x01010101u) & ~x & live(q);
      if (z) f = 4 * q + (__builtin_ctz(z) >> 3);
    }
    return f;
  }
};

// ----------------------------------------------------------- LDS tables
// k_level count table (per wave, cap = 2^bits entries): key (pattern -
// p_lo) * K + extension key, support, embeddings, the candidates' item sum,
// and the mask of the block's segments (by head lane) that hold the key.
// The support is settled at each block's end from the mask (a candidate
// only sets its segment's bit); bit 31 of sup then marks a mask whose bit 0
// belongs to a segment carried into the next block (counted already).
struct CountTable {
  uint32_t *key, *sup, *emb, *bsum;
  unsigned long long *mask;
  int *fill;
  int *failed;
  uint32_t cap, limit;
  int bits;
};

__device__ __forceinline__ CountTable count_table(uint32_t *dyn, int bits, int *scalars) {
  CountTable t;
  t.bits = bits;
  t.cap = 1u << bits;
  t.limit = t.cap - t.cap / 8;
  t.key = dyn;
  t.sup = dyn + t.cap;
  t.emb = dyn + 2 * t.cap;
  t.bsum = dyn + 3 * t.cap;
  t.mask = reinterpret_cast<unsigned long long *>(dyn + 4 * t.cap);  // 8-byte aligned
  t.fill = scalars;
  t.failed = scalars + 1;
  return t;
}

// Count one candidate (child item count `bound`) of a row whose segment
// head lane is `head`.
__device__ __forceinline__ void count_insert(const CountTable &t, uint32_t gk, int head, uint32_t bound) {
  if (__hip_atomic_load(t.failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) return;
  uint32_t h = hash_slot32(gk, t.bits);
  for (uint32_t probe = 0; probe < t.cap; ++probe) {
    uint32_t cur = __hip_atomic_load(&t.key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == kEmpty32) {
      if (__hip_atomic_load(t.fill, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= (int)t.limit) break;
      cur = atomicCAS(&t.key[h], kEmpty32, gk);
      if (cur == kEmpty32) {
        atomicAdd(t.fill, 1);
        cur = gk;
      }
    }
    if (cur == gk) {
      atomicOr(&t.mask[h], 1ull << head);  // results unused: the no-return forms
      atomicAdd(&t.emb[h], 1u);
      atomicAdd(&t.bsum[h], bound);
      return;
    }
    h = (h + 1) & (t.cap - 1);
  }
  __hip_atomic_store(t.failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ------------------------------------------------------------------ LEVEL
// One wave per chunk.  Per block of <= 64 segment-aligned rows, lane t owns
// row b + t: it loads the row (pattern, graph, tuple), then walks its items
// a few per step -- the rightmost vertex's adjacency, then each
// rightmost-path source's (the next source's adjacency range prefetched) --
// applies the extension rules, appends each candidate to the row's list
// region and counts it in the wave's table.
constexpr int kStepItems = 6;  // items of one source a lane takes per step
template <bool ROOT, int TW>
__global__ void __launch_bounds__(64, (TW <= 2 ? 6 : TW <= 8 ? 5 : 4)) k_level(StepArgs a) {  // waves per SIMD the registers allow
  // 16-byte aligned whatever static LDS precedes it: the table's segment
  // masks take 64-bit LDS atomics
  extern __shared__ __attribute__((aligned(16))) uint32_t dyn[];
  __shared__ int s_scal[2];
  const int lane = threadIdx.x;
  const int64_t ci = blockIdx.x;
  stamp_start(a, ci);
  int64_t start, end;
  chunk_bounds<ROOT>(a, ci, start, end);
  ChunkInfo *info = a.chunks + a.chunk0 + ci;
  if (start >= end) {
    if (lane == 0) *info = ChunkInfo{start, start, 0u, 0u, 0, 0u};
    return;
  }
  if (lane == 0) *info = ChunkInfo{start, end, 0u, 0u, a.p_lo, 0u};

  const CountTable tab = count_table(dyn, a.table_bits, s_scal);
  for (uint32_t i = lane; i < tab.cap; i += 64) {
    tab.key[i] = kEmpty32;
    tab.sup[i] = 0;
    tab.emb[i] = 0;
    tab.bsum[i] = 0;
    tab.mask[i] = 0;
  }
  if (lane == 0) {
    s_scal[0] = 0;
    s_scal[1] = 0;
  }
  wave_sync();

  const int NEL = a.NEL, NVL = a.NVL;
  const int fwd_base = a.S * NEL;
  const int nel_nvl = NEL * NVL;
  const int ms = a.mask_stride;
  uint32_t n_pruned = 0;        // per lane: at most its rows' items (< 2^32)
  unsigned long long listed = 0;  // candidates listed by the wave (uniform): with the pruned, the enumerated
  int err = 0;
  for (int64_t b = start; b < end;) {
    uint64_t starts;
    bool carry;
    const int nb = block_rows<ROOT>(a, b, end, starts, carry);
    const bool live = lane < nb;
    const int64_t r = b + lane;
    const int head = head_lane(starts, lane);
    const uint32_t rb = live ? row_items<ROOT>(a, r) : 0u;
    const unsigned long long lo = live ? row_list<ROOT>(a, r) : 0ull;
    int cnt = 0;  // candidates listed for this row
    // Each step a lane takes up to kStepItems items of its current source, their
    // adjacency entries loaded together, then evaluates them in order.
    if (ROOT) {
      int32_t ca = 0, ce = 0, lu = 0, job = 0, n = 0;
      if (live) {
        const int32_t g = a.vgraph[r];
        job = a.gjob[g];
        lu = a.vlab[r];
        ca = a.vadj[r];
        n = (int)rb;
        ce = ca + n;
      }
      const uint32_t kb = (uint32_t)((int64_t)(job - a.p_lo) * a.K);
      auto eval = [&](uint32_t e) {
        const int w = e & 0xFF, el = (e >> 8) & 0xFF, nl = (e >> 16) & 0xFF;
        if (lu <= nl) {
          const int key = (lu * NEL + el) * NVL + nl;
          const uint32_t bound = (uint32_t)n + (e >> 24);
          a.cand[lo + cnt] = make_uint2(cand_code(true, key, w), bound);
          ++cnt;
          count_insert(tab, kb + (uint32_t)key, head, bound);
        }
      };
      while (__ballot(ca < ce)) {
        if (ca < ce) {
          const int nstep = min(kStepItems, ce - ca);
          uint32_t e[kStepItems];
#pragma unroll
          for (int u = 0; u < kStepItems; ++u) e[u] = u < nstep ? a.adj[ca + u] : 0u;
#pragma unroll
          for (int u = 0; u < kStepItems; ++u)
            if (u < nstep) eval(e[u]);
          ca += nstep;
        }
      }
    } else {
      // the level's metadata layout follows its tuple stride (S = 4 * TW)
      constexpr bool kWide = 4 * TW > kNarrowPatV;
      using Meta = PatMetaT<kWide ? kMaxPatV : kNarrowPatV>;
      using VMask = typename std::conditional<kWide, uint64_t, uint32_t>::type;
      const Meta *metas = reinterpret_cast<const Meta *>(a.meta);
      int32_t p = 0, g = 0;
      uint4 hd = make_uint4(0, 0, 0, 0);
      Tuple<TW> t;
      uint32_t nrme = 0;  // the next source's rightmost-path edge
      // one-word tuples have the registers to hold the edge after it too, so
      // that a source switch need not wait on the pattern's metadata
      constexpr bool kDeep = TW <= 1;
      uint32_t nnrme = 0;
      VMask bwd_from = 0;
      if (live) {
        // the tuple (when its registers allow) and the pattern's first path
        // edge and backward sources load beside the metadata (none of them
        // depends on its flags)
        p = a.pid[r];
        g = a.rgraph[r];
        hd = *reinterpret_cast<const uint4 *>(metas + p);
        if constexpr (TW <= 4) {
          t.load(a.vt, r, a.S);
          nrme = metas[p].rm[0];
          if constexpr (kDeep) nnrme = metas[p].rm[1];
          bwd_from = metas[p].bwd_from;
        }
      }
      const bool work = live && !(hd.w & kPatDead);  // a dead pattern's rows list nothing
      const int nv = (int)(hd.y & 0xFF), nrm = (int)((hd.y >> 8) & 0xFF);
      const int minlabel = (int)((hd.y >> 16) & 0xFF), maxlabel = (int)(hd.y >> 24);
      VMask bwd_mask = hd.z;
      int32_t vb = 0, ca = 0, ce = 0, na = 0, ne = 0;
      // the pattern's metadata and path masks, addressed from p where they
      // are read (two 64-bit pointers held through the item loop cost the
      // registers of a seventh wave per SIMD)
      auto pat = [&]() -> const Meta & { return metas[p]; };
      auto pmask = [&](int i) { return a.pmask[(int64_t)p * 2 * ms + i]; };
      unsigned long long mk = ~0ull;
      const int n = (int)rb;
      if (work) {
        t.nv = nv;
        if constexpr (TW > 4) {
          t.load(a.vt, r, a.S);
          bwd_from = pat().bwd_from;
        }
        vb = g;  // the row's graph is stored as its first vertex
        if constexpr (kWide) {
          bwd_mask |= (VMask)pat().bwd_mask_hi << 32;
          bwd_from |= (VMask)pat().bwd_from_hi << 32;
        }
        const int rv = t.get(nv - 1);
        ca = a.vadj[vb + rv];
        ce = a.vadj[vb + rv + 1];
        if (nrm > 0) {
          if constexpr (TW > 4) nrme = pat().rm[0];
          const int u = t.get(nrme & 0xFF);
          na = a.vadj[vb + u];
          ne = a.vadj[vb + u + 1];
        }
        if (ms) mk = pmask(0);
      }
      const uint32_t kb = (uint32_t)((int64_t)(p - a.p_lo) * a.K);
      const int maxtoc = nv - 1;
      int s = 0;      // current source: 0 the rightmost vertex, q >= 1 rm[q - 1].from
      int pre_s = 0;  // items before the current source
      // the current source's forward keys start at src_kbase (its vertex's
      // key block); from the second source on, the order rule against its
      // rightmost-path edge rme (label bits 8-15, to-label bits 16-23)
      int src_kbase = fwd_base + maxtoc * nel_nvl;
      uint32_t rme = 0;
      int32_t cb = ca;  // first item of the current source
      auto eval = [&](uint32_t e) {
        const int w = e & 0xFF, el = (e >> 8) & 0xFF, nl = (e >> 16) & 0xFF;
        const uint32_t dw = e >> 24;
        const int fw = t.find(w);  // pattern vertex already mapped to w, or -1
        int key = -1, wout = 0;
        uint32_t bound = 0;
        bool keep = true;
        const int lab = el * NVL + nl;  // the edge and new-vertex labels' place in a key and a mask
        if (fw < 0) {
          // forward to a new vertex w, from the current source (the source's
          // order rule; none from the rightmost vertex)
          const int e1 = (int)((rme >> 8) & 0xFF), l1 = (int)((rme >> 16) & 0xFF);
          if (nl >= minlabel && (s == 0 || e1 < el || (e1 == el && l1 <= nl))) {
            key = src_kbase + lab;
            wout = w;
            bound = (uint32_t)(n - pre_s) + dw;  // sources s - 1 .. root, plus w
            keep = (mk >> lab) & 1ull;
          }
        } else if (s == 0 && ((bwd_from >> fw) & 1)) {
          // backward to a rightmost-path vertex f = rm[q].from, q >= 1
          const int q = pat().rmq[fw];
          const uint32_t rq = pat().rm[q];
          const int e1 = (rq >> 8) & 0xFF, l1 = (rq >> 16) & 0xFF;
          if (!((bwd_mask >> fw) & 1) && (e1 < el || (e1 == el && l1 <= maxlabel))) {
            key = fw * NEL + el;
            wout = -1;
            bound = (uint32_t)n;
            if (ms) keep = (pmask(ms + 1 + q) >> el) & 1ull;
          }
        }
        if (key >= 0) {
          if (!keep) {
            ++n_pruned;  // counted as a candidate, never listed
          } else if (cnt < n) {
            a.cand[lo + cnt] = make_uint2(cand_code(false, key, wout), bound);
            ++cnt;
            count_insert(tab, kb + (uint32_t)key, head, bound);
          } else {
            err = 1;  // more candidates than items: the row's item count is wrong
          }
        }
      };
      bool more = work && ca < ce;
      while (__ballot(more)) {
        if (more) {
          const int nstep = min(kStepItems, ce - ca);
          uint32_t e[kStepItems];
#pragma unroll
          for (int u = 0; u < kStepItems; ++u) e[u] = u < nstep ? a.adj[ca + u] : 0u;
#pragma unroll
          for (int u = 0; u < kStepItems; ++u)
            if (u < nstep) eval(e[u]);
          ca += nstep;
          if (ca >= ce) {  // next source (each source has at least one item)
            pre_s += ce - cb;
            ++s;
            if (s > nrm) {
              more = false;
              if (pre_s != n) err = 1;  // the row's item count disagrees with its sources
            } else {
              ca = cb = na;
              ce = ne;
              rme = nrme;
              src_kbase = fwd_base + (int)(rme & 0xFF) * nel_nvl;
              if (ms) mk = pmask(s);
              if (s < nrm) {
                if constexpr (kDeep) nrme = nnrme;
                else nrme = pat().rm[s];
                const int u = t.get(nrme & 0xFF);
                na = a.vadj[vb + u];
                ne = a.vadj[vb + u + 1];
                if constexpr (kDeep) nnrme = pat().rm[s + 1];  // s + 1 <= nrm < nv: in the array
              }
              more = ca < ce;
            }
          }
        }
      }
    }
    if (live) a.rccnt[r] = (uint16_t)cnt;
    listed += (unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane(wave_sum(cnt));
    wave_sync();
    // the block's segments are done (but a carried one, at head lane 0)
    // support: one per segment of the block that holds the key, less a
    // segment carried in from the previous block; a carried-out segment
    // keeps its bit (at head lane 0 of the next block)
    for (uint32_t i = lane; i < tab.cap; i += 64) {
      const unsigned long long m = tab.mask[i];
      if (m) {
        const uint32_t sv = tab.sup[i];
        const uint32_t keep = carry ? (uint32_t)(m & 1ull) : 0u;
        tab.sup[i] = ((sv & 0x7FFFFFFFu) + (uint32_t)__popcll(m) - (sv >> 31)) | (keep << 31);
        tab.mask[i] = keep;
      }
    }
    wave_sync();
    b += nb;
  }
  const bool failed = s_scal[1] != 0;
  if (!failed) {
    // commit the counts to the chunk's replica; keep the chunk's
    // (key, embeddings, item sum) records
    const int64_t gci = a.chunk0 + ci;
    const int64_t roff = (gci & (a.rep - 1)) * a.nkeys + (int64_t)(a.p_lo - a.p_base) * a.K;
    uint4 *rec = a.chunk_keys + gci * a.chunk_key_cap;
    uint32_t nrec = 0;
    for (uint32_t i0 = 0; i0 < tab.cap; i0 += 64) {
      const uint32_t i = i0 + lane;
      const uint32_t k = tab.key[i];
      const bool used = k != kEmpty32;
      const uint64_t m = __ballot(used);
      if (used) {
        const uint32_t emb = tab.emb[i], bs = tab.bsum[i];
        atomicAdd(&a.acc[roff + (int64_t)k], ((unsigned long long)tab.sup[i] << kSupShift) | (unsigned long long)emb);
        atomicAdd(&a.accb[roff + (int64_t)k], (unsigned long long)bs);
        rec[nrec + (uint32_t)__popcll(m & lanes_below(lane))] = make_uint4(k, emb, bs, 0u);
      }
      nrec += (uint32_t)__popcll(m);
    }
    if (lane == 0) info->nkeys = nrec;
  } else if (lane == 0) {
    const uint32_t at = atomicAdd(a.fail_count, 1u);
    a.fail_list[at] = (uint32_t)(a.chunk0 + ci);
    info->flags = kChunkFailed;
  }
  if (__ballot(err != 0) && lane == 0) atomicOr(a.error_flag, 1);
  const unsigned long long w_cand = listed + wave_sum((unsigned long long)n_pruned);
  const unsigned long long w_pruned = wave_sum((unsigned long long)n_pruned);
  if (lane == 0) {
    unsigned long long *line = a.spread + (ci & (kSpread - 1)) * kSpreadStride;
    if (w_cand) atomicAdd(&line[kSpreadCand], w_cand);
    if (w_pruned) atomicAdd(&line[kSpreadPruned], w_pruned);
    atomicMax(&line[kSpreadFill], (unsigned long long)s_scal[0]);
    if (failed) atomicAdd(&line[kSpreadFailed], 1ull);
  }
  stamp_end(a, ci);
}

// ------------------------------------------------------------- LEVEL FIX
// Chunks whose count table filled up, recounted from their lists (persistent
// grid over the failure list): each chunk in parts of about kFixPartRows
// rows cut at segment starts, one wave per part (a segment's pairs stay in
// one part); (segment, key) pairs are deduplicated in a 2^kFixTableBits-
// entry LDS table, one device atomic per distinct pair.  A pass that fills
// the table is split in two by the residue of the key (every key lands in
// exactly one committed pass).
constexpr int kFixPartRows = 64;
template <bool ROOT>
__global__ void __launch_bounds__(64) k_level_fix(StepArgs a) {
  constexpr uint32_t cap = 1u << kFixTableBits, limit = cap - cap / 4;
  __shared__ unsigned long long tkey[cap];
  __shared__ uint32_t tcnt[cap], tbs[cap];
  __shared__ uint32_t stack_mod[32], stack_res[32];
  __shared__ int s_fill, s_over;
  const int lane = threadIdx.x;
  const uint32_t nfail = *a.fail_count;
  const int64_t parts = (a.chunk_rows + kFixPartRows - 1) / kFixPartRows;
  const int64_t units = (int64_t)nfail * parts;
  for (int64_t unit = blockIdx.x; unit < units; unit += gridDim.x) {
    const uint32_t fi = (uint32_t)(unit / parts);
    const int64_t part = unit % parts;
    const uint32_t gci = a.fail_list[fi];
    ChunkInfo c = a.chunks[gci];
    // the part's rows: nominal cuts moved to segment starts (the last part
    // takes the rest of the chunk, which may run past its nominal rows)
    {
      const int64_t lo = part == 0 ? c.start : min(c.end, seg_at_or_after<ROOT>(a, c.start + part * kFixPartRows));
      const int64_t hi =
          part == parts - 1 ? c.end : min(c.end, seg_at_or_after<ROOT>(a, c.start + (part + 1) * kFixPartRows));
      c.start = lo;
      c.end = hi;
    }
    if (c.start >= c.end) continue;
    const int64_t roff = (int64_t)(gci & (uint32_t)(a.rep - 1)) * a.nkeys;
    int top = 1;
    if (lane == 0) {
      stack_mod[0] = 1;
      stack_res[0] = 0;
    }
    wave_sync();
    while (top > 0) {
      const uint32_t mod = stack_mod[top - 1], res = stack_res[top - 1];
      for (uint32_t i = lane; i < cap; i += 64) {
        tkey[i] = ~0ull;
        tcnt[i] = 0;
        tbs[i] = 0;
      }
      if (lane == 0) {
        s_fill = 0;
        s_over = 0;
      }
      wave_sync();
      int64_t seg_head = c.start;  // first row of the segment the block starts in
      for (int64_t b = c.start; b < c.end;) {
        const int nb = (int)min<int64_t>(64, c.end - b);
        const int64_t r = b + lane;
        const bool live = lane < nb;
        const uint64_t starts = __ballot(live && seg_start<ROOT>(a, r));
        const uint64_t upto = starts & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
        const int64_t hrow = upto ? b + (63 - __clzll((long long)upto)) : seg_head;
        const unsigned long long lo = live ? row_list<ROOT>(a, r) : 0ull;
        const int n = live ? (int)a.rccnt[r] : 0;
        const int maxn = wave_max(n);
        for (int j = 0; j < maxn; ++j) {
          if (j >= n) continue;
          int key, w;
          const uint2 ce = a.cand[lo + j];
          cand_decode(ROOT, ce.x, key, w);
          if (mod > 1 && ((uint32_t)key & (mod - 1)) != res) continue;
          if (__hip_atomic_load(&s_over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) continue;
          // the segment (head row) fixes the pattern
          const unsigned long long pk = ((unsigned long long)(hrow - c.start) << 32) | (uint32_t)key;
          uint32_t h = (uint32_t)((pk * 0x9E3779B97F4A7C15ull) >> (64 - kFixTableBits));
          bool done = false;
          for (uint32_t probe = 0; probe < cap && !done; ++probe) {
            const unsigned long long old = atomicCAS(&tkey[h], ~0ull, pk);
            if (old == ~0ull || old == pk) {
              atomicAdd(&tcnt[h], 1u);
              atomicAdd(&tbs[h], ce.y);
              if (old == ~0ull && (uint32_t)atomicAdd(&s_fill, 1) >= limit) atomicOr(&s_over, 1);
              done = true;
            }
            h = (h + 1) & (cap - 1);
          }
          if (!done) atomicOr(&s_over, 1);
        }
        const uint64_t last = starts ? 63 - __clzll((long long)starts) : 64;
        if (last < 64 && (int)last < nb) seg_head = b + (int64_t)last;
        b += nb;
        wave_sync();
      }
      wave_sync();
      const bool over = s_over != 0;
      if (over) {
        if (mod >= (1u << 24) || top + 1 > 32) {  // cannot happen: keys are below 2^24
          if (lane == 0) atomicOr(a.error_flag, 4);
          top = 0;
        } else {
          if (lane == 0) {
            stack_mod[top - 1] = mod * 2;
            stack_res[top - 1] = res + mod;
            stack_mod[top] = mod * 2;
            stack_res[top] = res;
          }
          top += 1;
        }
        wave_sync();
        continue;
      }
      for (uint32_t i = lane; i < cap; i += 64) {
        const unsigned long long k = tkey[i];
        if (k == ~0ull) continue;
        const int64_t hr = c.start + (int64_t)(k >> 32);
        const int32_t p = ROOT ? a.gjob[a.vgraph[hr]] : a.pid[hr];
        const int64_t at = roff + (int64_t)(p - a.p_base) * a.K + (int64_t)(uint32_t)k;
        atomicAdd(&a.acc[at], (1ull << kSupShift) | (unsigned long long)tcnt[i]);
        atomicAdd(&a.accb[at], (unsigned long long)tbs[i]);
      }
      top -= 1;
      wave_sync();
    }
  }
  if (blockIdx.x < (unsigned)kSpread) stamp_end(a, blockIdx.x);
}

// ----------------------------------------------------------------- REDUCE
// Sum every key's replicas; keys reaching their job's min support go to the
// frequent list as (key, support, embeddings, item sum), in key order, and
// keep their replicas (k_cursors splits the children's rows by them, then
// clears them); the other keys' replicas are cleared here.  Three kernels:
// k_reduce counts the frequent keys of each tile of kReduceTile keys,
// k_reduce_scan turns the counts into the tiles' first slots (and the
// list's length), k_reduce_list writes each tile's records in key order.
constexpr int kReduceThreads = 256;
constexpr int kReduceSteps = kReduceTile / kReduceThreads;
static_assert(kReduceSteps * kReduceThreads == kReduceTile, "reduce tile");

template <bool ROOT>
__device__ __forceinline__ bool reduce_frequent(const StepArgs &a, int64_t i, int32_t s) {
  if (s <= 0) return false;
  const int64_t p = i / a.K + a.p_base;
  const int32_t owner = ROOT ? (int32_t)p : meta_at(a.meta, a.S, p)->job;
  return s >= a.minsup_job[owner];
}

// sum of x over the workgroup (kReduceThreads lanes), read by every lane
__device__ __forceinline__ uint32_t block_sum(uint32_t x, uint32_t *s_part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  x = wave_sum(x);
  if (lane == 0) s_part[wave] = x;
  __syncthreads();
  uint32_t t = 0;
#pragma unroll
  for (int w = 0; w < kReduceThreads / 64; ++w) t += s_part[w];
  __syncthreads();
  return t;
}

template <bool ROOT>
__global__ void __launch_bounds__(kReduceThreads) k_reduce(StepArgs a) {
  __shared__ uint32_t s_part[kReduceThreads / 64];
  const int64_t n_tiles = reduce_tiles(a.nkeys);
  for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    uint32_t hits = 0;
    for (int j = 0; j < kReduceSteps; ++j) {
      const int64_t i = t * kReduceTile + (int64_t)j * kReduceThreads + threadIdx.x;
      if (i >= a.nkeys) continue;
      unsigned long long v = 0;
      for (int r = 0; r < a.rep; ++r) v += a.acc[(int64_t)r * a.nkeys + i];
      const int32_t s = (int32_t)(v >> kSupShift);
      a.sup[i] = s;
      const bool hit = reduce_frequent<ROOT>(a, i, s);
      hits += hit ? 1u : 0u;
      if (v && !hit)
        for (int r = 0; r < a.rep; ++r) {
          a.acc[(int64_t)r * a.nkeys + i] = 0;
          a.accb[(int64_t)r * a.nkeys + i] = 0;
        }
    }
    hits = block_sum(hits, s_part);
    if (threadIdx.x == 0) a.tile_count[t] = hits;
  }
}

// One workgroup: exclusive scan of the tile counts in place; the total is
// the list's length.
__global__ void __launch_bounds__(1024) k_reduce_scan(StepArgs a) {
  __shared__ uint32_t s_wave[16];
  const int64_t n = reduce_tiles(a.nkeys);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t carry = 0;
  for (int64_t b = 0; b < n; b += 1024) {
    const int64_t i = b + threadIdx.x;
    const uint32_t x = i < n ? a.tile_count[i] : 0u;
    uint32_t wt;
    const uint32_t ex = wave_exclusive_scan<uint32_t>(x, wt);
    if (lane == 0) s_wave[wave] = wt;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
      before += w < wave ? s_wave[w] : 0u;
      total += s_wave[w];
    }
    if (i < n) a.tile_count[i] = carry + before + ex;
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) *a.freq_count = carry;
}

template <bool ROOT>
__global__ void __launch_bounds__(kReduceThreads) k_reduce_list(StepArgs a) {
  __shared__ uint32_t s_part[kReduceThreads / 64];
  const int64_t n_tiles = reduce_tiles(a.nkeys);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    uint32_t next = a.tile_count[t];  // the tile's first slot
    for (int j = 0; j < kReduceSteps; ++j) {
      const int64_t i = t * kReduceTile + (int64_t)j * kReduceThreads + threadIdx.x;
      const int32_t s = i < a.nkeys ? a.sup[i] : 0;
      const bool hit = i < a.nkeys && reduce_frequent<ROOT>(a, i, s);
      // rank among the tile's keys of this step, in key order
      const uint64_t m = __ballot(hit);
      if (lane == 0) s_part[wave] = (uint32_t)__popcll(m);
      __syncthreads();
      uint32_t before = 0, step = 0;
#pragma unroll
      for (int w = 0; w < kReduceThreads / 64; ++w) {
        before += w < wave ? s_part[w] : 0u;
        step += s_part[w];
      }
      __syncthreads();
      if (hit) {
        const uint32_t slot = next + before + (uint32_t)__popcll(m & lanes_below(lane));
        if (slot < a.freq_cap) {
          unsigned long long v = 0, bs = 0;
          for (int r = 0; r < a.rep; ++r) {
            v += a.acc[(int64_t)r * a.nkeys + i];
            bs += a.accb[(int64_t)r * a.nkeys + i];
          }
          unsigned long long *f = a.freq + 4 * (int64_t)slot;
          f[0] = (unsigned long long)i;
          f[1] = (unsigned long long)s;
          f[2] = v & ((1ull << kSupShift) - 1);
          f[3] = bs;
        }
      }
      next += step;
    }
  }
}

// After the host chose the children: frequent key i (the key of record i of
// the frequent list) becomes child fchild[i] (>= 0; -1: dropped), its sup
// entry points at the child, and
// every (child, replica) gets its first row and list entry -- the child's
// bases plus the earlier replicas' embeddings and item sums -- and a zero
// reservation cursor; the frequent keys' replicas are cleared.  Also copies
// the children's upload.
__global__ void k_cursors(StepArgs a, unsigned long long *dst, const unsigned long long *src, int64_t n,
                          const int32_t *fchild, const unsigned long long *row_base, const unsigned long long *list_base,
                          int64_t F) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t i = t; i < n; i += stride) dst[i] = src[i];
  for (int64_t i = t; i < F; i += stride) {
    const int64_t key = (int64_t)a.freq[4 * i];
    const int32_t ch = fchild[i];
    unsigned long long re = ch >= 0 ? row_base[ch] : 0ull, rl = ch >= 0 ? list_base[ch] : 0ull;
    if (ch >= 0) a.sup[key] = -(ch + 1);
    for (int r = 0; r < a.rep; ++r) {
      const int64_t at = (int64_t)r * a.nkeys + key;
      const unsigned long long v = a.acc[at], bs = a.accb[at];
      if (ch >= 0) {
        a.cur_rows[(int64_t)ch * a.rep + r] = re;
        a.cur_list[(int64_t)ch * a.rep + r] = rl;
        a.cursor[(int64_t)ch * a.rep + r] = 0;
        re += v & ((1ull << kSupShift) - 1);
        rl += bs;
      }
      a.acc[at] = 0;
      a.accb[at] = 0;
    }
  }
}

// ------------------------------------------------------------------ WRITE
// One wave per chunk (the chunks k_level recorded).  Each child's rows of
// the chunk, and their list regions, are reserved with one atomic each on
// the child's cursors of the chunk's replica -- the counts come from the
// chunk's (key, embeddings, item sum) records -- then written with the
// candidates taken 64 at a time in list order (rows in order, each row's
// list in order): a lane's position is its child's reserved base plus the
// child's running count plus its rank among the window's lanes of the same
// child, its list offset likewise with the item counts.  The child's rows
// of one chunk are thus consecutive and in the parent's row order, so every
// child's rows stay graph-contiguous.
// A chunk that k_level_fix recounted has no records, and a chunk whose
// records name more children than the table holds cannot use them: either
// is counted here, a pass over the children c with c mod `mod` == res at a
// time (lane per row), and a pass whose table fills up is split in two by
// child residue before anything is reserved.
constexpr int kWriteUnroll = 2;  // candidate windows whose loads are issued together

template <bool ROOT, int TW>
__global__ void __launch_bounds__(64, (ROOT ? 5 : TW <= 8 ? 8 : 6)) k_write(StepArgs a) {
  constexpr int kBits = ROOT ? kWriteTableBitsRoot : kWriteTableBits;
  constexpr uint32_t cap = 1u << kBits;
  // keyed by the candidate's key relative to the chunk's first pattern
  // (which names one child), with that child's id in cch
  __shared__ uint32_t ckey[cap];
  __shared__ int32_t cch[cap];
  // per child: its rows and list entries in the chunk while counting, then
  // (cleared at the reservation) the rows and entries written so far
  __shared__ uint32_t ccnt[cap], cbs[cap];
  uint32_t *const crun = ccnt, *const crunl = cbs;
  __shared__ unsigned long long cbase[cap], clbase[cap];
  __shared__ int s_pre[64];
  __shared__ unsigned long long s_lo[64];
  __shared__ int32_t s_p[64], s_g[64], s_nvp[64];
  __shared__ uint32_t s_tup[64 * (TW > 0 ? TW : 1)];
  __shared__ uint32_t stack_mod[32], stack_res[32];
  __shared__ int s_fill, s_over;
  const int lane = threadIdx.x;
  const ChunkInfo c = a.chunks[blockIdx.x];
  if (c.start >= c.end) return;
  bool counted = !(c.flags & kChunkFailed);
  const int CW = a.CS >> 2;
  const int64_t rsel = (int64_t)(blockIdx.x & (unsigned)(a.rep - 1));  // the chunk's replica
  unsigned long long written = 0;
  int top = 1;
  if (lane == 0) {
    stack_mod[0] = 1;
    stack_res[0] = 0;
  }
  wave_sync();
  // child selected by the candidate (key of pattern / job p), or -1
  auto child_of = [&](int64_t pk) -> int32_t {
    const int32_t v = a.sup[pk];
    return v < 0 ? -v - 1 : -1;
  };
  // the table slot of key gk, or -1 (not a child of this pass)
  auto lookup = [&](uint32_t gk) -> int {
    uint32_t h = hash_slot32(gk, kBits);
    for (uint32_t probe = 0; probe < cap; ++probe) {
      const uint32_t cur = ckey[h];
      if (cur == gk) return (int)h;
      if (cur == kEmpty32) return -1;
      h = (h + 1) & (cap - 1);
    }
    return -1;
  };
  // add n rows of child ch, with items bs, to the table (false: past the
  // fill limit; a chunk's records, at most one per table entry of k_level,
  // may fill it completely)
  auto table_add = [&](uint32_t gk, int32_t ch, uint32_t n, uint32_t bs, uint32_t limit) -> bool {
    uint32_t h = hash_slot32(gk, kBits);
    for (uint32_t probe = 0; probe < cap; ++probe) {
      const uint32_t old = atomicCAS(&ckey[h], kEmpty32, gk);
      if (old == kEmpty32 || old == gk) {
        if (old == kEmpty32) cch[h] = ch;
        atomicAdd(&ccnt[h], n);
        atomicAdd(&cbs[h], bs);
        if (old == kEmpty32 && (uint32_t)atomicAdd(&s_fill, 1) >= limit) return false;
        return true;
      }
      h = (h + 1) & (cap - 1);
    }
    return false;
  };
  while (top > 0) {
    const uint32_t mod = stack_mod[top - 1], res = stack_res[top - 1];
    for (uint32_t i = lane; i < cap; i += 64) {
      ckey[i] = kEmpty32;
      ccnt[i] = 0;
      cbs[i] = 0;
    }
    if (lane == 0) {
      s_fill = 0;
      s_over = 0;
    }
    wave_sync();
    if (counted) {
      // the chunk's records: one per key, so one per child
      const uint4 *rec = a.chunk_keys + (int64_t)blockIdx.x * a.chunk_key_cap;
      for (uint32_t i = lane; i < c.nkeys; i += 64) {
        const uint4 kr = rec[i];
        const int32_t ch = child_of((int64_t)(c.p_lo - a.p_base) * a.K + (int64_t)kr.x);
        if (ch >= 0 && !table_add(kr.x, ch, kr.y, kr.z, cap)) atomicOr(&s_over, 1);
      }
    } else {
      // count: lane per row
      for (int64_t b = c.start; b < c.end;) {
        const int nb = (int)min<int64_t>(64, c.end - b);
        const int64_t r = b + lane;
        const bool live = lane < nb;
        const unsigned long long lo = live ? row_list<ROOT>(a, r) : 0ull;
        int n = 0, p = 0;
        if (live) {
          n = a.rccnt[r];
          p = ROOT ? a.gjob[a.vgraph[r]] : a.pid[r];
        }
        const int maxn = wave_max(n);
        for (int j = 0; j < maxn; ++j) {
          if (j >= n) continue;
          int key, w;
          const uint2 ce = a.cand[lo + j];
          cand_decode(ROOT, ce.x, key, w);
          const int32_t ch = child_of((int64_t)(p - a.p_base) * a.K + key);
          if (ch < 0 || (mod > 1 && ((uint32_t)ch & (mod - 1)) != res)) continue;
          if (__hip_atomic_load(&s_over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) continue;
          const uint32_t gk = (uint32_t)((int64_t)(p - c.p_lo) * a.K + key);
          if (!table_add(gk, ch, 1u, ce.y, cap - cap / 8)) atomicOr(&s_over, 1);
        }
        b += nb;
      }
    }
    wave_sync();
    if (s_over) {
      if (counted) {  // more children than the table holds: count them in split passes
        counted = false;
        wave_sync();
        continue;
      }
      if (mod >= (1u << 30) || top + 1 > 32) {  // cannot happen: a pass of one child fits
        if (lane == 0) atomicOr(a.error_flag, 8);
        break;
      }
      if (lane == 0) {
        stack_mod[top - 1] = mod * 2;
        stack_res[top - 1] = res + mod;
        stack_mod[top] = mod * 2;
        stack_res[top] = res;
      }
      top += 1;
      wave_sync();
      continue;
    }
    // reserve
    for (uint32_t i = lane; i < cap; i += 64)
      if (ckey[i] != kEmpty32) {
        // one atomic reserves both: rows << 32 | list entries
        const int64_t at = (int64_t)cch[i] * a.rep + rsel;
        const unsigned long long v =
            atomicAdd(&a.cursor[at], ((unsigned long long)ccnt[i] << 32) | (unsigned long long)cbs[i]);
        cbase[i] = a.cur_rows[at] + (v >> 32);
        clbase[i] = a.cur_list[at] + (v & 0xFFFFFFFFull);
        crun[i] = 0;  // the counts' storage becomes the write runs
        crunl[i] = 0;
      }
    wave_sync();
    // write, candidates 64 at a time in list order
    for (int64_t b = c.start; b < c.end;) {
      const int nb = (int)min<int64_t>(64, c.end - b);
      const int64_t r = b + lane;
      const bool live = lane < nb;
      const unsigned long long lo = live ? row_list<ROOT>(a, r) : 0ull;
      int n = 0;
      int32_t p = 0, g = 0, nvp = 0;
      if (live) {
        n = a.rccnt[r];
        if (ROOT) {
          const int32_t gi = a.vgraph[r];
          p = a.gjob[gi];
          g = a.gvbase[gi];  // the children's graph, as its first vertex
          nvp = (int)(r - g);  // the child tuple starts with the vertex
        } else {
          p = a.pid[r];
          g = a.rgraph[r];
          if (n) {
            nvp = meta_at(a.meta, a.S, p)->nv;
            const uint32_t *src = reinterpret_cast<const uint32_t *>(a.vt + r * a.S);
#pragma unroll
            for (int q = 0; q < TW; ++q) s_tup[lane * TW + q] = src[q];
          }
        }
      }
      int T;
      const int pre = wave_exclusive_scan<int>(n, T);
      s_pre[lane] = pre;
      s_lo[lane] = lo;
      s_p[lane] = p;
      s_g[lane] = g;
      s_nvp[lane] = nvp;
      wave_sync();
      for (int jb0 = 0; jb0 < T; jb0 += 64 * kWriteUnroll) {
        int kk[kWriteUnroll];
        uint2 ent[kWriteUnroll];
        int hv[kWriteUnroll];
#pragma unroll
        for (int u = 0; u < kWriteUnroll; ++u) {
          const int j = jb0 + u * 64 + lane;
          kk[u] = 0;
          ent[u] = make_uint2(0, 0);
          if (j < T) {
            int lo_k = 0, hi_k = nb;  // last row k with s_pre[k] <= j
            while (hi_k - lo_k > 1) {
              const int mid = (lo_k + hi_k) >> 1;
              if (s_pre[mid] <= j) lo_k = mid;
              else hi_k = mid;
            }
            kk[u] = lo_k;
            ent[u] = a.cand[s_lo[lo_k] + (unsigned long long)(j - s_pre[lo_k])];
          }
        }
#pragma unroll
        for (int u = 0; u < kWriteUnroll; ++u) {
          hv[u] = -1;
          if (jb0 + u * 64 + lane < T) {
            int key, w;
            cand_decode(ROOT, ent[u].x, key, w);
            hv[u] = lookup((uint32_t)((int64_t)(s_p[kk[u]] - c.p_lo) * a.K + key));
          }
        }
#pragma unroll
        for (int u = 0; u < kWriteUnroll; ++u) {
          const int k = kk[u];
          int key, w;
          cand_decode(ROOT, ent[u].x, key, w);
          const int32_t ch = hv[u] >= 0 ? cch[hv[u]] : -1;  // the pass's children only
          const uint32_t h = hv[u] >= 0 ? (uint32_t)hv[u] : 0u;
          uint64_t active = __ballot(ch >= 0);
          written += (unsigned long long)__popcll(active);
          // list regions: any disjoint order within the child's reserved range
          unsigned long long loff = 0;
          if (ch >= 0) loff = clbase[h] + (unsigned long long)atomicAdd(&crunl[h], ent[u].y);
          // row positions in list order: the lanes of the window with the
          // same child (the same table slot h) found with one ballot per
          // slot bit; each takes its rank among them, and the lowest
          // advances the child's run
          const bool sel = ch >= 0;
          uint64_t peers = active;
#pragma unroll
          for (int bit = 0; bit < kBits; ++bit) {
            const bool hb = (h >> bit) & 1u;
            const uint64_t m = __ballot(sel && hb);
            peers &= hb ? m : ~m;
          }
          const uint32_t rank = (uint32_t)__popcll(peers & lanes_below(lane));
          const uint32_t run = sel ? crun[h] : 0u;
          wave_sync();
          if (sel && rank == 0) crun[h] = run + (uint32_t)__popcll(peers);
          wave_sync();
          const unsigned long long pos = sel ? cbase[h] + (unsigned long long)run + rank : 0ull;
          // then the rows, every selected lane at once
          if (ch >= 0 && (int64_t)pos < a.ccap) {
            a.cpid[pos] = ch;
            a.crgraph[pos] = s_g[k];
            a.cbound[pos] = ent[u].y;
            a.crlist[pos] = loff;
            uint32_t *dst = reinterpret_cast<uint32_t *>(a.cvt + pos * a.CS);
            if (ROOT) {
              dst[0] = (uint32_t)s_nvp[k] | ((uint32_t)w << 8);
              for (int q = 1; q < CW; ++q) dst[q] = 0;
            } else {
              const int nvp = s_nvp[k];
              for (int q = 0; q < CW; ++q) {
                uint32_t word = q < TW ? s_tup[k * TW + q] : 0u;
                if (w >= 0 && (nvp >> 2) == q) {
                  const int sh = (nvp & 3) * 8;
                  word = (word & ~(0xFFu << sh)) | ((uint32_t)w << sh);
                }
                dst[q] = word;
              }
            }
          }
        }
      }
      wave_sync();  // the next block reuses the staged rows
      b += nb;
    }
    top -= 1;
    wave_sync();
  }
  if (lane == 0 && written)
    atomicAdd(&a.spread[(blockIdx.x & (kSpread - 1)) * kSpreadStride + kSpreadWritten], written);
}

// ----------------------------------------------------------- SUBSETS
// A packed database of selected graphs gathered from a resident one: one
// workgroup per selected graph t (source graph sel[t]) copies its vertex
// labels, its adjacency entries (local ids, labels and degrees unchanged)
// and its vertices' adjacency offsets rebased to dab[t]; graph t starts at
// vertex dvb[t] and belongs to job djob[t].  The last workgroup writes the
// closing offsets.
__global__ void __launch_bounds__(256) k_gather_graphs(GatherArgs g) {
  const int64_t t = blockIdx.x;
  const int32_t src = g.sel[t];
  const int32_t v0 = g.gvbase[src], n = g.gvbase[src + 1] - v0;
  const int32_t a0 = g.vadj[v0], na = g.vadj[v0 + n] - a0;
  const int32_t dv = g.dvb[t], da = g.dab[t];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    g.out_vlab[dv + i] = g.vlab[v0 + i];
    g.out_vadj[dv + i] = da + (g.vadj[v0 + i] - a0);
    g.out_vgraph[dv + i] = (int32_t)t;
  }
  for (int i = threadIdx.x; i < na; i += blockDim.x) g.out_adj[da + i] = g.adj[a0 + i];
  if (threadIdx.x == 0) {
    g.out_gvbase[t] = dv;
    g.out_gjob[t] = g.djob[t];
    if (t == g.n_sel - 1) {
      g.out_gvbase[t + 1] = dv + n;
      g.out_vadj[dv + n] = da + na;
    }
  }
}

// ------------------------------------------------------------ primitives
// Copies between device memory and mapped pinned host memory on the compute
// queue (a copy engine transfer between two kernels costs a cross-engine
// wait; the level step's transfers are small).
__global__ void k_copy8(unsigned long long *dst, const unsigned long long *src, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ void k_fill8(unsigned long long *dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = 0ull;
}

__global__ void __launch_bounds__(256) k_clear_counters(StepArgs a) {
  if (threadIdx.x == 0) {
    *a.freq_count = 0;
    *a.fail_count = 0;
    *a.error_flag = 0;
  }
  // the written-row slots keep counting across levels (the host compares
  // their total with the rows it asked for)
  for (int i = threadIdx.x; i < kSpread * kSpreadStride; i += blockDim.x)
    if (i % kSpreadStride != kSpreadWritten) a.spread[i] = 0;
}

template <bool ROOT>
hipError_t level_tw(const StepArgs &a, hipStream_t s, size_t lds) {
  const unsigned grid = (unsigned)a.n_chunks;
  if (ROOT) {
    hipLaunchKernelGGL((k_level<true, 0>), dim3(grid), dim3(64), lds, s, a);
    return hipGetLastError();
  }
  switch (a.S >> 2) {
    case 1: hipLaunchKernelGGL((k_level<false, 1>), dim3(grid), dim3(64), lds, s, a); break;
    case 2: hipLaunchKernelGGL((k_level<false, 2>), dim3(grid), dim3(64), lds, s, a); break;
    case 3: hipLaunchKernelGGL((k_level<false, 3>), dim3(grid), dim3(64), lds, s, a); break;
    case 4: hipLaunchKernelGGL((k_level<false, 4>), dim3(grid), dim3(64), lds, s, a); break;
    case 5: hipLaunchKernelGGL((k_level<false, 5>), dim3(grid), dim3(64), lds, s, a); break;
    case 6: hipLaunchKernelGGL((k_level<false, 6>), dim3(grid), dim3(64), lds, s, a); break;
    case 7: hipLaunchKernelGGL((k_level<false, 7>), dim3(grid), dim3(64), lds, s, a); break;
    case 8: hipLaunchKernelGGL((k_level<false, 8>), dim3(grid), dim3(64), lds, s, a); break;
    case 16: hipLaunchKernelGGL((k_level<false, 16>), dim3(grid), dim3(64), lds, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

// ------------------------------------------------------------ launchers
size_t level_lds_bytes(int table_bits) { return (size_t)24 << table_bits; }
// at most one record per table entry
int level_chunk_keys(int table_bits) { return 1 << table_bits; }

hipError_t launch_level(const StepArgs &a, bool root, hipStream_t s) {
  if (a.n_chunks <= 0) return hipSuccess;
  if (a.table_bits < 6 || a.table_bits > 12) return hipErrorInvalidValue;
  const size_t lds = level_lds_bytes(a.table_bits);
  return root ? level_tw<true>(a, s, lds) : level_tw<false>(a, s, lds);
}

hipError_t launch_level_fix(const StepArgs &a, bool root, hipStream_t s) {
  if (a.n_chunks <= 0) return hipSuccess;
  const unsigned grid = (unsigned)std::min<int64_t>(a.n_chunks, 512);
  if (root)
    hipLaunchKernelGGL(k_level_fix<true>, dim3(grid), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL(k_level_fix<false>, dim3(grid), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_reduce(const StepArgs &a, bool root, hipStream_t s) {
  if (a.nkeys <= 0) return hipSuccess;
  const unsigned g = (unsigned)std::min<int64_t>(reduce_tiles(a.nkeys), 8192);
  if (root)
    hipLaunchKernelGGL(k_reduce<true>, dim3(g), dim3(kReduceThreads), 0, s, a);
  else
    hipLaunchKernelGGL(k_reduce<false>, dim3(g), dim3(kReduceThreads), 0, s, a);
  hipLaunchKernelGGL(k_reduce_scan, dim3(1), dim3(1024), 0, s, a);
  if (root)
    hipLaunchKernelGGL(k_reduce_list<true>, dim3(g), dim3(kReduceThreads), 0, s, a);
  else
    hipLaunchKernelGGL(k_reduce_list<false>, dim3(g), dim3(kReduceThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_write(const StepArgs &a, bool root, hipStream_t s) {
  if (a.n_chunks <= 0) return hipSuccess;
  const unsigned grid = (unsigned)a.n_chunks;
  if (root) {
    hipLaunchKernelGGL((k_write<true, 0>), dim3(grid), dim3(64), 0, s, a);
    return hipGetLastError();
  }
  switch (a.S >> 2) {
    case 1: hipLaunchKernelGGL((k_write<false, 1>), dim3(grid), dim3(64), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_write<false, 2>), dim3(grid), dim3(64), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_write<false, 3>), dim3(grid), dim3(64), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_write<false, 4>), dim3(grid), dim3(64), 0, s, a); break;
    case 5: hipLaunchKernelGGL((k_write<false, 5>), dim3(grid), dim3(64), 0, s, a); break;
    case 6: hipLaunchKernelGGL((k_write<false, 6>), dim3(grid), dim3(64), 0, s, a); break;
    case 7: hipLaunchKernelGGL((k_write<false, 7>), dim3(grid), dim3(64), 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_write<false, 8>), dim3(grid), dim3(64), 0, s, a); break;
    case 16: hipLaunchKernelGGL((k_write<false, 16>), dim3(grid), dim3(64), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

static unsigned copy_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 2048)); }

hipError_t launch_copy8(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes % 8) return hipErrorInvalidValue;
  const int64_t n = (int64_t)(bytes / 8);
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy8, dim3(copy_grid(n)), dim3(256), 0, s, static_cast<unsigned long long *>(dst),
                     static_cast<const unsigned long long *>(src), n);
  return hipGetLastError();
}

hipError_t launch_cursors(const StepArgs &a, void *dst, const void *src, size_t bytes, const int32_t *fchild,
                          const unsigned long long *row_base, const unsigned long long *list_base, int64_t F,
                          hipStream_t s) {
  if (bytes % 8) return hipErrorInvalidValue;
  const int64_t n = (int64_t)(bytes / 8);
  if (n == 0 && F == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cursors, dim3(copy_grid(std::max(n, F))), dim3(256), 0, s, a,
                     static_cast<unsigned long long *>(dst), static_cast<const unsigned long long *>(src), n, fchild,
                     row_base, list_base, F);
  return hipGetLastError();
}

hipError_t launch_gather_graphs(const GatherArgs &g, hipStream_t s) {
  if (g.n_sel <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_graphs, dim3((unsigned)g.n_sel), dim3(256), 0, s, g);
  return hipGetLastError();
}

hipError_t launch_fill8(void *dst, size_t bytes, hipStream_t s) {
  if (bytes % 8) return hipErrorInvalidValue;
  const int64_t n = (int64_t)(bytes / 8);
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill8, dim3(copy_grid(n)), dim3(256), 0, s, static_cast<unsigned long long *>(dst), n);
  return hipGetLastError();
}

hipError_t launch_clear_counters(const StepArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(256), 0, s, a);
  return hipGetLastError();
}

// ------------------------------------------------------ static workspace
// The entry point that allocates nothing (gm_mine_noalloc) takes every
// buffer of its level steps from this array.  Each GPU the library runs on
// holds its own copy until the process ends.  In host code `scratch` only
// names the symbol: its device address is looked up once per device with
// hipGetSymbolAddress, with that device current, and only that address is
// handed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

hipError_t static_scratch(void **ptr, size_t *bytes) {
  static std::mutex mu;
  static void *addr[64] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[dev] = p;
  }
  *ptr = addr[dev];
  *bytes = kStaticScratchBytes;
  return hipSuccess;
}

}  // namespace gm
