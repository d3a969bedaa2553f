// kernels.hip -- HIP kernels of the level step (gfx950, wave64).
//
// Support: one workgroup (at the root, one wave) owns one database graph,
// so every support question ("how many distinct graphs have an embedding of
// P with extension x") is answered by deduplicating (pattern, key) pairs
// inside it (an LDS hash table) and adding 1 per distinct pair to a global
// counter.  Per level:
//   k_expand   lists every row's extension candidates with their child
//              bounds, row-parallel (later levels);
//   k_support  deduplicates each graph's candidates (from the lists, or by
//              enumerating rows where no lists are kept): distinct-graph
//              support and embedding count per (pattern, key); the root
//              level runs k_root_wave over 1-edge keys instead;
//   k_reduce   sums the counter replicas and lists the frequent keys;
//   children   the child rows (pattern id + tuple) of the selected keys:
//              a count pass, a placement scan and a write pass that also
//              lays out the next level's candidate lists -- row-parallel
//              (k_rows, k_rtop, k_gfix) where every candidate is listed,
//              per graph (k_children, k_scan_*) otherwise.
//
// Extension rules (pattern side precomputed in PatMeta, graph side read
// from the packed adjacency):
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
// pattern-level fact (bwd_mask) and "vertex not in the embedding" is a
// 256-bit mask over the graph's local ids built from the tuple.
//
// Candidate bound of a row: every candidate of an embedding is one
// adjacency entry of its rightmost vertex or of a rightmost-path source, so
// bound = deg(rightmost) + sum of deg(rightmost-path sources).  The
// enumeration derives it for each child from the parent's path:
//   backward           the parent's bound (same path, same rightmost);
//   pure fwd to w      parent sources + parent rightmost, plus deg(w);
//   fwd from source i  sources i..root, plus deg(w);
//   root edge (u, w)   deg(u) + deg(w).
// The count pass sums the bounds of a graph's children; the scan turns the
// sums into the regions that the next k_support fills with their lists.
//
// Register discipline: a row's vertex tuple lives in a per-thread LDS slot
// (byte reads), the visited mask in four 64-bit values selected
// arithmetically, and the graph view is specialised for LDS or global
// memory, so the kernels use no per-lane scratch and no flat accesses.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.hpp"
#include "kernels.hpp"

namespace gm {

namespace {

constexpr uint32_t kEmpty32 = 0xFFFFFFFFu;
constexpr int kPackShift = 40;  // packed scans: count << 40 | bound sum
constexpr long long kPackMask = (1ll << kPackShift) - 1;

// Graph adjacency: entry offsets vadj[v] are absolute; entries A[a - aoff].
struct GView {
  const int32_t *vadj;
  const uint32_t *A;
  int32_t aoff;
  __device__ __forceinline__ int32_t begin(int v) const { return vadj[v]; }
  __device__ __forceinline__ int32_t end(int v) const { return vadj[v + 1]; }
  __device__ __forceinline__ int deg(int v) const { return vadj[v + 1] - vadj[v]; }
  __device__ __forceinline__ uint32_t entry(int32_t a) const { return A[a - aoff]; }
};

// Wall-clock stamps of the extension + support stage: the first workgroups
// of a stage kernel record ~start (the max of complements is the earliest
// start), every working workgroup its end; the host reads them from the
// spread lines instead of timing events, which cost a queue gap each.
__device__ __forceinline__ void stamp_start(const LevelArgs &a) {
  if (blockIdx.x < kSpread && threadIdx.x == 0)
    atomicMax(&a.spread[blockIdx.x * kSpreadStride + kSpreadStart],
              ~(unsigned long long)__builtin_amdgcn_s_memrealtime());
}
__device__ __forceinline__ void stamp_end(const LevelArgs &a) {
  if (threadIdx.x == 0)
    atomicMax(&a.spread[(blockIdx.x & (kSpread - 1)) * kSpreadStride + kSpreadEnd],
              (unsigned long long)__builtin_amdgcn_s_memrealtime());
}

// Degree of an adjacency entry's neighbour (packed in its top byte).
__device__ __forceinline__ int entry_deg(uint32_t e) { return (int)(e >> 24); }

// Whether graph vertex w is one of the first nv bytes of a tuple (words
// rec[0..)): a zero byte of rec ^ (w in every byte) marks a match; bytes at
// or past nv (padding) are masked out, and a borrow only moves towards
// higher (masked) bytes.
__device__ __forceinline__ bool tuple_has(const uint32_t *rec, int nv, int w) {
  const uint32_t pat = (uint32_t)w * 0x01010101u;
  for (int i = 0; i * 4 < nv; ++i) {
    const uint32_t x = rec[i] ^ pat;
    const int valid = min(4, nv - 4 * i);
    const uint32_t live = valid == 4 ? 0x80808080u : (0x80808080u & ((1u << (8 * valid)) - 1u));
    if ((x - 0x01010101u) & ~x & live) return true;
  }
  return false;
}

// 256-bit set of graph-local vertex ids held in four 64-bit values.
struct VSet {
  uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
  __device__ __forceinline__ void add(int v) {
    const uint64_t b = 1ull << (v & 63);
    const int q = v >> 6;
    m0 |= q == 0 ? b : 0ull;
    m1 |= q == 1 ? b : 0ull;
    m2 |= q == 2 ? b : 0ull;
    m3 |= q == 3 ? b : 0ull;
  }
  __device__ __forceinline__ bool has(int v) const {
    const int q = v >> 6;
    const uint64_t w = (m0 & (0ull - (uint64_t)(q == 0))) | (m1 & (0ull - (uint64_t)(q == 1))) |
                       (m2 & (0ull - (uint64_t)(q == 2))) | (m3 & (0ull - (uint64_t)(q == 3)));
    return (w >> (v & 63)) & 1ull;
  }
};

// Copy row r's tuple (S bytes) from HBM into this thread's LDS slot.
__device__ __forceinline__ void load_slot(uint32_t *slot, const uint8_t *vt, int64_t r, int S) {
  const uint32_t *p = reinterpret_cast<const uint32_t *>(vt + r * S);
  const int nw = S >> 2;
  for (int i = 0; i < nw; ++i) slot[i] = p[i];
}

__device__ __forceinline__ int slot_get(const uint32_t *slot, int i) {
  return reinterpret_cast<const uint8_t *>(slot)[i];
}

__device__ __forceinline__ void slot_set(uint32_t *slot, int i, int v) {
  reinterpret_cast<uint8_t *>(slot)[i] = (uint8_t)v;
}

// Write the slot (CS bytes) as child row `pos`.
__device__ __forceinline__ void store_slot(uint8_t *cvt, int64_t pos, int CS, const uint32_t *slot) {
  uint32_t *p = reinterpret_cast<uint32_t *>(cvt + pos * CS);
  const int nw = CS >> 2;
  for (int i = 0; i < nw; ++i) p[i] = slot[i];
}

// Copy the pattern's rightmost-path words into this thread's LDS slot
// (rounded up to a multiple of 4 words; 16-byte loads, all independent).
__device__ __forceinline__ void stage_rm(const PatMeta *__restrict__ pm, uint32_t *rmv) {
  const int nrm = pm->nrm;
  const uint4 *src = reinterpret_cast<const uint4 *>(pm->rm);
  for (int i = 0; i < nrm; i += 4) {
    const uint4 v = src[i >> 2];
    rmv[i] = v.x;
    rmv[i + 1] = v.y;
    rmv[i + 2] = v.z;
    rmv[i + 3] = v.w;
  }
}

// Enumerate every valid extension of the embedding in `slot` (tuple bytes;
// the pattern's rightmost path staged at rmv by stage_rm).
// emit(key, w, bound): extension key, new vertex (-1 backward) and the
// child's candidate bound.
template <class F>
__device__ __forceinline__ void enumerate_row(const GView &G, const PatMeta *__restrict__ pm, const uint32_t *slot,
                                              const uint32_t *rmv, int S, int NEL, int NVL, F &&emit) {
  const uint4 head = *reinterpret_cast<const uint4 *>(pm);
  const int nv = (int)(head.y & 0xFF);
  const int nrm = (int)((head.y >> 8) & 0xFF);
  const int minlabel = (int)((head.y >> 16) & 0xFF);
  const int maxlabel = (int)(head.y >> 24);
  const uint32_t bwd_mask = head.z;
  const int maxtoc = nv - 1;
  const int fwd_base = S * NEL;
  VSet vis;
  for (int i = 0; i < nv; ++i) vis.add(slot_get(slot, i));
  const int rv = slot_get(slot, maxtoc);
  int rm_deg = 0;  // degrees of the rightmost-path sources
  for (int i = 0; i < nrm; ++i) rm_deg += G.deg(slot_get(slot, rmv[i] & 0xFF));
  const int own = rm_deg + G.deg(rv);  // this row's own bound

  // backward and pure forward: one sweep over the rightmost vertex's edges
  for (int32_t a = G.begin(rv), ae = G.end(rv); a < ae; ++a) {
    const uint32_t e = G.entry(a);
    const int w = e & 0xFF, el = (e >> 8) & 0xFF, nl = (e >> 16) & 0xFF;
    if (!vis.has(w)) {
      if (nl >= minlabel) emit(fwd_base + (maxtoc * NEL + el) * NVL + nl, w, own + entry_deg(e));
    } else {
      for (int i = 1; i < nrm; ++i) {
        const uint32_t rme = rmv[i];
        const int f = rme & 0xFF;
        if (slot_get(slot, f) != w) continue;
        const int e1 = (rme >> 8) & 0xFF, l1 = (rme >> 16) & 0xFF;
        if (!((bwd_mask >> f) & 1u) && (e1 < el || (e1 == el && l1 <= maxlabel))) emit(f * NEL + el, -1, own);
        break;
      }
    }
  }
  // forward from the rightmost path
  int below = 0;  // degrees of the sources deeper than i
  for (int i = 0; i < nrm; ++i) {
    const uint32_t rme = rmv[i];
    const int f = rme & 0xFF;
    const int u = slot_get(slot, f);
    const int e1 = (rme >> 8) & 0xFF, l1 = (rme >> 16) & 0xFF;
    const int keep = rm_deg - below;  // sources i .. root
    for (int32_t a = G.begin(u), ae = G.end(u); a < ae; ++a) {
      const uint32_t e = G.entry(a);
      const int w = e & 0xFF, el = (e >> 8) & 0xFF, nl = (e >> 16) & 0xFF;
      if (vis.has(w) || nl < minlabel) continue;
      if (e1 < el || (e1 == el && l1 <= nl)) emit(fwd_base + (f * NEL + el) * NVL + nl, w, keep + entry_deg(e));
    }
    below += G.deg(u);
  }
}

// Root extensions of vertex u (1-edge codes with fromlabel <= tolabel).
template <class F>
__device__ __forceinline__ void enumerate_root(const GView &G, int u, int lu, int NEL, int NVL, F &&emit) {
  const int du = G.deg(u);
  for (int32_t a = G.begin(u), ae = G.end(u); a < ae; ++a) {
    const uint32_t e = G.entry(a);
    const int w = e & 0xFF, el = (e >> 8) & 0xFF, nl = (e >> 16) & 0xFF;
    if (lu <= nl) emit((lu * NEL + el) * NVL + nl, w, du + entry_deg(e));
  }
}

// Candidate list entries: {code, child bound}.
//   root codes   key | w << 24          (key < 255^3 < 2^24)
//   later codes  key | (w + 1) << 23    (key < 32 * 255 * 256 < 2^23,
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

// Dynamic LDS layout of the level kernels (32-bit words):
//   [0, 2 * 2^tb)          dedup table (k_support): 2^tb keys, 2^tb counts
//   blockDim * slot_words  per-thread slots: the row's tuple (tuple_words)
//                          then its rightmost path
//   stage_v                staged vertex offsets
//   stage_a                staged adjacency entries
struct Lds {
  uint32_t *table;
  uint32_t *slots;
  int32_t *vadj;
  uint32_t *adj;
};

__device__ __forceinline__ Lds lds_layout(const LevelArgs &a, bool support) {
  extern __shared__ uint32_t dyn[];
  Lds l;
  uint32_t *p = dyn;
  l.table = p;
  if (support) p += 2u << a.table_bits;
  l.slots = p;
  p += blockDim.x * a.slot_words;
  l.vadj = reinterpret_cast<int32_t *>(p);
  p += a.stage_v;
  l.adj = p;
  return l;
}

// Whether graph g's adjacency fits the LDS stage (uniform per workgroup).
__device__ __forceinline__ bool graph_fits(const LevelArgs &a, int g) {
  const int32_t vb = a.gvbase[g], ve = a.gvbase[g + 1];
  return ve - vb + 1 <= a.stage_v && a.vadj[ve] - a.vadj[vb] <= a.stage_a;
}

template <bool LDS>
__device__ __forceinline__ GView graph_view(const LevelArgs &a, int g, const Lds &l) {
  const int32_t vb = a.gvbase[g], ve = a.gvbase[g + 1];
  const int32_t ab = a.vadj[vb], ae = a.vadj[ve];
  GView G;
  if (LDS) {
    for (int i = threadIdx.x; i <= ve - vb; i += blockDim.x) l.vadj[i] = a.vadj[vb + i];
    for (int i = threadIdx.x; i < ae - ab; i += blockDim.x) l.adj[i] = a.adj[ab + i];
    __syncthreads();
    G.vadj = l.vadj;
    G.A = l.adj;
    G.aoff = ab;
  } else {
    G.vadj = a.vadj + vb;
    G.A = a.adj;
    G.aoff = 0;
  }
  return G;
}

// Block-wide exclusive scan of one value per thread (wave64 shuffles + LDS;
// s holds blockDim / 64 + 1 values).
template <class T>
__device__ __forceinline__ T block_exclusive_scan(T x, T &total, T *s) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  T v = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  if (lane == 63) s[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    T acc = 0;
    for (int i = 0; i < nw; ++i) {
      T t = s[i];
      s[i] = acc;
      acc += t;
    }
    s[nw] = acc;
  }
  __syncthreads();
  T r = s[wid] + v - x;
  total = s[nw];
  __syncthreads();
  return r;
}

// Block-wide sum (the result in every thread; s holds blockDim / 64 values).
template <class T>
__device__ __forceinline__ T block_sum(T x, T *s) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = x;
  __syncthreads();
  T t = 0;
  for (int i = 0; i < (int)((blockDim.x + 63) >> 6); ++i) t += s[i];
  __syncthreads();
  return t;
}

// Wave-level helpers (one wave synchronises with itself only).
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


__device__ __forceinline__ uint32_t hash_slot32(uint32_t k, int bits) { return (k * 0x9E3779B1u) >> (32 - bits); }

// Rows of graph g at the current level (root level: its vertices).
template <bool ROOT>
__device__ __forceinline__ void graph_rows(const LevelArgs &a, int g, int64_t &r0, int64_t &r1) {
  if (ROOT) {
    r0 = 0;
    r1 = a.gvbase[g + 1] - a.gvbase[g];
  } else {
    r0 = a.gstart[g];
    r1 = r0 + a.gcount[g];
  }
}

// ---------------------------------------------------------------- SUPPORT
// One lane per row: the lane stages the row's tuple and rightmost path in
// its LDS slot, enumerates the row's extensions (enumerate_row) and inserts
// each candidate's key (p - p_lo) * K + key into the LDS hash table, which
// counts the candidates per key.  With candidate lists on, the lane also
// appends each candidate to the row's list (its region holds the row's
// bound, the most candidates the row can have).
//
// Passes and splits: one pass per residue class of pattern ids (root level:
// of the vertex label).  When the table passes 3/4 full the pass is
// abandoned and the class split in two; all keys of one pattern stay in one
// class, so the per-graph dedup stays exact.  The first pass writes every
// row's list even when it fails, so with lists on the split passes read the
// lists back instead of enumerating again.  A completed pass adds
// (1 << 40 | count) per distinct key to the graph's replica of the packed
// counters (replica = workgroup id mod a.rep; k_reduce sums the replicas).

constexpr int kSupShift = 40;  // packed counters: support << 40 | embeddings

struct SupportShared {
  int pre[kMaxBlock];  // list replay: prefix of the chunk's list lengths
  int ired[5];
  unsigned long long red[5];
  int fill, over, top, replay;
  uint32_t stack_mod[32], stack_res[32];
};

// Largest k in [0, n) with pre[k] <= j (pre non-decreasing, pre[0] = 0).
__device__ __forceinline__ int find_row(const int *pre, int j, int n) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= j) lo = mid;
    else hi = mid;
  }
  return lo;
}

// ----------------------------------------------------------------- EXPAND
// Later levels with candidate lists: the extensions of every row are listed
// before any deduplication, fully in parallel over the level's rows: one
// 64-lane workgroup (a single wave) per 64 consecutive rows, whatever their
// graphs.  Setup: lane t stages row t's record (tuple, rightmost path, item
// prefix per source, row facts) in LDS and its visited mask.  Items: the
// wave's items -- one per adjacency entry of each row's sources (the
// rightmost vertex, then the rightmost-path sources) -- are taken 64 at a
// time, one per lane (row by binary search over the item prefix); each lane
// applies the extension rule to its adjacency entry.  A row's candidates
// are appended to its list in item order, positions from a ballot over the
// window; lane t keeps row t's running count.  The adjacency is read from
// HBM (rows of one wave may belong to different graphs).
//
// Row record (32-bit words): [0, TW) tuple bytes; [TW, TW + RMW) rightmost
// path; [TW + RMW, + SPW) item prefix per source: 0, deg(rv), + deg(rm 0),
// ...; then pattern id, PatMeta word 1 (nv, nrm, minlabel, maxlabel),
// bwd_mask, list index (2 words); then the first adjacency entry of each
// source (SPW words).  Whether a neighbour is already in the embedding is
// read from the tuple bytes (tuple_has).
constexpr int kExpandRows = 64;
constexpr int kExpandUnroll = 4;  // item windows whose loads are issued together

__global__ void __launch_bounds__(kExpandRows) k_expand(LevelArgs a) {
  stamp_start(a);
  __shared__ int s_pre[kExpandRows];
  extern __shared__ uint32_t dyn[];
  const int t = threadIdx.x;
  const int TW = a.tuple_words, RMW = a.rm_words, SPW = a.src_words;
  uint32_t *recs = dyn;
  const int64_t r0 = (int64_t)blockIdx.x * kExpandRows;
  const int64_t r = r0 + t;
  const int nrow = (int)min<int64_t>(kExpandRows, a.n_rows - r0);
  const int NEL = a.NEL, NVL = a.NVL;
  const int fwd_base = a.S * NEL;
  uint32_t *mine = recs + t * a.slot_words;
  int n = 0;
  bool has_list = false;  // row t's graph kept lists (else k_support enumerates it)
  if (t < nrow) {
    const int g = a.rgraph[r];
    const long long gb = a.gcbase[g];
    const int32_t p = gb >= 0 ? a.pid[r] : 0;
    const PatMeta *pm = a.meta + p;
    const uint4 head = gb >= 0 ? *reinterpret_cast<const uint4 *>(pm) : make_uint4(0, 0, 0, 0);
    has_list = gb >= 0;
    if (gb >= 0 && !(head.w & kPatDead)) {  // a dead pattern's row lists nothing
      const int nv = (int)(head.y & 0xFF), nrm = (int)((head.y >> 8) & 0xFF);
      const int32_t vb = a.gvbase[g];
      GView G;
      G.vadj = a.vadj + vb;
      G.A = a.adj;
      G.aoff = 0;
      load_slot(mine, a.vt, r, a.S);
      uint32_t *rmv = mine + TW;
      stage_rm(pm, rmv);
      uint32_t *sp = rmv + RMW;
      uint32_t *misc = sp + SPW;
      uint32_t *sb = misc + 5;  // first adjacency entry of each source
      const int rvv = slot_get(mine, nv - 1);
      sb[0] = (uint32_t)G.begin(rvv);
      int acc = G.end(rvv) - (int)sb[0];
      sp[0] = 0;
      sp[1] = (uint32_t)acc;
      for (int i = 0; i < nrm; ++i) {
        const int u = slot_get(mine, rmv[i] & 0xFF);
        sb[i + 1] = (uint32_t)G.begin(u);
        acc += G.end(u) - (int)sb[i + 1];
        sp[i + 2] = (uint32_t)acc;
      }
      const unsigned long long li = (unsigned long long)gb + a.rcoff[r];
      misc[0] = (uint32_t)p;
      misc[1] = head.y;
      misc[2] = head.z;
      misc[3] = (uint32_t)li;
      misc[4] = (uint32_t)(li >> 32);
      n = acc;
    }
  }
  int T;
  const int pre = wave_exclusive_scan<int>(n, T);
  s_pre[t] = pre;
  wave_sync();
  int cnt = 0;  // candidates of row t listed so far
  unsigned long long cand = 0, bound_sum = 0, pruned = 0;
  const int ms = a.mask_stride;
  // kExpandUnroll windows of 64 items per step: every window's adjacency
  // entry is loaded first, then the windows are evaluated and compacted in
  // order (a row's list positions follow its items)
  for (int jb0 = 0; jb0 < T; jb0 += kExpandUnroll * kExpandRows) {
    int kk[kExpandUnroll], srcs[kExpandUnroll], ii[kExpandUnroll];
    uint32_t ee[kExpandUnroll];
#pragma unroll
    for (int u = 0; u < kExpandUnroll; ++u) {
      const int j = jb0 + u * kExpandRows + t;
      kk[u] = 0;
      srcs[u] = 0;
      ii[u] = 0;
      ee[u] = 0;
      if (j < T) {
        const int k = find_row(s_pre, j, nrow);
        const int i = j - s_pre[k];
        const uint32_t *sp = recs + k * a.slot_words + TW + RMW;
        const uint32_t *sb = sp + SPW + 5;
        int src = 0;  // 0: the rightmost vertex; s >= 1: rightmost-path source s - 1
        while ((int)sp[src + 1] <= i) ++src;
        kk[u] = k;
        srcs[u] = src;
        ii[u] = i;
        ee[u] = a.adj[sb[src] + (i - (int)sp[src])];
      }
    }
#pragma unroll
    for (int u = 0; u < kExpandUnroll; ++u) {
      const int jb = jb0 + u * kExpandRows;
      const int j = jb + t;
      const int k = kk[u];
      int key = -1, w = 0, bound = 0;
      bool drop = false;  // a valid extension the path rule removes
      if (j < T) {
        const int i = ii[u], src = srcs[u];
        const uint32_t e = ee[u];
        const uint32_t *rec = recs + k * a.slot_words;
        const uint32_t *rmv = rec + TW;
        const uint32_t *sp = rmv + RMW;
        const uint32_t *misc = sp + SPW;
        const uint32_t hy = misc[1];
        const int nv = (int)(hy & 0xFF), nrm = (int)((hy >> 8) & 0xFF);
        const int minlabel = (int)((hy >> 16) & 0xFF), maxlabel = (int)(hy >> 24);
        const int maxtoc = nv - 1;
        const int own = (int)sp[nrm + 1];
        const unsigned long long *mk = ms ? a.pmask + (int64_t)misc[0] * 2 * ms : nullptr;
        w = e & 0xFF;
        const int el = (e >> 8) & 0xFF, nl = (e >> 16) & 0xFF;
        if (src == 0) {
          if (!tuple_has(rec, nv, w)) {
            if (nl >= minlabel) {
              key = fwd_base + (maxtoc * NEL + el) * NVL + nl;
              bound = own + entry_deg(e);
              drop = mk && !((mk[0] >> (el * NVL + nl)) & 1ull);
            }
          } else {
            const uint32_t bwd_mask = misc[2];
            for (int q = 1; q < nrm; ++q) {
              const uint32_t rme = rmv[q];
              const int f = rme & 0xFF;
              if (slot_get(rec, f) != w) continue;
              const int e1 = (rme >> 8) & 0xFF, l1 = (rme >> 16) & 0xFF;
              if (!((bwd_mask >> f) & 1u) && (e1 < el || (e1 == el && l1 <= maxlabel))) {
                key = f * NEL + el;
                bound = own;
                drop = mk && !((mk[ms + 1 + q] >> el) & 1ull);
              }
              break;
            }
            w = -1;
          }
        } else {
          const uint32_t rme = rmv[src - 1];
          const int f = rme & 0xFF;
          const int e1 = (rme >> 8) & 0xFF, l1 = (rme >> 16) & 0xFF;
          if (nl >= minlabel && (e1 < el || (e1 == el && l1 <= nl)) && !tuple_has(rec, nv, w)) {
            key = fwd_base + (f * NEL + el) * NVL + nl;
            bound = own - (int)sp[src] + entry_deg(e);  // sources src - 1 .. root, plus w
            drop = mk && !((mk[src] >> (el * NVL + nl)) & 1ull);
          }
        }
        if (drop) {  // counted as a candidate, never listed
          key = -1;
          ++pruned;
        }
      }
      const bool emit = key >= 0;
      const unsigned long long mask = __ballot(emit);
      const int first = j < T ? max(s_pre[k], jb) - jb : 0;  // lane of row k's first item in the window
      const unsigned long long below = mask & ((1ull << t) - 1) & ~((1ull << first) - 1);
      const int listed = __shfl(cnt, k, kExpandRows);
      if (emit) {
        const uint32_t *misc = recs + k * a.slot_words + TW + RMW + SPW;
        const unsigned long long li = (unsigned long long)misc[3] | ((unsigned long long)misc[4] << 32);
        a.cand[li + listed + __popcll(below)] = make_uint2(cand_code(false, key, w), (uint32_t)bound);
        ++cand;
        bound_sum += (unsigned long long)bound;
      }
      // row t's items in this window: lanes [lo, hi)
      const int lo = max(pre, jb) - jb, hi = min(pre + n, jb + kExpandRows) - jb;
      if (hi > lo) {
        const unsigned long long range = (hi - lo == 64 ? ~0ull : (((1ull << (hi - lo)) - 1) << lo));
        cnt += __popcll(mask & range);
      }
    }
  }
  if (has_list) a.rccnt[r] = (uint16_t)cnt;
  // statistics: one atomic per wave each, spread over kSpread lines
  cand = wave_sum(cand + pruned);
  bound_sum = wave_sum(bound_sum);
  pruned = wave_sum(pruned);
  if (t == 0) {
    unsigned long long *line = a.spread + (blockIdx.x & (kSpread - 1)) * kSpreadStride;
    if (cand) atomicAdd(&line[kSpreadCand], cand);
    if (bound_sum) atomicAdd(&line[kSpreadBound], bound_sum);
    if (pruned) atomicAdd(&line[kSpreadPruned], pruned);
  }
}

constexpr int kReplayUnroll = 4;  // list entries per lane per step of the list loops

template <bool ROOT, bool LDS>
__device__ __forceinline__ void support_body(const LevelArgs &a, int g, int64_t r0, int64_t r1, const Lds &l,
                                             SupportShared &sh) {
  const int job = ROOT ? a.gjob[g] : 0;
  const int32_t vb = a.gvbase[g];
  uint2 *lists = nullptr;
  if (a.use_lists) {
    if (ROOT) {
      lists = a.cand;
    } else {
      const long long b = a.gcbase[g];
      if (b >= 0) lists = a.cand + b;
    }
  }
  // k_expand already listed this graph's candidates (and counted them):
  // only the dedup passes remain, reading the lists
  const bool listed = !ROOT && a.lists_ready && lists != nullptr;
  GView G{};
  if (!listed) G = graph_view<LDS>(a, g, l);
  const int B = blockDim.x;
  int tb = a.table_bits;
  if (listed) {
    // the graph's candidates bound its distinct keys: a table just large
    // enough for them is cheaper to clear and to commit
    unsigned long long ents = 0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
      const int32_t p = a.pid[r];
      if (p >= a.p_lo && p < a.p_hi) ents += a.rccnt[r];
    }
    ents = block_sum<unsigned long long>(ents, sh.red);
    int want = 6;
    while (want < tb && ((1ull << want) * 3) / 4 < ents) ++want;
    tb = want;
  }
  const uint32_t cap = 1u << tb, fill_limit = cap - cap / 4;
  uint32_t *table = l.table;
  uint32_t *tcnt = l.table + cap;
  uint32_t *slot = l.slots + threadIdx.x * a.slot_words;
  uint32_t *rmv = slot + a.tuple_words;
  unsigned long long *rep = a.acc_rep + (int64_t)(g & (a.rep - 1)) * a.nkeys;
  unsigned long long my_cand = 0, my_bound = 0;
  if (threadIdx.x == 0) {
    sh.top = 1;
    sh.stack_mod[0] = 1;
    sh.stack_res[0] = 0;
    sh.replay = listed ? 1 : 0;
  }
  for (;;) {
    __syncthreads();
    if (sh.top == 0) break;
    const uint32_t mod = sh.stack_mod[sh.top - 1], res = sh.stack_res[sh.top - 1];
    for (uint32_t i = threadIdx.x; i < cap; i += B) {
      table[i] = kEmpty32;
      tcnt[i] = 0;
    }
    if (threadIdx.x == 0) {
      sh.fill = 0;
      sh.over = 0;
    }
    __syncthreads();
    unsigned long long cand = 0, bound_sum = 0;
    int over = 0;  // this thread saw the table pass its fill limit
    // once the pass has failed (flag published in LDS) nothing more is
    // inserted: a nearly full table would make every probe sequence long
    auto failed_now = [&]() { return __hip_atomic_load(&sh.over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); };
    auto set_failed = [&]() { __hip_atomic_store(&sh.over, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); };
    auto insert = [&](uint32_t k) {
      if (over || failed_now()) return;
      uint32_t h = hash_slot32(k, tb);
      for (uint32_t probe = 0; probe < cap; ++probe) {
        const uint32_t old = atomicCAS(&table[h], kEmpty32, k);
        if (old == kEmpty32 || old == k) {
          atomicAdd(&tcnt[h], 1u);
          if (old == kEmpty32 && (uint32_t)atomicAdd(&sh.fill, 1) >= fill_limit) {
            over = 1;
            set_failed();
          }
          return;
        }
        h = (h + 1) & (cap - 1);
      }
      over = 1;
      set_failed();
    };
    if (!sh.replay) {
      for (int64_t r = r0 + threadIdx.x; r < r1; r += B) {
        int nc = 0;
        if (ROOT) {
          const int u = (int)r;
          const int lu = a.vlab[vb + u];
          if (mod > 1 && ((uint32_t)lu & (mod - 1)) != res) continue;
          const uint32_t base = (uint32_t)((int64_t)(job - a.p_lo) * a.K);
          uint2 *out = lists ? lists + G.begin(u) : nullptr;
          enumerate_root(G, u, lu, a.NEL, a.NVL, [&](int key, int w, int bound) {
            insert(base + (uint32_t)key);
            ++cand;
            bound_sum += (unsigned long long)bound;
            if (out) out[nc++] = make_uint2(cand_code(true, key, w), (uint32_t)bound);
          });
          if (out) a.rccnt[vb + u] = (uint16_t)nc;
        } else {
          const int32_t p = a.pid[r];
          if (p < a.p_lo || p >= a.p_hi) continue;
          if (mod > 1 && ((uint32_t)p & (mod - 1)) != res) continue;
          const PatMeta *pm = a.meta + p;
          uint2 *out = lists ? lists + a.rcoff[r] : nullptr;
          if (pm->flags & kPatDead) {  // not extended
            if (out) a.rccnt[r] = 0;
            continue;
          }
          load_slot(slot, a.vt, r, a.S);
          stage_rm(pm, rmv);
          const uint32_t base = (uint32_t)((int64_t)(p - a.p_lo) * a.K);
          enumerate_row(G, pm, slot, rmv, a.S, a.NEL, a.NVL, [&](int key, int w, int bound) {
            insert(base + (uint32_t)key);
            ++cand;
            bound_sum += (unsigned long long)bound;
            if (out) out[nc++] = make_uint2(cand_code(false, key, w), (uint32_t)bound);
          });
          if (out) a.rccnt[r] = (uint16_t)nc;
        }
      }
    } else {
      // read the lists back, lane-parallel: rows blockDim at a time, each
      // lane takes list entries found by binary search over the prefix
      for (int64_t base = r0; base < r1; base += B) {
        const int64_t r = base + threadIdx.x;
        int n = 0;
        if (r < r1) {
          if (ROOT) {
            const int u = (int)r;
            const int lu = a.vlab[vb + u];
            if (!(mod > 1 && ((uint32_t)lu & (mod - 1)) != res)) {
              n = a.rccnt[vb + u];
              slot[0] = (uint32_t)G.begin(u);
              slot[1] = (uint32_t)((int64_t)(job - a.p_lo) * a.K);
            }
          } else {
            const int32_t p = a.pid[r];
            if (p >= a.p_lo && p < a.p_hi && !(mod > 1 && ((uint32_t)p & (mod - 1)) != res)) {
              n = a.rccnt[r];
              slot[0] = a.rcoff[r];
              slot[1] = (uint32_t)((int64_t)(p - a.p_lo) * a.K);
            }
          }
        }
        int T;
        sh.pre[threadIdx.x] = block_exclusive_scan<int>(n, T, sh.ired);
        __syncthreads();
        const int nrow = (int)min<int64_t>(B, r1 - base);
        // kReplayUnroll entries per lane per step, their list loads issued
        // together (each entry's row is found by binary search first)
        for (int j0 = threadIdx.x; j0 < T; j0 += kReplayUnroll * B) {
          uint2 c[kReplayUnroll];
          uint32_t kb[kReplayUnroll];
#pragma unroll
          for (int u = 0; u < kReplayUnroll; ++u) {
            const int j = j0 + u * B;
            if (j < T) {
              const int k = find_row(sh.pre, j, nrow);
              const uint32_t *rec = l.slots + k * a.slot_words;
              c[u] = lists[rec[0] + (j - sh.pre[k])];
              kb[u] = rec[1];
            }
          }
#pragma unroll
          for (int u = 0; u < kReplayUnroll; ++u) {
            if (j0 + u * B < T) {
              int key, w;
              cand_decode(ROOT, c[u].x, key, w);
              insert(kb[u] + (uint32_t)key);
              if (!listed) {
                ++cand;
                bound_sum += c[u].y;
              }
            }
          }
        }
        __syncthreads();  // the next chunk reuses the slots
      }
    }
    if (over) atomicOr(&sh.over, 1);
    __syncthreads();
    const bool failed = sh.over != 0;
    __syncthreads();
    if (failed) {
      if (threadIdx.x == 0) {
        atomicAdd(a.split_passes, 1ull);
        if (lists) sh.replay = 1;  // every row's list is complete
        // the next level wants a larger table
        atomicMax(&a.spread[(g & (kSpread - 1)) * kSpreadStride + kSpreadFill], 2ull * cap);
        if (mod >= (1u << 30) || sh.top + 1 > 32) {
          atomicOr(a.error_flag, 1);
          sh.top = 0;
        } else {
          sh.stack_mod[sh.top - 1] = mod * 2;
          sh.stack_res[sh.top - 1] = res + mod;
          sh.stack_mod[sh.top] = mod * 2;
          sh.stack_res[sh.top] = res;
          sh.top += 1;
        }
      }
      continue;
    }
    my_cand += cand;
    my_bound += bound_sum;
    for (uint32_t i = threadIdx.x; i < cap; i += B) {
      const uint32_t k = table[i];
      if (k == kEmpty32) continue;
      atomicAdd(&rep[(int64_t)a.p_lo * a.K + (int64_t)k], (1ull << kSupShift) | (unsigned long long)tcnt[i]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      atomicMax(&a.spread[(g & (kSpread - 1)) * kSpreadStride + kSpreadFill], (unsigned long long)sh.fill);
      sh.top -= 1;
    }
  }
  // statistics: one atomic per workgroup each, spread over kSpread lines
  my_cand = block_sum<unsigned long long>(my_cand, sh.red);
  my_bound = block_sum<unsigned long long>(my_bound, sh.red);
  if (threadIdx.x == 0) {
    unsigned long long *line = a.spread + (g & (kSpread - 1)) * kSpreadStride;
    if (my_cand) atomicAdd(&line[kSpreadCand], my_cand);
    if (my_bound) atomicAdd(&line[kSpreadBound], my_bound);
  }
}

// Graph of this workgroup: the small-graph launch covers every graph id and
// leaves graphs of more than big_rows rows to the big-graph launch, which
// runs over the list k_scan built (-1: past the list's end).
template <bool BIG>
__device__ __forceinline__ int pick_graph(const LevelArgs &a) {
  if (!BIG) return blockIdx.x;
  return blockIdx.x < (unsigned)*a.nbig ? a.big[blockIdx.x] : -1;
}

template <bool ROOT, bool BIG>
__global__ void __launch_bounds__(kMaxBlock) k_support(LevelArgs a) {
  __shared__ SupportShared sh;
  if (!BIG) stamp_start(a);  // levels without k_expand start here
  const int g = pick_graph<BIG>(a);
  if (g < 0) return;
  if (ROOT) {
    const int job = a.gjob[g];
    if (job < a.p_lo || job >= a.p_hi) return;
  }
  int64_t r0, r1;
  graph_rows<ROOT>(a, g, r0, r1);
  if (r0 == r1) return;
  if (!ROOT && !BIG && r1 - r0 > a.big_rows) return;
  const Lds l = lds_layout(a, true);
  if (graph_fits(a, g))
    support_body<ROOT, true>(a, g, r0, r1, l, sh);
  else
    support_body<ROOT, false>(a, g, r0, r1, l, sh);
  stamp_end(a);
}

// Merge one graph's count for key k into a workgroup's accumulation table
// (keys, graphs, embeddings; 2^bits entries each).  New keys stop being
// admitted at 3/4 fill; returns false when k is not in the table then (the
// caller commits it to HBM itself).
__device__ __forceinline__ bool acc_add(uint32_t *acc, int bits, int *fill, uint32_t k, uint32_t emb) {
  const uint32_t cap = 1u << bits, limit = cap - cap / 4;
  uint32_t h = hash_slot32(k, bits);
  for (uint32_t probe = 0; probe < cap; ++probe) {
    uint32_t cur = __hip_atomic_load(&acc[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == kEmpty32) {
      if (__hip_atomic_load(fill, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= (int)limit) return false;
      cur = atomicCAS(&acc[h], kEmpty32, k);
      if (cur == kEmpty32) {
        atomicAdd(fill, 1);
        cur = k;
      }
    }
    if (cur == k) {
      atomicAdd(&acc[cap + h], 1u);
      atomicAdd(&acc[2 * cap + h], emb);
      return true;
    }
    h = (h + 1) & (cap - 1);
  }
  return false;
}

// ------------------------------------------------------------ ROOT WAVES
// The root level deduplicates one graph per wave: a persistent grid of
// workgroups, each over a contiguous range of graph ids that its waves take
// in turn, each wave with its own LDS table, so no workgroup barrier is
// needed and the kernel arguments are loaded once per wave rather than once
// per graph -- databases of many small graphs (the batched job mix:
// hundreds of thousands) spend most of a per-graph workgroup's life on that
// fixed cost.  A graph's distinct keys are merged into the workgroup's LDS
// accumulation table (graphs of one job share their root keys), committed
// to the packed counters once at the end: one HBM atomic per key and
// workgroup instead of one per key and graph.  Each lane enumerates one vertex
// (adjacency read from HBM) and writes its list at the vertex's adjacency
// offset.  Overflow splits the graph's pass by residue class of the vertex
// label as in support_body; with lists, the split passes read them back.
constexpr int kWaveBlock = 256;    // 4 waves per workgroup
constexpr int kWaveTableBits = 9;  // per-wave dedup table: 512 keys
constexpr int kRootAccBits = 9;    // workgroup accumulation table: 512 keys

// Per-wave LDS (32-bit words): table keys [cap], counts [cap], item prefix
// [64], row records [64 x 2] (list offset, key base), split stack [2 x 32],
// fill and overflow flags.
constexpr int kWaveWords = (2 << kWaveTableBits) + 64 + 128 + 64 + 2;

__global__ void __launch_bounds__(kWaveBlock) k_root_wave(LevelArgs a, int n_graphs) {
  __shared__ uint32_t wl[(kWaveBlock / 64) * kWaveWords];
  __shared__ uint32_t racc[3 << kRootAccBits];
  __shared__ int racc_fill;
  for (int i = threadIdx.x; i < (1 << kRootAccBits); i += kWaveBlock) {
    racc[i] = kEmpty32;
    racc[(1 << kRootAccBits) + i] = 0;
    racc[(2 << kRootAccBits) + i] = 0;
  }
  if (threadIdx.x == 0) racc_fill = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t *base = wl + wid * kWaveWords;
  uint32_t *table = base;
  uint32_t *tcnt = base + (1 << kWaveTableBits);
  int *pre = reinterpret_cast<int *>(base + (2 << kWaveTableBits));
  uint32_t *rec = base + (2 << kWaveTableBits) + 64;
  uint32_t *stack_mod = rec + 128, *stack_res = rec + 160;
  int *fill = reinterpret_cast<int *>(rec + 192);
  int *over_flag = fill + 1;
  const int NEL = a.NEL, NVL = a.NVL;
  unsigned long long my_cand = 0, my_bound = 0, my_fill = 0;
  const int chunk = (n_graphs + gridDim.x - 1) / gridDim.x;
  const int g_end = min(n_graphs, (int)(blockIdx.x + 1) * chunk);
  for (int g = blockIdx.x * chunk + wid; g < g_end; g += kWaveBlock / 64) {
    const int32_t vb = a.gvbase[g];
    const int nvert = a.gvbase[g + 1] - vb;
    if (nvert == 0) continue;
    const int job = a.gjob[g];
    if (job < a.p_lo || job >= a.p_hi) continue;
    uint2 *lists = a.use_lists ? a.cand : nullptr;
    GView G;
    G.vadj = a.vadj + vb;
    G.A = a.adj;
    G.aoff = 0;
    // the graph's adjacency entries bound its distinct root keys
    const unsigned long long ents = (unsigned long long)(G.end(nvert - 1) - G.begin(0));
    int tb = 6;
    while (tb < kWaveTableBits && ((1ull << tb) * 3) / 4 < ents) ++tb;
    const uint32_t cap = 1u << tb, fill_limit = cap - cap / 4;
    unsigned long long *rep = a.acc_rep + (int64_t)(g & (a.rep - 1)) * a.nkeys;
    const uint32_t kb = (uint32_t)((int64_t)(job - a.p_lo) * a.K);
    bool replay = false;
    int top = 1;
    if (lane == 0) {
      stack_mod[0] = 1;
      stack_res[0] = 0;
    }
    wave_sync();
    while (top > 0) {
      const uint32_t mod = stack_mod[top - 1], res = stack_res[top - 1];
      for (uint32_t i = lane; i < cap; i += 64) {
        table[i] = kEmpty32;
        tcnt[i] = 0;
      }
      if (lane == 0) {
        *fill = 0;
        *over_flag = 0;
      }
      wave_sync();
      unsigned long long cand = 0, bound_sum = 0;
      int over = 0;
      auto insert = [&](uint32_t k) {
        if (over || __hip_atomic_load(over_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) return;
        uint32_t h = hash_slot32(k, tb);
        for (uint32_t probe = 0; probe < cap; ++probe) {
          const uint32_t old = atomicCAS(&table[h], kEmpty32, k);
          if (old == kEmpty32 || old == k) {
            atomicAdd(&tcnt[h], 1u);
            if (old == kEmpty32 && (uint32_t)atomicAdd(fill, 1) >= fill_limit) {
              over = 1;
              __hip_atomic_store(over_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            return;
          }
          h = (h + 1) & (cap - 1);
        }
        over = 1;
        __hip_atomic_store(over_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      };
      if (!replay) {  // one vertex per lane
        for (int u = lane; u < nvert; u += 64) {
          const int lu = a.vlab[vb + u];
          if (mod > 1 && ((uint32_t)lu & (mod - 1)) != res) continue;
          uint2 *out = lists ? lists + G.begin(u) : nullptr;
          int nc = 0;
          enumerate_root(G, u, lu, NEL, NVL, [&](int key, int w, int bound) {
            insert(kb + (uint32_t)key);
            ++cand;
            bound_sum += (unsigned long long)bound;
            if (out) out[nc++] = make_uint2(cand_code(true, key, w), (uint32_t)bound);
          });
          if (out) a.rccnt[vb + u] = (uint16_t)nc;
        }
      } else {  // read the lists back, one entry per lane
        for (int cb = 0; cb < nvert; cb += 64) {
          const int u = cb + lane;
          int n = 0;
          if (u < nvert) {
            const int lu = a.vlab[vb + u];
            if (!(mod > 1 && ((uint32_t)lu & (mod - 1)) != res)) {
              n = a.rccnt[vb + u];
              rec[2 * lane] = (uint32_t)G.begin(u);
            }
          }
          int T;
          pre[lane] = wave_exclusive_scan<int>(n, T);
          wave_sync();
          const int nrow = min(64, nvert - cb);
          for (int j = lane; j < T; j += 64) {
            const int k = find_row(pre, j, nrow);
            const uint2 c = lists[rec[2 * k] + (j - pre[k])];
            int key, w;
            cand_decode(true, c.x, key, w);
            insert(kb + (uint32_t)key);
            ++cand;
            bound_sum += c.y;
          }
          wave_sync();  // the next chunk reuses pre and rec
        }
      }
      wave_sync();
      const bool failed = __hip_atomic_load(over_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0 ||
                          __any(over);
      if (failed) {
        const bool give_up = mod >= (1u << 30) || top + 1 > 32;
        if (lane == 0) {
          atomicAdd(a.split_passes, 1ull);
          if (give_up) {
            atomicOr(a.error_flag, 1);
          } else {
            stack_mod[top - 1] = mod * 2;
            stack_res[top - 1] = res + mod;
            stack_mod[top] = mod * 2;
            stack_res[top] = res;
          }
        }
        my_fill = max(my_fill, 2ull * cap);
        if (give_up) break;
        top += 1;
        if (lists) replay = true;  // every vertex's list is complete
        wave_sync();
        continue;
      }
      my_cand += cand;
      my_bound += bound_sum;
      for (uint32_t i = lane; i < cap; i += 64) {
        const uint32_t k = table[i];
        if (k == kEmpty32) continue;
        if (acc_add(racc, kRootAccBits, &racc_fill, k, tcnt[i])) continue;
        atomicAdd(&rep[(int64_t)a.p_lo * a.K + (int64_t)k], (1ull << kSupShift) | (unsigned long long)tcnt[i]);
      }
      my_fill = max(my_fill, (unsigned long long)*fill);
      top -= 1;
      wave_sync();
    }
  }
  // the workgroup's accumulated keys
  __syncthreads();
  {
    unsigned long long *rep = a.acc_rep + (int64_t)(blockIdx.x & (a.rep - 1)) * a.nkeys;
    for (int i = threadIdx.x; i < (1 << kRootAccBits); i += kWaveBlock) {
      const uint32_t k = racc[i];
      if (k == kEmpty32) continue;
      atomicAdd(&rep[(int64_t)a.p_lo * a.K + (int64_t)k],
                ((unsigned long long)racc[(1 << kRootAccBits) + i] << kSupShift) |
                    (unsigned long long)racc[(2 << kRootAccBits) + i]);
    }
  }
  // statistics: one atomic per wave each, spread over kSpread lines
  my_cand = wave_sum(my_cand);
  my_bound = wave_sum(my_bound);
  if (lane == 0) {
    unsigned long long *line =
        a.spread + ((blockIdx.x * (kWaveBlock / 64) + wid) & (kSpread - 1)) * kSpreadStride;
    if (my_cand) atomicAdd(&line[kSpreadCand], my_cand);
    if (my_bound) atomicAdd(&line[kSpreadBound], my_bound);
    if (my_fill) atomicMax(&line[kSpreadFill], my_fill);
  }
}

// Sum the replicas of every key (and clear them for the next level); keys
// reaching their job's min support go to the frequent list as
// (key, support, embeddings).
template <bool ROOT>
__global__ void __launch_bounds__(256) k_reduce(LevelArgs a) {
  // the row-parallel children passes' tile sums start at zero
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n_wtiles; i += gridDim.x * blockDim.x) {
    a.trows[(int64_t)i * kTileStride] = 0;
    a.tents[(int64_t)i * kTileStride] = 0;
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.nkeys; i += (int64_t)gridDim.x * blockDim.x) {
    // eight replicas' loads in flight before any of them is cleared (a
    // load after a store to the same array waits for it)
    unsigned long long v = 0;
    int r = 0;
    for (; r + 8 <= a.rep; r += 8) {
      unsigned long long x[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) x[q] = a.acc_rep[(int64_t)(r + q) * a.nkeys + i];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        v += x[q];
        a.acc_rep[(int64_t)(r + q) * a.nkeys + i] = 0;
      }
    }
    for (; r < a.rep; ++r) {
      const int64_t j = (int64_t)r * a.nkeys + i;
      v += a.acc_rep[j];
      a.acc_rep[j] = 0;
    }
    const int32_t s = (int32_t)(v >> kSupShift);
    const unsigned long long c = v & ((1ull << kSupShift) - 1);
    a.sup[i] = s;
    const int64_t p = i / a.K;
    const int32_t owner = ROOT ? (int32_t)p : a.meta[p].job;
    const bool hit = s >= a.minsup_job[owner];
    // one atomic per wave for its frequent keys
    const unsigned long long mask = __ballot(hit);
    const int lane = threadIdx.x & 63;
    uint32_t first = 0;
    const int leader = mask ? __ffsll((long long)mask) - 1 : 0;
    if (mask && lane == leader) first = atomicAdd(a.freq_count, (uint32_t)__popcll(mask));
    first = __shfl(first, leader, 64);
    if (hit) {
      const uint32_t slot = first + (uint32_t)__popcll(mask & ((1ull << lane) - 1));
      if (slot < a.freq_cap) {
        unsigned long long *f = a.freq + 3 * (int64_t)slot;
        f[0] = (unsigned long long)i;
        f[1] = (unsigned long long)s;
        f[2] = c;
      }
    }
  }
}

// ------------------------------------------------------------------ WRITE
// Child rows of the selected keys (sup[key] < 0 encodes child -(sup+1)),
// in three launches with no contended atomics:
//   k_count  per graph: children and the sum of their candidate bounds
//            (enumeration mode also keeps both per row);
//   k_scan   one workgroup: exclusive scans of both over the graphs give each
//            graph's first child row and its next-level list region;
//   k_write  per graph: the rows (pattern id, tuple) and each child's list
//            offset within its graph's region.
// Child rows are graph-major in graph order, so the layout is the same on
// every run.

struct WriteShared {
  long long red[5];          // block sums and scans (<= 4 waves + total)
  int ired[5];
  int pre[kMaxBlock];        // list mode: prefix of the chunk's list lengths
  long long off[kMaxBlock];  // list mode: list offset of each chunk row
  int pid[kMaxBlock];        // list mode: pattern of each chunk row
  int nvp[kMaxBlock];        // list mode (write): its vertex count
};

// Write one child row at `pos` (parent tuple in slot; root: vertex u).
template <bool ROOT>
__device__ __forceinline__ void put_child(const LevelArgs &a, int g, int64_t pos, int32_t child, int u, int w,
                                          int nvp, uint32_t *slot, bool keep_lists, uint32_t off) {
  a.cpid[pos] = child;
  a.crgraph[pos] = g;
  if (ROOT) {
    uint32_t *p = reinterpret_cast<uint32_t *>(a.cvt + pos * a.CS);
    p[0] = (uint32_t)u | ((uint32_t)w << 8);
    for (int i = 1; i < (a.CS >> 2); ++i) p[i] = 0;
  } else {
    if (w >= 0) slot_set(slot, nvp, w);
    store_slot(a.cvt, pos, a.CS, slot);
  }
  if (keep_lists) a.nrcoff[pos] = off;
}

// List mode: the graph's candidates are processed lane-parallel.  Rows are
// taken blockDim at a time; their list lengths are scanned into LDS and each
// lane finds the row of its flat candidate index by binary search.
template <bool ROOT, bool WRITE>
__device__ __forceinline__ void children_lists(const LevelArgs &a, int g, int64_t r0, int64_t r1, const Lds &l,
                                               WriteShared &sh) {
  const int B = blockDim.x;
  const int job = ROOT ? a.gjob[g] : 0;
  const int32_t vb = a.gvbase[g];
  const uint2 *lists = ROOT ? a.cand : a.cand + a.gcbase[g];
  uint32_t *slot = l.slots + threadIdx.x * a.slot_words;
  // rows [base, base + B): list prefix and offsets into sh; returns the
  // chunk's candidate total
  auto chunk = [&](int64_t base) -> int {
    const int64_t r = base + threadIdx.x;
    const bool live = r < r1;
    const int n = live ? (int)a.rccnt[ROOT ? vb + r : r] : 0;
    const int p = !ROOT && live ? a.pid[r] : 0;
    int T;
    const int pre = block_exclusive_scan<int>(n, T, sh.ired);
    sh.pre[threadIdx.x] = pre;
    sh.off[threadIdx.x] = live ? (ROOT ? (long long)a.vadj[vb + r] : (long long)a.rcoff[r]) : 0;
    sh.pid[threadIdx.x] = p;
    if (WRITE && !ROOT) sh.nvp[threadIdx.x] = live && n ? (int)a.meta[p].nv : 0;
    __syncthreads();
    return T;
  };
  auto find = [&](int j, int nrow) {
    int lo = 0, hi = nrow;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (sh.pre[mid] <= j) lo = mid;
      else hi = mid;
    }
    return lo;
  };
  if (!WRITE) {
    long long cnt = 0, bsum = 0;
    for (int64_t base = r0; base < r1; base += B) {
      const int T = chunk(base);
      const int nrow = (int)min<int64_t>(B, r1 - base);
      for (int j0 = threadIdx.x; j0 < T; j0 += kReplayUnroll * B) {
        uint2 c[kReplayUnroll];
        int kk[kReplayUnroll];
#pragma unroll
        for (int u = 0; u < kReplayUnroll; ++u) {
          const int j = j0 + u * B;
          if (j < T) {
            const int k = find(j, nrow);
            kk[u] = k;
            c[u] = lists[sh.off[k] + (j - sh.pre[k])];
          }
        }
        int32_t v[kReplayUnroll];
#pragma unroll
        for (int u = 0; u < kReplayUnroll; ++u) {
          v[u] = 0;
          if (j0 + u * B < T) {
            int key, w;
            cand_decode(ROOT, c[u].x, key, w);
            const uint64_t kbase = ROOT ? (uint64_t)job * a.K : (uint64_t)sh.pid[kk[u]] * (uint64_t)a.K;
            v[u] = a.sup[kbase + key];
          }
        }
#pragma unroll
        for (int u = 0; u < kReplayUnroll; ++u) {
          if (v[u] < 0) {
            ++cnt;
            bsum += c[u].y;
          }
        }
      }
      __syncthreads();
    }
    cnt = block_sum<long long>(cnt, sh.red);
    bsum = block_sum<long long>(bsum, sh.red);
    if (threadIdx.x == 0) {
      a.cgcount[g] = cnt;
      a.gbsum[g] = (unsigned long long)bsum;
    }
    return;
  }
  const long long nb = a.next_lists ? a.ngcbase[g] : -1;
  int64_t out = a.cgstart[g];
  long long boff = 0;
  for (int64_t base = r0; base < r1; base += B) {
    const int T = chunk(base);
    const int nrow = (int)min<int64_t>(B, r1 - base);
    for (int jb = 0; jb < T; jb += B) {
      const int j = jb + threadIdx.x;
      int32_t v = 0;
      int key = 0, w = 0, k = 0;
      uint32_t bound = 0;
      if (j < T) {
        k = find(j, nrow);
        const uint2 c = lists[sh.off[k] + (j - sh.pre[k])];
        cand_decode(ROOT, c.x, key, w);
        bound = c.y;
        const uint64_t kbase = ROOT ? (uint64_t)job * a.K : (uint64_t)sh.pid[k] * (uint64_t)a.K;
        v = a.sup[kbase + key];
      }
      const long long x = v < 0 ? ((1ll << kPackShift) | (long long)bound) : 0;
      long long tot;
      const long long pre = block_exclusive_scan<long long>(x, tot, sh.red);
      if (v < 0) {
        const int64_t r = base + k;
        int nvp = 0;
        if (!ROOT) {
          nvp = sh.nvp[k];
          load_slot(slot, a.vt, r, a.S);
        }
        put_child<ROOT>(a, g, out + (pre >> kPackShift), -v - 1, (int)r, w, nvp, slot, nb >= 0,
                        (uint32_t)(boff + (pre & kPackMask)));
      }
      out += tot >> kPackShift;
      boff += tot & kPackMask;
    }
    __syncthreads();
  }
}

// Enumeration mode: rows are enumerated again (adjacency staged in LDS or
// read from HBM), one row per lane.
template <bool ROOT, bool LDS, bool WRITE>
__device__ __forceinline__ void children_enum(const LevelArgs &a, int g, int64_t r0, int64_t r1, const Lds &l,
                                              WriteShared &sh) {
  const GView G = graph_view<LDS>(a, g, l);
  const int job = ROOT ? a.gjob[g] : 0;
  const int32_t vb = a.gvbase[g];
  uint32_t *slot = l.slots + threadIdx.x * a.slot_words;
  const int64_t cbase = ROOT ? (int64_t)vb : 0;  // rowcnt / rowbsum index base
  uint32_t *rmv = slot + a.tuple_words;
  // every candidate of row r (tuple already in the slot)
  auto candidates = [&](int64_t r, auto &&f) {
    if (ROOT) {
      enumerate_root(G, (int)r, a.vlab[vb + (int)r], a.NEL, a.NVL, f);
    } else {
      const PatMeta *pm = a.meta + a.pid[r];
      if (pm->flags & kPatDead) return;  // never extended: no candidate is selected
      stage_rm(pm, rmv);
      enumerate_row(G, pm, slot, rmv, a.S, a.NEL, a.NVL, f);
    }
  };
  if (!WRITE) {
    long long cnt = 0, bsum = 0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
      int n = 0;
      uint32_t b = 0;
      const uint64_t kbase = ROOT ? (uint64_t)job * a.K : (uint64_t)a.pid[r] * (uint64_t)a.K;
      if (!ROOT) load_slot(slot, a.vt, r, a.S);
      candidates(r, [&](int key, int, int bound) {
        if (a.sup[kbase + key] < 0) {
          ++n;
          b += (uint32_t)bound;
        }
      });
      a.rowcnt[cbase + r] = (uint16_t)n;
      a.rowbsum[cbase + r] = b;
      cnt += n;
      bsum += b;
    }
    cnt = block_sum<long long>(cnt, sh.red);
    bsum = block_sum<long long>(bsum, sh.red);
    if (threadIdx.x == 0) {
      a.cgcount[g] = cnt;
      a.gbsum[g] = (unsigned long long)bsum;
    }
    return;
  }
  const long long nb = a.next_lists ? a.ngcbase[g] : -1;
  int64_t out = a.cgstart[g];
  long long boff = 0;
  for (int64_t base = r0; base < r1; base += blockDim.x) {
    const int64_t r = base + threadIdx.x;
    const bool live = r < r1;
    const int n = live ? (int)a.rowcnt[cbase + r] : 0;
    const long long x = live ? (((long long)n << kPackShift) | (long long)a.rowbsum[cbase + r]) : 0;
    long long tot;
    const long long pre = block_exclusive_scan<long long>(x, tot, sh.red);
    if (n) {
      int64_t pos = out + (pre >> kPackShift);
      long long off = boff + (pre & kPackMask);
      const uint64_t kbase = ROOT ? (uint64_t)job * a.K : (uint64_t)a.pid[r] * (uint64_t)a.K;
      const int nvp = ROOT ? 0 : a.meta[a.pid[r]].nv;
      if (!ROOT) load_slot(slot, a.vt, r, a.S);
      candidates(r, [&](int key, int w, int bound) {
        const int32_t v = a.sup[kbase + key];
        if (v < 0) {
          put_child<ROOT>(a, g, pos, -v - 1, (int)r, w, nvp, slot, nb >= 0, (uint32_t)off);
          ++pos;
          off += bound;
        }
      });
    }
    out += tot >> kPackShift;
    boff += tot & kPackMask;
  }
}

// WRITE = false: k_count; WRITE = true: k_write.
template <bool ROOT, bool WRITE, bool BIG>
__global__ void __launch_bounds__(kMaxBlock) k_children(LevelArgs a) {
  __shared__ WriteShared sh;
  const int g = pick_graph<BIG>(a);
  if (g < 0) return;
  int64_t r0, r1;
  graph_rows<ROOT>(a, g, r0, r1);
  if (!ROOT && !BIG && r1 - r0 > a.big_rows) return;
  if (r0 == r1) {
    if (!WRITE && threadIdx.x == 0) {
      a.cgcount[g] = 0;
      a.gbsum[g] = 0;
    }
    return;
  }
  if (WRITE && a.cgcount[g] == 0) return;
  const Lds l = lds_layout(a, false);
  const bool lists = a.use_lists && (ROOT || a.gcbase[g] >= 0);
  if (lists)
    children_lists<ROOT, WRITE>(a, g, r0, r1, l, sh);
  else if (graph_fits(a, g))
    children_enum<ROOT, true, WRITE>(a, g, r0, r1, l, sh);
  else
    children_enum<ROOT, false, WRITE>(a, g, r0, r1, l, sh);
}

// Placement of the children, three launches over tiles of kScanTile graphs
// (one graph per thread): k_scan_tiles sums each tile's child rows, list
// entries (gbsum) and big graphs; k_scan_top (one workgroup) turns the tile
// sums into exclusive tile offsets and writes the totals (*cursor = rows,
// *ncand_cursor = list entries, *nnbig = big graphs of the next level,
// cursor[2] = the largest graph's rows); k_scan_apply writes per graph
// cgstart = exclusive scan of cgcount, ngcbase = exclusive scan of gbsum
// (-1 where the region would pass the store or the 32-bit offsets) and the
// next level's big-graph list in graph order.  A graph whose rows would pass
// ccap gets no rows (the host compares the total with the embedding count
// and fails the call).
constexpr int kScanTile = kScanTileGraphs;

__device__ __forceinline__ void tile_values(const LevelArgs &a, int g, int n, long long &c, unsigned long long &b,
                                            int &big) {
  c = g < n ? a.cgcount[g] : 0;
  b = g < n ? a.gbsum[g] : 0;
  big = c > a.big_rows ? 1 : 0;
}

__global__ void __launch_bounds__(kScanTile) k_scan_tiles(LevelArgs a, int n) {
  __shared__ long long s[kScanTile / 64];
  __shared__ unsigned long long su[kScanTile / 64];
  __shared__ int si[kScanTile / 64];
  const int g = blockIdx.x * kScanTile + threadIdx.x;
  long long c;
  unsigned long long b;
  int big;
  tile_values(a, g, n, c, b, big);
  long long m = c;
  for (int o = 32; o > 0; o >>= 1) m = max(m, (long long)__shfl_down(m, o, 64));
  const long long rows = block_sum<long long>(c, s);
  const unsigned long long ents = block_sum<unsigned long long>(b, su);
  const int bigs = block_sum<int>(big, si);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kScanTile / 64; ++i) m = max(m, s[i]);
    a.tile_rows[blockIdx.x] = rows;
    a.tile_ents[blockIdx.x] = ents;
    a.tile_big[blockIdx.x] = bigs;
    a.tile_max[blockIdx.x] = m;
  }
}

// Zero the counters the next level step's support stage accumulates (frequent
// key count, error flag, split passes, spread statistics).  The level's own
// values were downloaded before its write pass began.
__device__ __forceinline__ void clear_level_counters(const LevelArgs &a) {
  if (threadIdx.x == 0) {
    *a.freq_count = 0;
    *a.error_flag = 0;
    *a.split_passes = 0;
  }
  for (int i = threadIdx.x; i < kSpread * kSpreadStride; i += blockDim.x) a.spread[i] = 0;
}

__global__ void __launch_bounds__(256) k_clear_counters(LevelArgs a) { clear_level_counters(a); }

__global__ void __launch_bounds__(kScanTile) k_scan_top(LevelArgs a, int ntiles) {
  __shared__ long long s[kScanTile / 64 + 1];
  __shared__ unsigned long long su[kScanTile / 64 + 1];
  __shared__ int si[kScanTile / 64 + 1];
  long long crow = 0, cmax = 0;
  unsigned long long cent = 0;
  int cbig = 0;
  for (int base = 0; base < ntiles; base += kScanTile) {
    const int t = base + threadIdx.x;
    const bool live = t < ntiles;
    long long tr, tm = live ? a.tile_max[t] : 0;
    unsigned long long te;
    int tb;
    const long long r = block_exclusive_scan<long long>(live ? a.tile_rows[t] : 0, tr, s);
    const unsigned long long e = block_exclusive_scan<unsigned long long>(live ? a.tile_ents[t] : 0, te, su);
    const int bg = block_exclusive_scan<int>(live ? (int)a.tile_big[t] : 0, tb, si);
    if (live) {
      a.tile_rows[t] = crow + r;
      a.tile_ents[t] = cent + e;
      a.tile_big[t] = cbig + bg;
    }
    crow += tr;
    cent += te;
    cbig += tb;
    for (int o = 32; o > 0; o >>= 1) tm = max(tm, (long long)__shfl_down(tm, o, 64));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = tm;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = 0; i < kScanTile / 64; ++i) cmax = max(cmax, s[i]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.cursor[0] = (unsigned long long)crow;
    *a.ncand_cursor = cent;
    a.cursor[2] = (unsigned long long)cmax;  // largest graph's child rows
    *a.nnbig = (unsigned long long)cbig;
  }
  clear_level_counters(a);
}

// The three scan kernels in one workgroup, for up to kScanTile *
// kScanSinglePer graphs: thread t places the consecutive graphs
// [t * per, (t + 1) * per).
constexpr int kScanSinglePer = kScanSingleGraphs / kScanTile;

__global__ void __launch_bounds__(kScanTile) k_scan_single(LevelArgs a, int n) {
  __shared__ long long s[kScanTile / 64 + 1];
  __shared__ unsigned long long su[kScanTile / 64 + 1];
  __shared__ int si[kScanTile / 64 + 1];
  const int per = (n + kScanTile - 1) / kScanTile;  // <= kScanSinglePer (the launcher checks)
  const int g0 = min(n, (int)threadIdx.x * per), g1 = min(n, g0 + per);
  // all of the thread's loads first (one memory latency, not per graph)
  long long cv[kScanSinglePer];
  unsigned long long bv[kScanSinglePer];
#pragma unroll
  for (int k = 0; k < kScanSinglePer; ++k) {
    const int g = g0 + k;
    const bool in = g < g1;
    cv[k] = in ? a.cgcount[g] : 0;
    bv[k] = in ? a.gbsum[g] : 0;
  }
  long long csum = 0, cmax = 0;
  unsigned long long bsum = 0;
  int bigs = 0;
#pragma unroll
  for (int k = 0; k < kScanSinglePer; ++k) {
    csum += cv[k];
    bsum += bv[k];
    bigs += cv[k] > a.big_rows ? 1 : 0;
    cmax = max(cmax, cv[k]);
  }
  long long tr;
  unsigned long long te;
  int tb;
  long long rows = block_exclusive_scan<long long>(csum, tr, s);
  unsigned long long ents = block_exclusive_scan<unsigned long long>(bsum, te, su);
  int at = block_exclusive_scan<int>(bigs, tb, si);
#pragma unroll
  for (int k = 0; k < kScanSinglePer; ++k) {
    const int g = g0 + k;
    if (g >= g1) break;
    const long long c = cv[k];
    const unsigned long long b = bv[k];
    a.cgstart[g] = rows;
    if (rows + c > a.ccap) a.cgcount[g] = 0;  // past the allocation: not written; the host sees the total
    long long nb = -1;
    if (a.next_lists && b > 0 && b <= 0xFFFFFFFFull && ents + b <= (unsigned long long)a.ncand_cap) nb = (long long)ents;
    a.ngcbase[g] = nb;
    if (c > a.big_rows) a.nbig_list[at++] = g;
    rows += c;
    ents += b;
  }
  for (int o = 32; o > 0; o >>= 1) cmax = max(cmax, (long long)__shfl_down(cmax, o, 64));
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = cmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kScanTile / 64; ++i) cmax = max(cmax, s[i]);
    a.cursor[0] = (unsigned long long)tr;
    *a.ncand_cursor = te;
    a.cursor[2] = (unsigned long long)cmax;  // largest graph's child rows
    *a.nnbig = (unsigned long long)tb;
  }
  clear_level_counters(a);
}

__global__ void __launch_bounds__(kScanTile) k_scan_apply(LevelArgs a, int n) {
  __shared__ long long s[kScanTile / 64 + 1];
  __shared__ unsigned long long su[kScanTile / 64 + 1];
  __shared__ int si[kScanTile / 64 + 1];
  const int g = blockIdx.x * kScanTile + threadIdx.x;
  long long c;
  unsigned long long b;
  int big;
  tile_values(a, g, n, c, b, big);
  long long tr;
  unsigned long long te;
  int tb;
  const long long rows = a.tile_rows[blockIdx.x] + block_exclusive_scan<long long>(c, tr, s);
  const unsigned long long ents = a.tile_ents[blockIdx.x] + block_exclusive_scan<unsigned long long>(b, te, su);
  const int at = (int)a.tile_big[blockIdx.x] + block_exclusive_scan<int>(big, tb, si);
  if (g >= n) return;
  a.cgstart[g] = rows;
  if (rows + c > a.ccap) a.cgcount[g] = 0;  // past the allocation: not written; the host sees the total
  long long nb = -1;
  if (a.next_lists && b > 0 && b <= 0xFFFFFFFFull && ents + b <= (unsigned long long)a.ncand_cap) nb = (long long)ents;
  a.ngcbase[g] = nb;
  if (big) a.nbig_list[at] = g;
}

// --------------------------------------------------------- ROW-PARALLEL WRITE
// Later levels whose candidates k_expand listed (every graph with
// candidates then has lists): the children are counted, placed and written
// over rows instead of graphs -- one single-wave workgroup per 64
// consecutive rows whatever their graphs, so a graph of thousands of rows
// spreads over many waves rather than running serially in one workgroup,
// and no workgroup barrier is needed.
// (The root level runs the same passes with the database's vertices as
// rows: their lists sit at their adjacency offsets.)
//   k_rows<false>  per row: selected candidates and their bound sum (rowcnt,
//                  rowbsum); per wave: the totals (wrows, wents), also added
//                  to its tile's sums (trows, tents: 64 waves a tile);
//   k_rtop         one workgroup: exclusive scans of the tile sums, in
//                  place; the grand totals go to the cursors;
//   k_rows<true>   per row: its children at the wave's base (its tile's
//                  offset plus the tile's earlier waves) plus the row's
//                  prefix, in list order, their list offsets absolute (the
//                  next level's per-graph list base is 0, -1 without lists);
//                  a graph's first and last row record its child-row and
//                  list-entry ranges (cgstart, gbsum; cgcount, ngcbase);
//   k_gfix         per graph: child count, list flag, next big-graph list,
//                  largest graph.
// Child rows come out graph-major in graph order, as from the per-graph
// passes.
constexpr int kRowWave = kRowWaveRows;
constexpr int kRowUnroll = 4;  // item windows whose loads are issued together

template <bool ROOT, bool WRITE>
__global__ void __launch_bounds__(kRowWave) k_rows(LevelArgs a) {
  // root level: the rows are the database's vertices (graph from vgraph,
  // list at the vertex's adjacency offset, keys per job, child tuple (u, w))
  __shared__ int s_pre[kRowWave];
  __shared__ uint32_t s_li[kRowWave];
  __shared__ int32_t s_p[kRowWave];
  __shared__ int32_t s_g[kRowWave];
  __shared__ int32_t s_nvp[kRowWave];
  __shared__ long long s_pos[kRowWave];
  __shared__ unsigned long long s_ent[kRowWave];
  __shared__ int32_t s_ms[kRowWave];  // count by support: the row's job's min support
  extern __shared__ uint32_t tup[];  // WRITE: row t's tuple at tup[t * TW, + TW)
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * kRowWave;
  const int64_t r = r0 + t;
  const int nrow = (int)min<int64_t>(kRowWave, a.n_rows - r0);
  const int TW = a.S >> 2;
  int n = 0, g = -1;
  int32_t p = 0;
  uint32_t li = 0;
  int uv = 0;  // root level: the vertex's index in its graph
  if (t < nrow) {
    if (ROOT) {
      g = a.vgraph[r];
      p = a.gjob[g];
      n = a.rccnt[r];
      li = (uint32_t)a.vadj[r];
      uv = (int)(r - a.gvbase[g]);
    } else {
      g = a.rgraph[r];
      p = a.pid[r];
      const long long gb = a.gcbase[g];
      if (gb >= 0) {  // graphs without lists have no candidates
        n = a.rccnt[r];
        li = (uint32_t)(gb + a.rcoff[r]);
      }
      if (WRITE) load_slot(tup + t * TW, a.vt, r, a.S);
    }
  }
  int T;
  const int pre = wave_exclusive_scan<int>(n, T);
  s_pre[t] = pre;
  s_li[t] = li;
  s_p[t] = p;
  // a speculative level's count runs before the child ids are scattered:
  // every key reaching its job's min support becomes a child
  const bool by_support = !WRITE && a.count_by_support;
  if (by_support) s_ms[t] = n ? a.minsup_job[ROOT ? p : a.meta[p].job] : 0;
  if (WRITE) {
    // the row's first child row and first list entry
    uint32_t rc = 0, rb = 0;
    int nvp = 0;
    if (t < nrow) {
      rc = a.rowcnt[r];
      rb = a.rowbsum[r];
      if (ROOT)
        nvp = uv;  // the child's tuple starts with the vertex
      else if (rc)
        nvp = a.meta[p].nv;
    }
    // the wave's base: its tile's offset plus the tile's earlier waves
    const int64_t w0 = (int64_t)blockIdx.x & ~(int64_t)63;
    const bool earlier = t < (int)(blockIdx.x & 63);
    const int64_t tile = (int64_t)(blockIdx.x >> 6) * kTileStride;
    const unsigned long long br = a.trows[tile] + wave_sum<unsigned long long>(earlier ? a.wrows[w0 + t] : 0);
    const unsigned long long be = a.tents[tile] + wave_sum<unsigned long long>(earlier ? a.wents[w0 + t] : 0);
    unsigned long long tc, tb;
    const long long pos0 = (long long)br + (long long)wave_exclusive_scan<unsigned long long>(rc, tc);
    const unsigned long long ent0 = be + wave_exclusive_scan<unsigned long long>(rb, tb);
    s_g[t] = g;
    s_nvp[t] = nvp;
    s_pos[t] = pos0;
    s_ent[t] = ent0;
    // graph boundaries (rows are graph-major): the neighbours' graphs
    int gprev = __shfl(g, t > 0 ? t - 1 : 0, kRowWave);
    int gnext = __shfl(g, t < kRowWave - 1 ? t + 1 : t, kRowWave);
    const int32_t *rg = ROOT ? a.vgraph : a.rgraph;
    if (t == 0) gprev = r > 0 ? rg[r - 1] : -1;
    if (t == nrow - 1) gnext = r + 1 < a.n_rows ? rg[r + 1] : -1;
    if (t < nrow) {
      if (gprev != g) {
        a.cgstart[g] = pos0;
        a.gbsum[g] = ent0;
      }
      if (gnext != g) {  // ends; k_gfix turns them into the count and the list flag
        a.cgcount[g] = pos0 + (long long)rc;
        a.ngcbase[g] = (long long)(ent0 + rb);
      }
    }
  }
  wave_sync();
  const int CW = a.CS >> 2;
  uint32_t cnt = 0, bsum = 0;  // row t's selected candidates so far, their bound sum
  for (int jb0 = 0; jb0 < T; jb0 += kRowUnroll * kRowWave) {
    int kk[kRowUnroll];
    uint2 c[kRowUnroll];
    int32_t v[kRowUnroll];
#pragma unroll
    for (int u = 0; u < kRowUnroll; ++u) {
      const int j = jb0 + u * kRowWave + t;
      kk[u] = 0;
      c[u] = make_uint2(0, 0);
      if (j < T) {
        const int k = find_row(s_pre, j, nrow);
        kk[u] = k;
        c[u] = a.cand[(uint64_t)s_li[k] + (uint64_t)(j - s_pre[k])];
      }
    }
#pragma unroll
    for (int u = 0; u < kRowUnroll; ++u) {
      v[u] = 0;
      if (jb0 + u * kRowWave + t < T) {
        int key, w;
        cand_decode(ROOT, c[u].x, key, w);
        v[u] = a.sup[(uint64_t)s_p[kk[u]] * (uint64_t)a.K + (uint64_t)key];
      }
    }
#pragma unroll
    for (int u = 0; u < kRowUnroll; ++u) {
      const int jb = jb0 + u * kRowWave;
      const bool live = jb + t < T;
      const bool sel = live && (by_support ? v[u] >= s_ms[kk[u]] : v[u] < 0);
      const unsigned long long mask = __ballot(sel);
      const uint32_t x = sel ? c[u].y : 0u;
      uint32_t xt;
      const uint32_t ex = wave_exclusive_scan<uint32_t>(x, xt);
      const int k = kk[u];
      if (WRITE) {
        const int first = live ? max(s_pre[k], jb) - jb : 0;  // lane of row k's first item here
        const uint32_t kc = __shfl(cnt, k, kRowWave), kb = __shfl(bsum, k, kRowWave);
        const uint32_t exf = __shfl(ex, first, kRowWave);
        if (sel) {
          const unsigned long long below = mask & ((1ull << t) - 1) & ~((1ull << first) - 1);
          const long long pos = s_pos[k] + (long long)kc + (long long)__popcll(below);
          if (pos < a.ccap) {
            int key, w;
            cand_decode(ROOT, c[u].x, key, w);
            a.cpid[pos] = -v[u] - 1;
            a.crgraph[pos] = s_g[k];
            uint32_t *dst = reinterpret_cast<uint32_t *>(a.cvt + pos * a.CS);
            const int nvp = s_nvp[k];
            if (ROOT) {  // (u, w)
              dst[0] = (uint32_t)nvp | ((uint32_t)w << 8);
              for (int i = 1; i < CW; ++i) dst[i] = 0;
            } else {
              const uint32_t *src = tup + k * TW;
              for (int i = 0; i < CW; ++i) {
                uint32_t word = i < TW ? src[i] : 0u;
                if (w >= 0 && (nvp >> 2) == i) {
                  const int sh = (nvp & 3) * 8;
                  word = (word & ~(0xFFu << sh)) | ((uint32_t)w << sh);
                }
                dst[i] = word;
              }
            }
            if (a.next_lists) a.nrcoff[pos] = (uint32_t)(s_ent[k] + kb + (ex - exf));
          }
        }
      }
      // row t's items in this window: lanes [lo, hi)
      const int lo = max(pre, jb) - jb, hi = min(pre + n, jb + kRowWave) - jb;
      const uint32_t hi_inc = __shfl(ex + x, hi > lo ? hi - 1 : 0, kRowWave);
      const uint32_t lo_ex = __shfl(ex, hi > lo ? lo : 0, kRowWave);
      if (hi > lo) {
        const unsigned long long range = (hi - lo == 64 ? ~0ull : (((1ull << (hi - lo)) - 1) << lo));
        cnt += (uint32_t)__popcll(mask & range);
        bsum += hi_inc - lo_ex;
      }
    }
  }
  if (!WRITE) {
    if (t < nrow) {
      a.rowcnt[r] = (uint16_t)cnt;
      a.rowbsum[r] = bsum;
    }
    const unsigned long long wr = wave_sum<unsigned long long>(cnt), we = wave_sum<unsigned long long>(bsum);
    if (t == 0) {
      a.wrows[blockIdx.x] = wr;
      a.wents[blockIdx.x] = we;
      if (wr) atomicAdd(&a.trows[(int64_t)(blockIdx.x >> 6) * kTileStride], wr);
      if (we) atomicAdd(&a.tents[(int64_t)(blockIdx.x >> 6) * kTileStride], we);
    }
  }
}

// Exclusive scans of the tile sums in place (n tiles of 64 waves); the
// totals to the cursors, the next level's big-graph count and largest graph
// zeroed for k_gfix.
__global__ void __launch_bounds__(kScanTile) k_rtop(LevelArgs a, int n) {
  __shared__ long long s[kScanTile / 64 + 1];
  __shared__ unsigned long long su[kScanTile / 64 + 1];
  long long crow = 0;
  unsigned long long cent = 0;
  for (int base = 0; base < n; base += kScanTile) {
    const int i = base + threadIdx.x;
    const bool live = i < n;
    long long tr;
    unsigned long long te;
    const int64_t at = (int64_t)i * kTileStride;
    const long long r = block_exclusive_scan<long long>(live ? (long long)a.trows[at] : 0, tr, s);
    const unsigned long long e = block_exclusive_scan<unsigned long long>(live ? a.tents[at] : 0, te, su);
    if (live) {
      a.trows[at] = (unsigned long long)(crow + r);
      a.tents[at] = cent + e;
    }
    crow += tr;
    cent += te;
  }
  if (threadIdx.x == 0) {
    a.cursor[0] = (unsigned long long)crow;
    *a.ncand_cursor = cent;
    a.cursor[2] = 0;
    *a.nnbig = 0;
  }
  clear_level_counters(a);
}

// Per graph after k_rows<true>: the child count from the recorded range
// (graphs without rows at this level: none), the next level's list flag
// (0: lists at absolute offsets, -1: none), the big-graph list (wave-
// aggregated append; its order does not matter) and the largest graph.
template <bool ROOT>
__global__ void __launch_bounds__(256) k_gfix(LevelArgs a, int n_graphs) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  long long c = 0;
  bool lists = false;
  if (g < n_graphs) {
    const bool rows = ROOT ? a.gvbase[g + 1] > a.gvbase[g] : a.gcount[g] > 0;
    if (rows) {
      const long long s0 = a.cgstart[g], s1 = a.cgcount[g];
      const unsigned long long e0 = a.gbsum[g], e1 = (unsigned long long)a.ngcbase[g];
      c = s1 > a.ccap ? 0 : s1 - s0;  // past the allocation: not written; the host sees the total
      const unsigned long long b = e1 - e0;
      lists = a.next_lists && b > 0 && e1 <= 0xFFFFFFFFull && e1 <= (unsigned long long)a.ncand_cap;
    } else {
      a.cgstart[g] = 0;
    }
    a.cgcount[g] = c;
    a.ngcbase[g] = lists ? 0 : -1;
  }
  const bool big = c > a.big_rows;
  const unsigned long long bm = __ballot(big);
  const int lane = threadIdx.x & 63;
  uint32_t at = 0;
  const int leader = bm ? __ffsll((long long)bm) - 1 : 0;
  if (bm && lane == leader) at = (uint32_t)atomicAdd(a.nnbig, (unsigned long long)__popcll(bm));
  at = __shfl(at, leader, 64);
  if (big) a.nbig_list[at + (uint32_t)__popcll(bm & ((1ull << lane) - 1))] = g;
  long long m = c;
  for (int o = 32; o > 0; o >>= 1) m = max(m, (long long)__shfl_xor(m, o, 64));
  if (lane == 0 && m > 0) atomicMax(&a.cursor[2], (unsigned long long)m);
}

// ------------------------------------------------------------ primitives
// Copies between device memory and mapped pinned host memory on the compute
// queue (a copy engine transfer between two kernels costs a cross-engine
// wait; the level step's transfers are small).
__global__ void k_copy8(unsigned long long *dst, const unsigned long long *src, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// The children's upload (copied to device memory) and sup[key] = child id
// for the selected keys (read from the upload's source).
__global__ void k_upload_scatter(unsigned long long *dst, const unsigned long long *src, int64_t n, int32_t *sup,
                                 const uint64_t *keys, const int32_t *vals, int64_t C) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t i = t; i < n; i += stride) dst[i] = src[i];
  for (int64_t i = t; i < C; i += stride) sup[keys[i]] = vals[i];
}



}  // namespace

// ------------------------------------------------------------ launchers
static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

size_t level_lds_bytes(const LevelArgs &a, bool support) {
  return 4ull * ((support ? (2u << a.table_bits) : 0u) + (size_t)a.block * a.slot_words + a.stage_v + a.stage_a);
}

hipError_t launch_support(const LevelArgs &a, int n_graphs, bool root, hipStream_t s) {
  if (n_graphs <= 0) return hipSuccess;
  if (root) {
    // per-wave kernel: enough waves to fill the device, never more than graphs
    const unsigned wave_grid = (unsigned)std::min<int64_t>((n_graphs + 3) / 4, a.wave_grid);
    hipLaunchKernelGGL(k_root_wave, dim3(wave_grid), dim3(kWaveBlock), 0, s, a, n_graphs);
  } else {
    hipLaunchKernelGGL((k_support<false, false>), dim3(n_graphs), dim3(a.block), level_lds_bytes(a, true), s, a);
    if (a.big_grid > 0) {
      LevelArgs b = a;
      b.block = a.big_block;
      b.table_bits = a.big_table_bits;
      hipLaunchKernelGGL((k_support<false, true>), dim3(a.big_grid), dim3(b.block), level_lds_bytes(b, true), s, b);
    }
  }
  return hipGetLastError();
}

size_t expand_lds_bytes(const LevelArgs &a) { return 4ull * kExpandRows * a.slot_words; }

hipError_t launch_expand(const LevelArgs &a, hipStream_t s) {
  if (a.n_rows <= 0) return hipSuccess;
  const unsigned grid = (unsigned)((a.n_rows + kExpandRows - 1) / kExpandRows);
  hipLaunchKernelGGL(k_expand, dim3(grid), dim3(kExpandRows), expand_lds_bytes(a), s, a);
  return hipGetLastError();
}

hipError_t launch_reduce(const LevelArgs &a, bool root, hipStream_t s) {
  if (a.nkeys <= 0) return hipSuccess;
  const unsigned g = (unsigned)std::min<int64_t>((a.nkeys + 255) / 256, 16384);
  if (root)
    hipLaunchKernelGGL(k_reduce<true>, dim3(g), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_reduce<false>, dim3(g), dim3(256), 0, s, a);
  return hipGetLastError();
}

static unsigned copy_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 2048)); }

// One word to mapped host memory, released at system scope after every
// earlier kernel of the stream: the host polls it instead of synchronising
// the stream, so work queued after it keeps running.
__global__ void k_signal(unsigned long long *flag, unsigned long long v) {
  __hip_atomic_store(flag, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

hipError_t launch_signal(unsigned long long *flag, unsigned long long v, hipStream_t s) {
  hipLaunchKernelGGL(k_signal, dim3(1), dim3(1), 0, s, flag, v);
  return hipGetLastError();
}

hipError_t launch_count(const LevelArgs &a, bool root, hipStream_t s) {
  const int64_t nw = (a.n_rows + kRowWave - 1) / kRowWave;
  if (nw <= 0) return hipSuccess;
  if ((nw + 63) / 64 > a.n_wtiles) return hipErrorInvalidValue;  // tile sums not sized (or not zeroed) for nw
  if (root)
    hipLaunchKernelGGL((k_rows<true, false>), dim3((unsigned)nw), dim3(kRowWave), 0, s, a);
  else
    hipLaunchKernelGGL((k_rows<false, false>), dim3((unsigned)nw), dim3(kRowWave), 0, s, a);
  // the placement scan needs only the count's tile sums
  hipLaunchKernelGGL(k_rtop, dim3(1), dim3(kScanTile), 0, s, a, (int)((nw + 63) / 64));
  return hipGetLastError();
}

hipError_t launch_copy8(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes % 8) return hipErrorInvalidValue;
  const int64_t n = (int64_t)(bytes / 8);
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy8, dim3(copy_grid(n)), dim3(256), 0, s, static_cast<unsigned long long *>(dst),
                     static_cast<const unsigned long long *>(src), n);
  return hipGetLastError();
}

hipError_t launch_upload_scatter(void *dst, const void *src, size_t bytes, int32_t *sup, const uint64_t *keys,
                                 const int32_t *vals, int64_t C, hipStream_t s) {
  if (bytes % 8) return hipErrorInvalidValue;
  const int64_t n = (int64_t)(bytes / 8);
  if (n == 0 && C == 0) return hipSuccess;
  hipLaunchKernelGGL(k_upload_scatter, dim3(copy_grid(std::max(n, C))), dim3(256), 0, s,
                     static_cast<unsigned long long *>(dst), static_cast<const unsigned long long *>(src), n, sup,
                     keys, vals, C);
  return hipGetLastError();
}

hipError_t launch_clear_counters(const LevelArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_write(const LevelArgs &a, int n_graphs, bool root, hipStream_t s) {
  if (n_graphs <= 0) return hipSuccess;
  if (a.rows_mode) {
    const int64_t nw = (a.n_rows + kRowWave - 1) / kRowWave;
    if (nw <= 0) return hipSuccess;
    if ((nw + 63) / 64 > a.n_wtiles) return hipErrorInvalidValue;  // tile sums not sized (or not zeroed) for nw
    const size_t tup_lds = root ? 0 : sizeof(uint32_t) * kRowWave * (a.S >> 2);
    if (!a.count_done) {  // count pass and placement scan
      const hipError_t e = launch_count(a, root, s);
      if (e != hipSuccess) return e;
    }
    if (root) {
      hipLaunchKernelGGL((k_rows<true, true>), dim3((unsigned)nw), dim3(kRowWave), tup_lds, s, a);
      hipLaunchKernelGGL(k_gfix<true>, dim3((unsigned)((n_graphs + 255) / 256)), dim3(256), 0, s, a, n_graphs);
    } else {
      hipLaunchKernelGGL((k_rows<false, true>), dim3((unsigned)nw), dim3(kRowWave), tup_lds, s, a);
      hipLaunchKernelGGL(k_gfix<false>, dim3((unsigned)((n_graphs + 255) / 256)), dim3(256), 0, s, a, n_graphs);
    }
    return hipGetLastError();
  }
  const size_t lds = level_lds_bytes(a, false);
  LevelArgs b = a;
  b.block = a.big_block;
  const size_t lds_big = level_lds_bytes(b, false);
  const bool big = !root && a.big_grid > 0;
  if (root) {
    hipLaunchKernelGGL((k_children<true, false, false>), dim3(n_graphs), dim3(a.block), lds, s, a);
  } else {
    hipLaunchKernelGGL((k_children<false, false, false>), dim3(n_graphs), dim3(a.block), lds, s, a);
    if (big) hipLaunchKernelGGL((k_children<false, false, true>), dim3(a.big_grid), dim3(b.block), lds_big, s, b);
  }
  if (n_graphs <= kScanSingleGraphs) {
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kScanTile), 0, s, a, n_graphs);
  } else {
    const unsigned tiles = (unsigned)((n_graphs + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanTile), 0, s, a, n_graphs);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanTile), 0, s, a, (int)tiles);
    hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(kScanTile), 0, s, a, n_graphs);
  }
  if (root) {
    hipLaunchKernelGGL((k_children<true, true, false>), dim3(n_graphs), dim3(a.block), lds, s, a);
  } else {
    hipLaunchKernelGGL((k_children<false, true, false>), dim3(n_graphs), dim3(a.block), lds, s, a);
    if (big) hipLaunchKernelGGL((k_children<false, true, true>), dim3(a.big_grid), dim3(b.block), lds_big, s, b);
  }
  return hipGetLastError();
}



}  // namespace gm
