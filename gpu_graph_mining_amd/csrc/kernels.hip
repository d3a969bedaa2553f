// kernels.hip -- HIP kernels of the level step (gfx950, wave64).
//
// One workgroup owns one database graph for the whole level step, so every
// support question ("how many distinct graphs have an embedding of P with
// extension x") is answered by deduplicating (pattern, key) pairs inside the
// workgroup (an LDS hash table) and adding 1 per distinct pair to a global
// counter.  The same enumeration routine drives three kernels:
//   k_support          distinct-graph support of every (pattern, key)
//   k_children<false>  child rows per graph for the selected keys (count)
//   k_children<true>   the child rows themselves (pattern id + tuple)
// and, for the first level, the same three over 1-edge keys.
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

constexpr int kStageV = 257;   // vertex offsets staged in LDS (<= 256 vertices)
constexpr int kStageA = 2048;  // adjacency entries staged in LDS
constexpr int kSlotWords = 9;  // per-thread tuple slot: 36 bytes (odd word stride)
constexpr uint32_t kEmpty32 = 0xFFFFFFFFu;

// Graph adjacency: entry offsets vadj[v] are absolute; entries A[a - aoff].
struct GView {
  const int32_t *vadj;
  const uint32_t *A;
  int32_t aoff;
  __device__ __forceinline__ int32_t begin(int v) const { return vadj[v]; }
  __device__ __forceinline__ int32_t end(int v) const { return vadj[v + 1]; }
  __device__ __forceinline__ uint32_t entry(int32_t a) const { return A[a - aoff]; }
};

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

// Enumerate every valid extension of the embedding in `slot`.
// emit(key, w): extension key (see KeySpace) and new vertex (-1 backward).
template <class F>
__device__ __forceinline__ void enumerate_row(const GView &G, const PatMeta *__restrict__ pm, const uint32_t *slot,
                                              int S, int NEL, int NVL, F &&emit) {
  const int nv = pm->nv;
  const int maxtoc = nv - 1;
  const int minlabel = pm->minlabel;
  const int maxlabel = pm->maxlabel;
  const int nrm = pm->nrm;
  const uint32_t bwd_mask = pm->bwd_mask;
  const int fwd_base = S * NEL;
  VSet vis;
  for (int i = 0; i < nv; ++i) vis.add(slot_get(slot, i));
  const int rv = slot_get(slot, maxtoc);

  // backward and pure forward: one sweep over the rightmost vertex's edges
  for (int32_t a = G.begin(rv), ae = G.end(rv); a < ae; ++a) {
    const uint32_t e = G.entry(a);
    const int w = e & 0xFF, el = (e >> 8) & 0xFF, nl = (e >> 16) & 0xFF;
    if (!vis.has(w)) {
      if (nl >= minlabel) emit(fwd_base + (maxtoc * NEL + el) * NVL + nl, w);
    } else {
      for (int i = 1; i < nrm; ++i) {
        const uint32_t rme = pm->rm[i];
        const int f = rme & 0xFF;
        if (slot_get(slot, f) != w) continue;
        const int e1 = (rme >> 8) & 0xFF, l1 = (rme >> 16) & 0xFF;
        if (!((bwd_mask >> f) & 1u) && (e1 < el || (e1 == el && l1 <= maxlabel))) emit(f * NEL + el, -1);
        break;
      }
    }
  }
  // forward from the rightmost path
  for (int i = 0; i < nrm; ++i) {
    const uint32_t rme = pm->rm[i];
    const int f = rme & 0xFF;
    const int u = slot_get(slot, f);
    const int e1 = (rme >> 8) & 0xFF, l1 = (rme >> 16) & 0xFF;
    for (int32_t a = G.begin(u), ae = G.end(u); a < ae; ++a) {
      const uint32_t e = G.entry(a);
      const int w = e & 0xFF, el = (e >> 8) & 0xFF, nl = (e >> 16) & 0xFF;
      if (vis.has(w) || nl < minlabel) continue;
      if (e1 < el || (e1 == el && l1 <= nl)) emit(fwd_base + (f * NEL + el) * NVL + nl, w);
    }
  }
}

// Root extensions of vertex u (1-edge codes with fromlabel <= tolabel).
template <class F>
__device__ __forceinline__ void enumerate_root(const GView &G, int u, int lu, int NEL, int NVL, F &&emit) {
  for (int32_t a = G.begin(u), ae = G.end(u); a < ae; ++a) {
    const uint32_t e = G.entry(a);
    const int w = e & 0xFF, el = (e >> 8) & 0xFF, nl = (e >> 16) & 0xFF;
    if (lu <= nl) emit((lu * NEL + el) * NVL + nl, w);
  }
}

struct Stage {
  int32_t vadj[kStageV];
  uint32_t adj[kStageA];
};

// Whether graph g's adjacency fits the LDS stage (uniform per workgroup).
__device__ __forceinline__ bool graph_fits(const LevelArgs &a, int g) {
  const int32_t vb = a.gvbase[g], ve = a.gvbase[g + 1];
  return ve - vb + 1 <= kStageV && a.vadj[ve] - a.vadj[vb] <= kStageA;
}

template <bool LDS>
__device__ __forceinline__ GView graph_view(const LevelArgs &a, int g, Stage &st) {
  const int32_t vb = a.gvbase[g], ve = a.gvbase[g + 1];
  const int32_t ab = a.vadj[vb], ae = a.vadj[ve];
  GView G;
  if (LDS) {
    for (int i = threadIdx.x; i <= ve - vb; i += blockDim.x) st.vadj[i] = a.vadj[vb + i];
    for (int i = threadIdx.x; i < ae - ab; i += blockDim.x) st.adj[i] = a.adj[ab + i];
    __syncthreads();
    G.vadj = st.vadj;
    G.A = st.adj;
    G.aoff = ab;
  } else {
    G.vadj = a.vadj + vb;
    G.A = a.adj;
    G.aoff = 0;
  }
  return G;
}

// Block-wide exclusive scan of one int per thread (wave64 shuffles + LDS).
__device__ __forceinline__ int block_exclusive_scan(int x, int &total, int *s_waves) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  int v = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  if (lane == 63) s_waves[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int i = 0; i < nw; ++i) {
      int t = s_waves[i];
      s_waves[i] = acc;
      acc += t;
    }
    s_waves[nw] = acc;
  }
  __syncthreads();
  int r = s_waves[wid] + v - x;
  total = s_waves[nw];
  __syncthreads();
  return r;
}

__device__ __forceinline__ uint32_t hash_slot32(uint32_t k) { return (k * 0x9E3779B1u) >> (32 - kTableBits); }

// ---------------------------------------------------------------- SUPPORT
// One pass per residue class of pattern ids (p mod `mod` == `res`; root
// level: of the source vertex label): every candidate key
// (p - p_lo) * K + key is inserted into an LDS hash table; when the table
// passes 3/4 full the pass is abandoned and the class split in two (all keys
// of one pattern stay in one class, so the per-graph dedup stays exact);
// a completed pass adds 1 per distinct key to the global supports.
template <bool ROOT, bool LDS>
__device__ __forceinline__ void support_body(const LevelArgs &a, int g, int64_t r0, int64_t r1, Stage &st,
                                             uint32_t *table, uint32_t *slot, int *s_misc,
                                             uint32_t *s_stack_mod, uint32_t *s_stack_res) {
  const GView G = graph_view<LDS>(a, g, st);
  const int job = ROOT ? a.gjob[g] : 0;
  const int32_t vb = a.gvbase[g];
  int &s_fill = s_misc[0], &s_over = s_misc[1], &s_top = s_misc[2];
  unsigned long long my_cand = 0;
  if (threadIdx.x == 0) {
    s_top = 1;
    s_stack_mod[0] = 1;
    s_stack_res[0] = 0;
  }
  for (;;) {
    __syncthreads();
    if (s_top == 0) break;
    const uint32_t mod = s_stack_mod[s_top - 1], res = s_stack_res[s_top - 1];
    for (int i = threadIdx.x; i < kTableCap; i += blockDim.x) table[i] = kEmpty32;
    if (threadIdx.x == 0) {
      s_fill = 0;
      s_over = 0;
    }
    __syncthreads();
    unsigned long long cand = 0;
    int over = 0;  // this thread saw the table pass its fill limit
    auto insert = [&](uint32_t k) {
      ++cand;
      uint32_t h = hash_slot32(k);
      for (int probe = 0; probe < kTableCap; ++probe) {
        const uint32_t old = atomicCAS(&table[h], kEmpty32, k);
        if (old == kEmpty32) {
          if (atomicAdd(&s_fill, 1) >= kTableFill) over = 1;
          return;
        }
        if (old == k) return;
        h = (h + 1) & (kTableCap - 1);
      }
      over = 1;
    };
    for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
      if (ROOT) {
        const int u = (int)r;
        const int lu = a.vlab[vb + u];
        if (mod > 1 && ((uint32_t)lu & (mod - 1)) != res) continue;
        const uint32_t base = (uint32_t)((int64_t)(job - a.p_lo) * a.K);
        enumerate_root(G, u, lu, a.NEL, a.NVL, [&](int key, int) { insert(base + (uint32_t)key); });
      } else {
        const int32_t p = a.pid[r];
        if (p < a.p_lo || p >= a.p_hi) continue;
        if (mod > 1 && ((uint32_t)p & (mod - 1)) != res) continue;
        load_slot(slot, a.vt, r, a.S);
        const uint32_t base = (uint32_t)((int64_t)(p - a.p_lo) * a.K);
        enumerate_row(G, a.meta + p, slot, a.S, a.NEL, a.NVL, [&](int key, int) { insert(base + (uint32_t)key); });
      }
    }
    if (over) atomicOr(&s_over, 1);
    __syncthreads();
    const bool failed = s_over != 0;
    __syncthreads();
    if (failed) {
      if (threadIdx.x == 0) {
        if (mod >= (1u << 30) || s_top + 1 > 32) {
          atomicOr(a.error_flag, 1);
          s_top = 0;
        } else {
          s_stack_mod[s_top - 1] = mod * 2;
          s_stack_res[s_top - 1] = res + mod;
          s_stack_mod[s_top] = mod * 2;
          s_stack_res[s_top] = res;
          s_top += 1;
        }
      }
      continue;
    }
    my_cand += cand;
    for (int i = threadIdx.x; i < kTableCap; i += blockDim.x) {
      const uint32_t k = table[i];
      if (k == kEmpty32) continue;
      const int64_t gk = (int64_t)a.p_lo * a.K + (int64_t)k;
      const int32_t old = atomicAdd(&a.sup[gk], 1);
      const int64_t owner_p = gk / a.K;
      const int32_t owner = ROOT ? (int32_t)owner_p : a.meta[owner_p].job;
      if (old + 1 == a.minsup_job[owner]) {
        const uint32_t slot_i = atomicAdd(a.freq_count, 1u);
        if (slot_i < a.freq_cap) a.freq[slot_i] = (unsigned long long)gk;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_top -= 1;
  }
  // candidate counter (E_cand): one atomic per wave
  for (int o = 32; o > 0; o >>= 1) my_cand += __shfl_down(my_cand, o, 64);
  if ((threadIdx.x & 63) == 0 && my_cand) atomicAdd(a.ecand, my_cand);
}

template <bool ROOT>
__global__ void __launch_bounds__(kBlock) k_support(LevelArgs a) {
  __shared__ Stage st;
  __shared__ uint32_t table[kTableCap];
  __shared__ uint32_t slots[kBlock * kSlotWords];
  __shared__ int s_misc[4];
  __shared__ uint32_t s_stack_mod[32], s_stack_res[32];
  const int g = blockIdx.x;
  int64_t r0 = 0, r1;
  if (ROOT) {
    const int job = a.gjob[g];
    if (job < a.p_lo || job >= a.p_hi) return;
    r1 = a.gvbase[g + 1] - a.gvbase[g];
  } else {
    r0 = a.grow[g];
    r1 = a.grow[g + 1];
  }
  if (r0 == r1) return;
  uint32_t *slot = slots + threadIdx.x * kSlotWords;
  if (graph_fits(a, g))
    support_body<ROOT, true>(a, g, r0, r1, st, table, slot, s_misc, s_stack_mod, s_stack_res);
  else
    support_body<ROOT, false>(a, g, r0, r1, st, table, slot, s_misc, s_stack_mod, s_stack_res);
}

// ------------------------------------------------------------ COUNT/WRITE
template <bool ROOT, bool WRITE, bool LDS>
__device__ __forceinline__ void children_body(const LevelArgs &a, int g, int64_t r0, int64_t r1, Stage &st,
                                              uint32_t *slot, int *s_waves, long long *s_red) {
  const GView G = graph_view<LDS>(a, g, st);
  const int job = ROOT ? a.gjob[g] : 0;
  const int32_t vb = a.gvbase[g];
  const int64_t nrows = r1 - r0;
  int64_t out = WRITE ? a.cgrow[g] : 0;
  long long total_count = 0;
  for (int64_t base = 0; base < nrows; base += blockDim.x) {
    const int64_t r = r0 + base + threadIdx.x;
    const bool live = base + threadIdx.x < nrows;
    int cnt = 0;
    const PatMeta *pm = a.meta;
    uint64_t kbase = 0;
    int lu = 0;
    if (live) {
      if (ROOT) {
        lu = a.vlab[vb + (int)r];
        kbase = (uint64_t)job * a.K;
        enumerate_root(G, (int)r, lu, a.NEL, a.NVL, [&](int key, int) {
          if (a.sup[kbase + key] < 0) ++cnt;
        });
      } else {
        const int32_t p = a.pid[r];
        pm = a.meta + p;
        load_slot(slot, a.vt, r, a.S);
        kbase = (uint64_t)p * (uint64_t)a.K;
        enumerate_row(G, pm, slot, a.S, a.NEL, a.NVL, [&](int key, int) {
          if (a.sup[kbase + key] < 0) ++cnt;
        });
      }
    }
    if (!WRITE) {
      total_count += cnt;
      continue;
    }
    int tot;
    const int pre = block_exclusive_scan(cnt, tot, s_waves);
    if (live && cnt) {
      int64_t pos = out + pre;
      if (ROOT) {
        const int u = (int)r;
        enumerate_root(G, u, lu, a.NEL, a.NVL, [&](int key, int w) {
          const int32_t v = a.sup[kbase + key];
          if (v < 0) {
            a.cpid[pos] = -v - 1;
            uint32_t *p = reinterpret_cast<uint32_t *>(a.cvt + pos * a.CS);
            p[0] = (uint32_t)u | ((uint32_t)w << 8);
            for (int i = 1; i < (a.CS >> 2); ++i) p[i] = 0;
            ++pos;
          }
        });
      } else {
        const int nvp = pm->nv;
        enumerate_row(G, pm, slot, a.S, a.NEL, a.NVL, [&](int key, int w) {
          const int32_t v = a.sup[kbase + key];
          if (v < 0) {
            a.cpid[pos] = -v - 1;
            if (w >= 0) slot_set(slot, nvp, w);
            store_slot(a.cvt, pos, a.CS, slot);
            ++pos;
          }
        });
      }
    }
    out += tot;
  }
  if (!WRITE) {
    long long v = total_count;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      long long s = 0;
      for (int i = 0; i < (int)((blockDim.x + 63) >> 6); ++i) s += s_red[i];
      a.ccount[g] = s;
    }
  }
}

template <bool ROOT, bool WRITE>
__global__ void __launch_bounds__(kBlock) k_children(LevelArgs a) {
  __shared__ Stage st;
  __shared__ uint32_t slots[kBlock * kSlotWords];
  __shared__ int s_waves[17];
  __shared__ long long s_red[16];
  const int g = blockIdx.x;
  int64_t r0 = 0, r1;
  if (ROOT)
    r1 = a.gvbase[g + 1] - a.gvbase[g];
  else {
    r0 = a.grow[g];
    r1 = a.grow[g + 1];
  }
  if (r0 == r1) {
    if (!WRITE && threadIdx.x == 0) a.ccount[g] = 0;
    return;
  }
  uint32_t *slot = slots + threadIdx.x * kSlotWords;
  if (graph_fits(a, g))
    children_body<ROOT, WRITE, true>(a, g, r0, r1, st, slot, s_waves, s_red);
  else
    children_body<ROOT, WRITE, false>(a, g, r0, r1, st, slot, s_waves, s_red);
}

// ------------------------------------------------------------ primitives
// Exclusive scan of n int64 values (n+1 outputs; out[n] = total), three
// phases: per-block sums, one-block scan of the sums, per-block rescan.
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ long long block_excl_scan_ll(long long x, long long &total, long long *s) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  long long v = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    long long y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  if (lane == 63) s[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long acc = 0;
    for (int i = 0; i < kScanBlock / 64; ++i) {
      long long t = s[i];
      s[i] = acc;
      acc += t;
    }
    s[kScanBlock / 64] = acc;
  }
  __syncthreads();
  long long r = s[wid] + v - x;
  total = s[kScanBlock / 64];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_sums(const int64_t *in, int64_t n, int64_t *sums) {
  __shared__ long long s[kScanBlock / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  long long acc = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) acc += in[base + i];
  long long tot;
  block_excl_scan_ll(acc, tot, s);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_block_sums(int64_t *sums, int64_t nb) {
  __shared__ long long s[kScanBlock / 64 + 1];
  long long carry = 0;
  for (int64_t base = 0; base < nb; base += kScanBlock) {
    const int64_t i = base + threadIdx.x;
    long long x = i < nb ? sums[i] : 0;
    long long tot;
    long long ex = block_excl_scan_ll(x, tot, s);
    if (i < nb) sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_apply(const int64_t *in, int64_t n, const int64_t *sums,
                                                           int64_t *out) {
  __shared__ long long s[kScanBlock / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  long long vals[kScanItems];
  long long acc = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    vals[i] = base + i < n ? in[base + i] : 0;
    acc += vals[i];
  }
  long long tot;
  long long ex = block_excl_scan_ll(acc, tot, s) + sums[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = ex;
    ex += vals[i];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = sums[gridDim.x];
}

__global__ void k_gather_i32(const int32_t *src, const uint64_t *idx, int64_t n, int32_t *dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

__global__ void k_scatter_i32(int32_t *dst, const uint64_t *idx, const int32_t *val, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = val[i];
}

__global__ void k_fill_i32(int32_t *dst, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = v;
}

__global__ void k_pattern_hist(const int32_t *pid, int64_t n, unsigned long long *hist) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&hist[pid[i]], 1ull);
}

}  // namespace

// ------------------------------------------------------------ launchers
static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_support(const LevelArgs &a, int n_graphs, bool root, hipStream_t s) {
  if (n_graphs <= 0) return hipSuccess;
  if (root)
    hipLaunchKernelGGL(k_support<true>, dim3(n_graphs), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_support<false>, dim3(n_graphs), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_children(const LevelArgs &a, int n_graphs, bool root, bool write, hipStream_t s) {
  if (n_graphs <= 0) return hipSuccess;
  if (root) {
    if (write)
      hipLaunchKernelGGL((k_children<true, true>), dim3(n_graphs), dim3(kBlock), 0, s, a);
    else
      hipLaunchKernelGGL((k_children<true, false>), dim3(n_graphs), dim3(kBlock), 0, s, a);
  } else {
    if (write)
      hipLaunchKernelGGL((k_children<false, true>), dim3(n_graphs), dim3(kBlock), 0, s, a);
    else
      hipLaunchKernelGGL((k_children<false, false>), dim3(n_graphs), dim3(kBlock), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_exclusive_scan(const int64_t *in, int64_t n, int64_t *out, int64_t *tmp, hipStream_t s) {
  // tmp must hold num_tiles + 1 values
  const int64_t tiles = std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)tiles), dim3(kScanBlock), 0, s, in, n, tmp);
  hipLaunchKernelGGL(k_scan_block_sums, dim3(1), dim3(kScanBlock), 0, s, tmp, tiles);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)tiles), dim3(kScanBlock), 0, s, in, n, tmp, out);
  return hipGetLastError();
}

int64_t scan_tmp_elems(int64_t n) { return std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile) + 1; }

hipError_t launch_gather_i32(const int32_t *src, const uint64_t *idx, int64_t n, int32_t *dst, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_i32, dim3(grid1(n, 256)), dim3(256), 0, s, src, idx, n, dst);
  return hipGetLastError();
}

hipError_t launch_scatter_i32(int32_t *dst, const uint64_t *idx, const int32_t *val, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter_i32, dim3(grid1(n, 256)), dim3(256), 0, s, dst, idx, val, n);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t *dst, int64_t n, int32_t v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned g = (unsigned)std::min<int64_t>(grid1(n, 256), 8192);
  hipLaunchKernelGGL(k_fill_i32, dim3(g), dim3(256), 0, s, dst, n, v);
  return hipGetLastError();
}

hipError_t launch_pattern_hist(const int32_t *pid, int64_t n, unsigned long long *hist, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned g = (unsigned)std::min<int64_t>(grid1(n, 256), 8192);
  hipLaunchKernelGGL(k_pattern_hist, dim3(g), dim3(256), 0, s, pid, n, hist);
  return hipGetLastError();
}

}  // namespace gm
