// kernels.hip -- HIP kernels of the level step (gfx950, wave64).
//
// One workgroup owns one database graph for the whole level step, so every
// support question ("how many distinct graphs have an embedding of P with
// extension x") is answered by deduplicating (pattern, key) pairs inside the
// workgroup (an LDS hash table) and adding 1 per distinct pair to a global
// counter.  The same enumeration routine drives two kernels:
//   k_support  distinct-graph support and embedding count of every
//              (pattern, key)
//   k_write    the child rows (pattern id + tuple) of the selected keys
// and, for the first level, the same two over 1-edge keys.
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

// Dynamic LDS layout of the level kernels (bytes):
//   [0, 8*2^tb)                dedup table: keys, then candidate counts
//                              (k_support only)
//   then kBlock * 36           per-thread tuple slots
//   then 4*stage_v             staged vertex offsets
//   then 4*stage_a             staged adjacency entries
struct Lds {
  uint32_t *table;
  uint32_t *slots;
  int32_t *vadj;
  uint32_t *adj;
};

__device__ __forceinline__ Lds lds_layout(const LevelArgs &a, bool with_table) {
  extern __shared__ uint32_t dyn[];
  Lds l;
  uint32_t *p = dyn;
  l.table = p;
  if (with_table) p += (2u << a.table_bits);
  l.slots = p;
  p += blockDim.x * kSlotWords;
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

__device__ __forceinline__ uint32_t hash_slot32(uint32_t k, int bits) { return (k * 0x9E3779B1u) >> (32 - bits); }

// ---------------------------------------------------------------- SUPPORT
// One pass per residue class of pattern ids (p mod `mod` == `res`; root
// level: of the source vertex label): every candidate key
// (p - p_lo) * K + key is inserted into the LDS hash table, whose second
// half counts the candidates per key; when the table passes 3/4 full the
// pass is abandoned and the class split in two (all keys of one pattern stay
// in one class, so the per-graph dedup stays exact).  A completed pass adds
// 1 per distinct key to the distinct-graph counters and the key's candidate
// count to the embedding counters of the graph's replica
// (replica = workgroup id mod a.rep; k_reduce sums the replicas).
template <bool ROOT, bool LDS>
__device__ __forceinline__ void support_body(const LevelArgs &a, int g, int64_t r0, int64_t r1, const Lds &l,
                                             int *s_misc, uint32_t *s_stack_mod, uint32_t *s_stack_res) {
  const GView G = graph_view<LDS>(a, g, l);
  const int job = ROOT ? a.gjob[g] : 0;
  const int32_t vb = a.gvbase[g];
  const int tb = a.table_bits;
  const uint32_t cap = 1u << tb, fill_limit = cap - cap / 4;
  uint32_t *table = l.table;
  uint32_t *tcnt = l.table + cap;
  uint32_t *slot = l.slots + threadIdx.x * kSlotWords;
  const int64_t rbase = (int64_t)(g & (a.rep - 1)) * a.nkeys;
  int32_t *rep = a.sup_rep + rbase;
  unsigned long long *repc = a.proj_rep + rbase;
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
    for (uint32_t i = threadIdx.x; i < cap; i += blockDim.x) {
      table[i] = kEmpty32;
      tcnt[i] = 0;
    }
    if (threadIdx.x == 0) {
      s_fill = 0;
      s_over = 0;
    }
    __syncthreads();
    unsigned long long cand = 0;
    int over = 0;  // this thread saw the table pass its fill limit
    auto insert = [&](uint32_t k) {
      ++cand;
      uint32_t h = hash_slot32(k, tb);
      for (uint32_t probe = 0; probe < cap; ++probe) {
        const uint32_t old = atomicCAS(&table[h], kEmpty32, k);
        if (old == kEmpty32 || old == k) {
          atomicAdd(&tcnt[h], 1u);
          if (old == kEmpty32 && (uint32_t)atomicAdd(&s_fill, 1) >= fill_limit) over = 1;
          return;
        }
        h = (h + 1) & (cap - 1);
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
        atomicAdd(a.split_passes, 1ull);
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
    for (uint32_t i = threadIdx.x; i < cap; i += blockDim.x) {
      const uint32_t k = table[i];
      if (k == kEmpty32) continue;
      const int64_t gk = (int64_t)a.p_lo * a.K + (int64_t)k;
      atomicAdd(&rep[gk], 1);
      atomicAdd(&repc[gk], (unsigned long long)tcnt[i]);
    }
    __syncthreads();
    if (threadIdx.x == 0) s_top -= 1;
  }
  // candidate counter (E_cand): one atomic per wave
  for (int o = 32; o > 0; o >>= 1) my_cand += __shfl_down(my_cand, o, 64);
  if ((threadIdx.x & 63) == 0 && my_cand) atomicAdd(a.ecand, my_cand);
}

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

template <bool ROOT>
__global__ void __launch_bounds__(kMaxBlock) k_support(LevelArgs a) {
  __shared__ int s_misc[4];
  __shared__ uint32_t s_stack_mod[32], s_stack_res[32];
  const int g = blockIdx.x;
  if (ROOT) {
    const int job = a.gjob[g];
    if (job < a.p_lo || job >= a.p_hi) return;
  }
  int64_t r0, r1;
  graph_rows<ROOT>(a, g, r0, r1);
  if (r0 == r1) return;
  const Lds l = lds_layout(a, true);
  if (graph_fits(a, g))
    support_body<ROOT, true>(a, g, r0, r1, l, s_misc, s_stack_mod, s_stack_res);
  else
    support_body<ROOT, false>(a, g, r0, r1, l, s_misc, s_stack_mod, s_stack_res);
}

// Sum the replicas of every key; keys reaching their job's min support go to
// the frequent list as (key, support, embeddings).
template <bool ROOT>
__global__ void __launch_bounds__(256) k_reduce(LevelArgs a) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.nkeys; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t s = 0;
    unsigned long long c = 0;
    for (int r = 0; r < a.rep; ++r) {
      s += a.sup_rep[(int64_t)r * a.nkeys + i];
      c += a.proj_rep[(int64_t)r * a.nkeys + i];
    }
    a.sup[i] = s;
    const int64_t p = i / a.K;
    const int32_t owner = ROOT ? (int32_t)p : a.meta[p].job;
    if (s >= a.minsup_job[owner]) {
      const uint32_t slot = atomicAdd(a.freq_count, 1u);
      if (slot < a.freq_cap) {
        a.freq[3 * slot] = (unsigned long long)i;
        a.freq[3 * slot + 1] = (unsigned long long)s;
        a.freq[3 * slot + 2] = c;
      }
    }
  }
}

// ------------------------------------------------------------------ WRITE
// Child rows of graph g for the selected keys (sup[key] < 0 encodes child
// -(sup+1)): pass 1 counts each row's children into rowcnt, one atomic on
// the level cursor reserves the graph's contiguous output segment, pass 2
// scans the counts and writes (pattern id, tuple) rows.
template <bool ROOT, bool LDS>
__device__ __forceinline__ void write_body(const LevelArgs &a, int g, int64_t r0, int64_t r1, const Lds &l,
                                           int *s_waves, long long *s_red) {
  const GView G = graph_view<LDS>(a, g, l);
  const int job = ROOT ? a.gjob[g] : 0;
  const int32_t vb = a.gvbase[g];
  uint32_t *slot = l.slots + threadIdx.x * kSlotWords;
  const int64_t cbase = ROOT ? (int64_t)vb : 0;  // rowcnt index base
  // pass 1: count
  long long mine = 0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    int cnt = 0;
    if (ROOT) {
      const int lu = a.vlab[vb + (int)r];
      const uint64_t kbase = (uint64_t)job * a.K;
      enumerate_root(G, (int)r, lu, a.NEL, a.NVL, [&](int key, int) {
        if (a.sup[kbase + key] < 0) ++cnt;
      });
    } else {
      const int32_t p = a.pid[r];
      load_slot(slot, a.vt, r, a.S);
      const uint64_t kbase = (uint64_t)p * (uint64_t)a.K;
      enumerate_row(G, a.meta + p, slot, a.S, a.NEL, a.NVL, [&](int key, int) {
        if (a.sup[kbase + key] < 0) ++cnt;
      });
    }
    a.rowcnt[cbase + r] = (uint16_t)cnt;
    mine += cnt;
  }
  for (int o = 32; o > 0; o >>= 1) mine += __shfl_down(mine, o, 64);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long total = 0;
    for (int i = 0; i < (int)((blockDim.x + 63) >> 6); ++i) total += s_red[i];
    long long start = 0;
    if (total > 0) {
      start = (long long)atomicAdd(a.cursor, (unsigned long long)total);
      if (start + total > a.ccap) {
        atomicOr(a.error_flag, 2);
        total = 0;  // nothing is written; the host reports the overflow
      }
    }
    a.cgstart[g] = start;
    a.cgcount[g] = total;
    s_red[0] = start;
    s_red[1] = total;
  }
  __syncthreads();
  const long long start = s_red[0];
  if (s_red[1] == 0) return;
  // pass 2: write
  int64_t out = start;
  for (int64_t base = r0; base < r1; base += blockDim.x) {
    const int64_t r = base + threadIdx.x;
    const bool live = r < r1;
    const int cnt = live ? (int)a.rowcnt[cbase + r] : 0;
    int tot;
    const int pre = block_exclusive_scan(cnt, tot, s_waves);
    if (live && cnt) {
      int64_t pos = out + pre;
      if (ROOT) {
        const int u = (int)r;
        const int lu = a.vlab[vb + u];
        const uint64_t kbase = (uint64_t)job * a.K;
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
        const int32_t p = a.pid[r];
        const PatMeta *pm = a.meta + p;
        const int nvp = pm->nv;
        load_slot(slot, a.vt, r, a.S);
        const uint64_t kbase = (uint64_t)p * (uint64_t)a.K;
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
}

template <bool ROOT>
__global__ void __launch_bounds__(kMaxBlock) k_write(LevelArgs a) {
  __shared__ int s_waves[17];
  __shared__ long long s_red[16];
  const int g = blockIdx.x;
  int64_t r0, r1;
  graph_rows<ROOT>(a, g, r0, r1);
  if (r0 == r1) {
    if (threadIdx.x == 0) {
      a.cgstart[g] = 0;
      a.cgcount[g] = 0;
    }
    return;
  }
  const Lds l = lds_layout(a, false);
  if (graph_fits(a, g))
    write_body<ROOT, true>(a, g, r0, r1, l, s_waves, s_red);
  else
    write_body<ROOT, false>(a, g, r0, r1, l, s_waves, s_red);
}

// ------------------------------------------------------------ primitives
__global__ void k_scatter_i32(int32_t *dst, const uint64_t *idx, const int32_t *val, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = val[i];
}

__global__ void k_fill_i32(int32_t *dst, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = v;
}

}  // namespace

// ------------------------------------------------------------ launchers
static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

size_t level_lds_bytes(const LevelArgs &a, bool with_table) {
  return 4ull * ((with_table ? (2u << a.table_bits) : 0u) + (size_t)a.block * kSlotWords + a.stage_v + a.stage_a);
}

hipError_t launch_support(const LevelArgs &a, int n_graphs, bool root, hipStream_t s) {
  if (n_graphs <= 0) return hipSuccess;
  const size_t lds = level_lds_bytes(a, true);
  if (root)
    hipLaunchKernelGGL(k_support<true>, dim3(n_graphs), dim3(a.block), lds, s, a);
  else
    hipLaunchKernelGGL(k_support<false>, dim3(n_graphs), dim3(a.block), lds, s, a);
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

hipError_t launch_write(const LevelArgs &a, int n_graphs, bool root, hipStream_t s) {
  if (n_graphs <= 0) return hipSuccess;
  const size_t lds = level_lds_bytes(a, false);
  if (root)
    hipLaunchKernelGGL(k_write<true>, dim3(n_graphs), dim3(a.block), lds, s, a);
  else
    hipLaunchKernelGGL(k_write<false>, dim3(n_graphs), dim3(a.block), lds, s, a);
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

}  // namespace gm
