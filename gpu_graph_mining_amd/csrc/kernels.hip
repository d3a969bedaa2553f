// kernels.hip -- HIP kernels of the level step (gfx950, wave64).
//
// One workgroup owns one database graph for the whole level step, so every
// support question ("how many distinct graphs have an embedding of P with
// extension x") is answered by deduplicating (pattern, key) pairs inside the
// workgroup (an LDS hash table) and adding 1 per distinct pair to a global
// counter.  The same enumeration routine drives three kernels:
//   SUPPORT  count distinct graphs per (pattern, key)         [kernel "A"]
//   COUNT    count child rows per graph for selected keys      [kernel "B1"]
//   WRITE    write child rows (pattern id + vertex tuple)      [kernel "B2"]
// and, for the first level, the same three over 1-edge keys.
//
// Extension rules (pattern side precomputed in PatMeta, graph side read
// from the packed adjacency):
//   backward  rightmost vertex -> image of rightmost-path vertex rm_from[i],
//             i >= 1, edge not yet in the pattern, and
//             rm_el[i] < e || (rm_el[i] == e && rm_tl[i] <= label(rightmost))
//             (src/globals/graph_types.cpp:389-413, gspan.cpp:220-225);
//   pure fwd  rightmost vertex -> unvisited w, label(w) >= minlabel
//             (graph_types.cpp:337-355, gspan.cpp:233-237);
//   rm fwd    rm_from[i] -> unvisited w, label(w) >= minlabel and
//             rm_el[i] < e || (rm_el[i] == e && rm_tl[i] <= label(w))
//             (graph_types.cpp:305-323, gspan.cpp:240-246);
//   root      u -> w with label(u) <= label(w) (graph_types.cpp:365-375).
// Because the database graphs are simple, "edge not in the embedding" is a
// pattern-level fact (bwd_mask) and "vertex not in the embedding" is a
// bitmask over the graph's local ids built from the tuple.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace gm {

namespace {

struct Tuple {
  uint64_t w0, w1, w2, w3;
};

__device__ __forceinline__ int tget(const Tuple &t, int i) {
  uint64_t w = (i < 16) ? ((i < 8) ? t.w0 : t.w1) : ((i < 24) ? t.w2 : t.w3);
  return (int)((w >> ((i & 7) * 8)) & 0xFF);
}

__device__ __forceinline__ void tset(Tuple &t, int i, int v) {
  int q = i >> 3, sh = (i & 7) * 8;
  uint64_t clr = ~(0xFFull << sh), val = (uint64_t)(v & 0xFF) << sh;
  t.w0 = q == 0 ? ((t.w0 & clr) | val) : t.w0;
  t.w1 = q == 1 ? ((t.w1 & clr) | val) : t.w1;
  t.w2 = q == 2 ? ((t.w2 & clr) | val) : t.w2;
  t.w3 = q == 3 ? ((t.w3 & clr) | val) : t.w3;
}

struct VMask {
  uint64_t m0, m1, m2, m3;
};

__device__ __forceinline__ void vset(VMask &m, int v) {
  uint64_t b = 1ull << (v & 63);
  int q = v >> 6;
  m.m0 |= q == 0 ? b : 0;
  m.m1 |= q == 1 ? b : 0;
  m.m2 |= q == 2 ? b : 0;
  m.m3 |= q == 3 ? b : 0;
}

__device__ __forceinline__ bool vhas(const VMask &m, int v) {
  uint64_t w = (v < 128) ? ((v < 64) ? m.m0 : m.m1) : ((v < 192) ? m.m2 : m.m3);
  return (w >> (v & 63)) & 1;
}

__device__ __forceinline__ Tuple load_tuple(const uint8_t *vt, int64_t r, int S) {
  const uint32_t *p = reinterpret_cast<const uint32_t *>(vt + r * S);
  const int nw = S >> 2;
  uint32_t w[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) w[i] = i < nw ? p[i] : 0u;
  Tuple t;
  t.w0 = w[0] | ((uint64_t)w[1] << 32);
  t.w1 = w[2] | ((uint64_t)w[3] << 32);
  t.w2 = w[4] | ((uint64_t)w[5] << 32);
  t.w3 = w[6] | ((uint64_t)w[7] << 32);
  return t;
}

__device__ __forceinline__ void store_tuple(uint8_t *vt, int64_t r, int S, const Tuple &t) {
  uint32_t *p = reinterpret_cast<uint32_t *>(vt + r * S);
  const int nw = S >> 2;
  uint32_t w[8] = {(uint32_t)t.w0, (uint32_t)(t.w0 >> 32), (uint32_t)t.w1, (uint32_t)(t.w1 >> 32),
                   (uint32_t)t.w2, (uint32_t)(t.w2 >> 32), (uint32_t)t.w3, (uint32_t)(t.w3 >> 32)};
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (i < nw) p[i] = w[i];
}

// Graph adjacency view: vadj[] indexed by local vertex (absolute entry
// offsets), entries A[a - aoff].  Either LDS-staged or straight from HBM.
struct GView {
  const int32_t *vadj;
  const uint32_t *A;
  int64_t aoff;
  __device__ __forceinline__ int32_t begin(int v) const { return vadj[v]; }
  __device__ __forceinline__ int32_t end(int v) const { return vadj[v + 1]; }
  __device__ __forceinline__ uint32_t entry(int32_t a) const { return A[a - aoff]; }
};

// Enumerate every valid extension of one embedding.  emit(key, w) gets the
// extension key (see KeySpace) and the new vertex's local id (-1 backward).
template <class F>
__device__ __forceinline__ void enumerate_row(const GView &G, const PatMeta *__restrict__ pm, const Tuple &t,
                                              int S, int NEL, int NVL, F &&emit) {
  const int nv = pm->nv;
  const int maxtoc = nv - 1;
  const int minlabel = pm->minlabel;
  const int maxlabel = pm->maxlabel;
  const int nrm = pm->nrm;
  const uint32_t bwd_mask = pm->bwd_mask;
  const int fwd_base = S * NEL;
  VMask vis{0, 0, 0, 0};
  for (int i = 0; i < nv; ++i) vset(vis, tget(t, i));
  const int rv = tget(t, maxtoc);

  // backward and pure forward: one sweep over the rightmost vertex's edges
  for (int32_t a = G.begin(rv), ae = G.end(rv); a < ae; ++a) {
    const uint32_t e = G.entry(a);
    const int w = e & 0xFF, el = (e >> 8) & 0xFF, nl = (e >> 16) & 0xFF;
    if (!vhas(vis, w)) {
      if (nl >= minlabel) emit(fwd_base + (maxtoc * NEL + el) * NVL + nl, w);
    } else {
      for (int i = 1; i < nrm; ++i) {
        const uint32_t rme = pm->rm[i];
        const int f = rme & 0xFF;
        if (tget(t, f) != w) continue;
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
    const int u = tget(t, f);
    const int e1 = (rme >> 8) & 0xFF, l1 = (rme >> 16) & 0xFF;
    for (int32_t a = G.begin(u), ae = G.end(u); a < ae; ++a) {
      const uint32_t e = G.entry(a);
      const int w = e & 0xFF, el = (e >> 8) & 0xFF, nl = (e >> 16) & 0xFF;
      if (vhas(vis, w) || nl < minlabel) continue;
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

constexpr int kStageV = 257;
constexpr int kStageA = 2048;
constexpr uint64_t kEmpty = ~0ull;

struct Shared {
  int32_t vadj[kStageV];
  uint32_t adj[kStageA];
};

// Stage graph g's adjacency in LDS when it fits; returns the view to use.
__device__ __forceinline__ GView stage_graph(const LevelArgs &a, int g, Shared &sh, int &nv_out) {
  const int32_t vb = a.gvbase[g], ve = a.gvbase[g + 1];
  const int nv = ve - vb;
  nv_out = nv;
  const int32_t ab = a.vadj[vb], ae = a.vadj[ve];
  GView G;
  if (nv + 1 <= kStageV && ae - ab <= kStageA) {
    for (int i = threadIdx.x; i <= nv; i += blockDim.x) sh.vadj[i] = a.vadj[vb + i];
    for (int i = threadIdx.x; i < ae - ab; i += blockDim.x) sh.adj[i] = a.adj[ab + i];
    G.vadj = sh.vadj;
    G.A = sh.adj;
    G.aoff = ab;
  } else {
    G.vadj = a.vadj + vb;
    G.A = a.adj;
    G.aoff = 0;
  }
  __syncthreads();
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

__device__ __forceinline__ uint32_t hash_slot(uint64_t k, int bits) {
  return (uint32_t)((k * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

// ---------------------------------------------------------------- SUPPORT
template <bool ROOT>
__global__ void __launch_bounds__(256) k_support(LevelArgs a) {
  __shared__ Shared sh;
  __shared__ unsigned long long table[kHashCap];
  __shared__ int s_fill, s_over, s_top;
  __shared__ uint32_t s_stack_mod[32], s_stack_res[32];
  const int g = blockIdx.x;
  int nv;
  int64_t r0, r1;
  if (ROOT) {
    r0 = 0;
    r1 = 0;  // rows are the graph's vertices
  } else {
    r0 = a.grow[g];
    r1 = a.grow[g + 1];
    if (r0 == r1) return;
  }
  GView G = stage_graph(a, g, sh, nv);
  if (ROOT && nv == 0) return;
  const int job = ROOT ? a.gjob[g] : 0;
  const int32_t vb = a.gvbase[g];
  unsigned long long my_cand = 0;
  if (threadIdx.x == 0) {
    s_top = 1;
    s_stack_mod[0] = 1;
    s_stack_res[0] = 0;
  }
  __syncthreads();
  while (true) {
    __syncthreads();
    if (s_top == 0) break;
    const uint32_t mod = s_stack_mod[s_top - 1], res = s_stack_res[s_top - 1];
    for (int i = threadIdx.x; i < kHashCap; i += blockDim.x) table[i] = kEmpty;
    if (threadIdx.x == 0) {
      s_fill = 0;
      s_over = 0;
    }
    __syncthreads();
    unsigned long long cand = 0;
    auto insert = [&](uint64_t key) {
      ++cand;
      if (s_over) return;
      uint32_t h = hash_slot(key, kHashBits);
      for (int probe = 0; probe < kHashCap; ++probe) {
        unsigned long long old = atomicCAS(&table[h], kEmpty, (unsigned long long)key);
        if (old == kEmpty) {
          if (atomicAdd(&s_fill, 1) >= kHashCap * 3 / 4) s_over = 1;
          return;
        }
        if (old == key) return;
        h = (h + 1) & (kHashCap - 1);
      }
      s_over = 1;
    };
    if (ROOT) {
      const uint64_t base = (uint64_t)job * a.K;
      for (int u = threadIdx.x; u < nv; u += blockDim.x) {
        const int lu = a.vlab[vb + u];
        if (mod > 1 && ((uint32_t)lu & (mod - 1)) != res) continue;
        enumerate_root(G, u, lu, a.NEL, a.NVL, [&](int key, int) { insert(base + (uint64_t)key); });
      }
    } else {
      for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const int32_t p = a.pid[r];
        if (mod > 1 && ((uint32_t)p & (mod - 1)) != res) continue;
        const Tuple t = load_tuple(a.vt, r, a.S);
        const uint64_t base = (uint64_t)p * (uint64_t)a.K;
        enumerate_row(G, a.meta + p, t, a.S, a.NEL, a.NVL, [&](int key, int) { insert(base + (uint64_t)key); });
      }
    }
    __syncthreads();
    const bool over = s_over != 0;
    __syncthreads();
    if (over) {
      // Split the residue class: keys of one pattern (or one root label)
      // never straddle classes, so per-graph dedup stays exact.
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
    for (int i = threadIdx.x; i < kHashCap; i += blockDim.x) {
      const unsigned long long k = table[i];
      if (k == kEmpty) continue;
      const int32_t old = atomicAdd(&a.sup[k], 1);
      const int32_t owner = ROOT ? (int32_t)(k / (uint64_t)a.K) : a.meta[k / (uint64_t)a.K].job;
      if (old + 1 == a.minsup_job[owner]) {
        const uint32_t slot = atomicAdd(a.freq_count, 1u);
        if (slot < a.freq_cap) a.freq[slot] = k;
      }
    }
    if (threadIdx.x == 0) s_top -= 1;
  }
  // candidate counter (E_cand): one atomic per wave
  for (int o = 32; o > 0; o >>= 1) my_cand += __shfl_down(my_cand, o, 64);
  if ((threadIdx.x & 63) == 0 && my_cand) atomicAdd(a.ecand, my_cand);
}

// ------------------------------------------------------------ COUNT/WRITE
template <bool ROOT, bool WRITE>
__global__ void __launch_bounds__(256) k_children(LevelArgs a) {
  __shared__ Shared sh;
  __shared__ int s_waves[17];
  const int g = blockIdx.x;
  int nv;
  int64_t r0, r1;
  if (ROOT) {
    r0 = 0;
    r1 = 0;
  } else {
    r0 = a.grow[g];
    r1 = a.grow[g + 1];
    if (r0 == r1) {
      if (!WRITE && threadIdx.x == 0) a.ccount[g] = 0;
      return;
    }
  }
  GView G = stage_graph(a, g, sh, nv);
  const int job = ROOT ? a.gjob[g] : 0;
  const int32_t vb = a.gvbase[g];
  if (ROOT) {
    r0 = 0;
    r1 = nv;
  }
  const int64_t nrows = r1 - r0;
  int64_t out = WRITE ? a.cgrow[g] : 0;
  long long total_count = 0;
  for (int64_t base = 0; base < nrows; base += blockDim.x) {
    const int64_t r = r0 + base + threadIdx.x;
    const bool live = base + threadIdx.x < nrows;
    int cnt = 0;
    Tuple t{0, 0, 0, 0};
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
        t = load_tuple(a.vt, r, a.S);
        kbase = (uint64_t)p * (uint64_t)a.K;
        enumerate_row(G, pm, t, a.S, a.NEL, a.NVL, [&](int key, int) {
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
            Tuple c{0, 0, 0, 0};
            tset(c, 0, u);
            tset(c, 1, w);
            store_tuple(a.cvt, pos, a.CS, c);
            ++pos;
          }
        });
      } else {
        const int nvp = pm->nv;
        enumerate_row(G, pm, t, a.S, a.NEL, a.NVL, [&](int key, int w) {
          const int32_t v = a.sup[kbase + key];
          if (v < 0) {
            a.cpid[pos] = -v - 1;
            Tuple c = t;
            if (w >= 0) tset(c, nvp, w);
            store_tuple(a.cvt, pos, a.CS, c);
            ++pos;
          }
        });
      }
    }
    out += tot;
  }
  if (!WRITE) {
    // block reduction of the per-thread counts
    long long v = total_count;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    __shared__ long long s_red[16];
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      long long s = 0;
      for (int i = 0; i < (int)((blockDim.x + 63) >> 6); ++i) s += s_red[i];
      a.ccount[g] = s;
    }
  }
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
