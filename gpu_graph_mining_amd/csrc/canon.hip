// canon.hip -- minimum DFS code check on the device (gfx950, wave64).
//
// One 64-lane wave decides for one DFS code C of n edges whether C is the
// code that gSpan's greedy is_min grows on C's own pattern graph
// (src/gspan/ismin.cpp:36-140; src/globals/dfs_ismin.cpp:23-160), with the
// same steps as the host check in canonical.cpp:
//   1. the root tuple is the smallest (fromlabel, elabel, tolabel) over the
//      directed pattern edges with fromlabel <= tolabel (not compared with
//      C[0]);
//   2. with the embeddings of the code grown so far, the next tuple is the
//      backward edge from the rightmost vertex to the rightmost-path vertex
//      nearest the root that an embedding reaches (smallest elabel, under
//      get_backward's order rule, graph_types.cpp:389-413); else the
//      smallest (elabel, tolabel) forward edge from the rightmost vertex
//      (graph_types.cpp:337-355); else from the deepest rightmost-path
//      vertex that has one (graph_types.cpp:305-323);
//   3. the first tuple that differs from C answers "not minimal"; reaching
//      the end answers "minimal".
// The pattern graph (an edge-id matrix and neighbour masks) and two
// embedding lists (vertex map, used-edge and used-vertex masks) live in
// LDS; the lanes take the embeddings, and each step's smallest tuple is a
// wave-wide minimum.  A code beyond the LDS capacity (more than kCanonE
// edges, kCanonV - 1 vertices, kCanonCap embeddings, or a repeated edge)
// is left undecided (kCanonHost) for the host check.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace gm {
namespace {

constexpr int kCV = kCanonV;
constexpr int kCE = kCanonE;
constexpr int kCap = kCanonCap;
constexpr unsigned long long kNone = ~0ull;

struct Emb {
  unsigned long long used, vused;  // pattern edges / vertices matched
  uint8_t map[kCV];                // code vertex -> pattern vertex
};
static_assert(sizeof(Emb) == 16 + kCV && kCV % 8 == 0, "Emb layout");

struct CanonLds {
  unsigned long long nb[kCV];  // neighbour mask per pattern vertex
  int32_t lab[kCV];            // vertex labels
  int32_t wel[kCE], wtl[kCE];  // C: edge label, to label per tuple
  uint8_t wfrom[kCE], wto[kCE];
  uint8_t rm[kCE];             // rightmost path of the code grown so far, last forward tuple first
  alignas(4) int8_t eid[kCV * kCV];  // tuple index of the edge between two vertices, -1: none
  int32_t nrm;
  Emb emb[2][kCap];
};

// word copy of an embedding (a struct copy goes through private memory)
__device__ __forceinline__ void copy_emb(Emb &dst, const Emb &src) {
  const unsigned long long *s = reinterpret_cast<const unsigned long long *>(&src);
  unsigned long long *d = reinterpret_cast<unsigned long long *>(&dst);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(Emb) / 8); ++k) d[k] = s[k];
}

// (a, b) as one unsigned key with the order of the signed pair
__device__ __forceinline__ unsigned long long pair_key(int32_t a, int32_t b) {
  return ((unsigned long long)((uint32_t)a ^ 0x80000000u) << 32) | (unsigned long long)((uint32_t)b ^ 0x80000000u);
}
__device__ __forceinline__ int32_t key_hi(unsigned long long k) { return (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u); }
__device__ __forceinline__ int32_t key_lo(unsigned long long k) { return (int32_t)((uint32_t)k ^ 0x80000000u); }

__device__ __forceinline__ unsigned long long wave_min64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long t = __shfl_xor(v, o, 64);
    v = t < v ? t : v;
  }
  return v;
}

__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

__device__ __forceinline__ uint64_t below(int lane) { return (1ull << lane) - 1ull; }

// The grown code's tuple i: tuple 0 is the chosen root (0, 1, el0); the
// later ones equal C's (the check stops at the first difference).
struct Grown {
  int32_t el0;
  __device__ __forceinline__ int from(const CanonLds &L, int i) const { return i ? L.wfrom[i] : 0; }
  __device__ __forceinline__ int to(const CanonLds &L, int i) const { return i ? L.wto[i] : 1; }
  __device__ __forceinline__ int32_t el(const CanonLds &L, int i) const { return i ? L.wel[i] : el0; }
};

// The check of C (wfrom/wto/wel/wtl for n tuples, lab for nv vertices, set
// by the caller).  Every lane returns the same answer.
__device__ __forceinline__ int canon_core(CanonLds &L, int n, int nv, int lane) {
  if (n <= 1) return kCanonMin;
  if (n > kCE || nv > kCV - 1) return kCanonHost;
  {
    uint32_t *e32 = reinterpret_cast<uint32_t *>(L.eid);
    for (int i = lane; i < kCV * kCV / 4; i += 64) e32[i] = 0xFFFFFFFFu;
  }
  for (int v = lane; v < kCV; v += 64) L.nb[v] = 0;
  __syncthreads();
  // the pattern graph, one tuple per lane; a tuple outside nv vertices or a
  // repeated vertex pair leaves the code to the host
  int bad = 0;
  for (int i = lane; i < n; i += 64) {
    const int a = L.wfrom[i], b = L.wto[i];
    if (a >= nv || b >= nv || a == b) {
      bad = 1;
      continue;
    }
    L.eid[a * kCV + b] = (int8_t)i;
    L.eid[b * kCV + a] = (int8_t)i;
    atomicOr(&L.nb[a], 1ull << b);
    atomicOr(&L.nb[b], 1ull << a);
  }
  __syncthreads();
  for (int i = lane; i < n; i += 64) {
    const int a = L.wfrom[i], b = L.wto[i];
    if (a < nv && b < nv && a != b && (L.eid[a * kCV + b] != i || L.eid[b * kCV + a] != i)) bad = 1;
  }
  if (__ballot(bad != 0)) return kCanonHost;

  // 1. root tuple: smallest (fromlabel, elabel), then smallest tolabel
  unsigned long long k1 = kNone;
  for (int d = lane; d < 2 * n; d += 64) {
    const int i = d >> 1;
    const int u = (d & 1) ? L.wto[i] : L.wfrom[i], v = (d & 1) ? L.wfrom[i] : L.wto[i];
    if (L.lab[u] <= L.lab[v]) {
      const unsigned long long k = pair_key(L.lab[u], L.wel[i]);
      k1 = k < k1 ? k : k1;
    }
  }
  k1 = wave_min64(k1);
  const int32_t bfl = key_hi(k1), bel = key_lo(k1);
  unsigned long long k2 = kNone;
  for (int d = lane; d < 2 * n; d += 64) {
    const int i = d >> 1;
    const int u = (d & 1) ? L.wto[i] : L.wfrom[i], v = (d & 1) ? L.wfrom[i] : L.wto[i];
    if (L.lab[u] == bfl && L.wel[i] == bel && L.lab[u] <= L.lab[v]) {
      const unsigned long long k = pair_key(L.lab[v], 0);
      k2 = k < k2 ? k : k2;
    }
  }
  k2 = wave_min64(k2);
  const int32_t btl = key_hi(k2);
  int ncur = 0;
  for (int base = 0; base < 2 * n; base += 64) {
    const int d = base + lane;
    bool hit = false;
    int u = 0, v = 0, i = 0;
    if (d < 2 * n) {
      i = d >> 1;
      u = (d & 1) ? L.wto[i] : L.wfrom[i];
      v = (d & 1) ? L.wfrom[i] : L.wto[i];
      hit = L.lab[u] == bfl && L.wel[i] == bel && L.lab[v] == btl;
    }
    const unsigned long long m = __ballot(hit);
    if (hit) {
      Emb &e = L.emb[0][ncur + __popcll(m & below(lane))];
      e.used = 1ull << i;
      e.vused = (1ull << u) | (1ull << v);
      e.map[0] = (uint8_t)u;
      e.map[1] = (uint8_t)v;
    }
    ncur += __popcll(m);
  }
  const Grown g{bel};
  const int32_t minlabel = bfl;
  int cur = 0;

  for (int mpos = 1; mpos < n; ++mpos) {
    __syncthreads();
    if (lane == 0) {  // rightmost path of tuples [0, mpos) (dfs_code.cpp:75-90)
      int nrm = 0, old_from = -1;
      for (int i = mpos - 1; i >= 0; --i) {
        const int f = g.from(L, i), t = g.to(L, i);
        if (f < t && (nrm == 0 || old_from == t)) {
          L.rm[nrm++] = (uint8_t)i;
          old_from = f;
        }
      }
      L.nrm = nrm;
    }
    __syncthreads();
    const int nrm = L.nrm;
    const int maxtoc = g.to(L, L.rm[0]);
    if (maxtoc + 1 >= kCV) return kCanonHost;
    const Emb *E = L.emb[cur];
    Emb *X = L.emb[cur ^ 1];

    // backward edges from the rightmost vertex: the path vertex nearest the
    // root that any embedding reaches, then the smallest edge label -- one
    // minimum of (nrm - path index, elabel) over every embedding's first hit
    // from the root end
    unsigned long long bk = kNone;
    for (int j = lane; j < ncur; j += 64) {
      const int r = E[j].map[maxtoc];
      for (int ii = nrm - 1; ii >= 1; --ii) {
        const int e1 = L.rm[ii];
        const int id = L.eid[r * kCV + E[j].map[g.from(L, e1)]];
        if (id < 0 || ((E[j].used >> id) & 1ull)) continue;
        const int32_t le = L.wel[id], e1el = g.el(L, e1);
        if (e1el < le || (e1el == le && L.lab[E[j].map[g.to(L, e1)]] <= L.lab[r])) {
          const unsigned long long k = pair_key(nrm - ii, le);
          bk = k < bk ? k : bk;
          break;
        }
      }
    }
    bk = wave_min64(bk);
    if (bk != kNone) {
      const int ii = nrm - key_hi(bk);
      const int32_t bl = key_lo(bk);
      const int e1 = L.rm[ii];
      const int e1f = g.from(L, e1), e1t = g.to(L, e1);
      const int32_t e1el = g.el(L, e1);
      if (!(L.wfrom[mpos] == maxtoc && L.wto[mpos] == e1f && L.wel[mpos] == bl)) return kCanonNotMin;
      int nn = 0;
      for (int base = 0; base < ncur; base += 64) {
        const int j = base + lane;
        bool hit = false;
        int id = 0;
        if (j < ncur) {
          const int r = E[j].map[maxtoc], target = E[j].map[e1f];
          id = L.eid[r * kCV + target];
          if (id >= 0 && !((E[j].used >> id) & 1ull)) {
            const int32_t le = L.wel[id];
            hit = le == bl && (e1el < le || (e1el == le && L.lab[E[j].map[e1t]] <= L.lab[r]));
          }
        }
        const unsigned long long m = __ballot(hit);
        if (hit) {
          Emb &x = X[nn + __popcll(m & below(lane))];
          copy_emb(x, E[j]);
          x.used |= 1ull << id;
        }
        nn += __popcll(m);
      }
      ncur = nn;
      cur ^= 1;
      continue;
    }

    // forward edges: from the rightmost vertex (source level -1), else from
    // the deepest rightmost-path vertex that has one (level = path index);
    // first the smallest level any embedding has, then the smallest
    // (elabel, tolabel) at that level.  cand(j, lvl, w) is the rule.
    auto fwd_ok = [&](int j, int lvl, int w, int32_t &le, int32_t &tl) -> bool {
      tl = L.lab[w];
      if (minlabel > tl) return false;
      if (lvl < 0) {
        le = L.wel[L.eid[E[j].map[maxtoc] * kCV + w]];
        return true;
      }
      const int e = L.rm[lvl];
      const int u = E[j].map[g.from(L, e)], x = E[j].map[g.to(L, e)];
      if (w == x) return false;
      le = L.wel[L.eid[u * kCV + w]];
      const int32_t eel = g.el(L, e);
      return eel < le || (eel == le && L.lab[x] <= tl);
    };
    auto src_of = [&](int j, int lvl) -> int { return lvl < 0 ? E[j].map[maxtoc] : E[j].map[g.from(L, L.rm[lvl])]; };
    int lv_lane = 1 << 30;
    for (int j = lane; j < ncur; j += 64) {
      for (int lvl = -1; lvl < nrm && lvl < lv_lane; ++lvl) {
        unsigned long long c = L.nb[src_of(j, lvl)] & ~E[j].vused;
        bool any = false;
        while (c && !any) {
          const int w = __ffsll((long long)c) - 1;
          c &= c - 1;
          int32_t le, tl;
          any = fwd_ok(j, lvl, w, le, tl);
        }
        if (any) lv_lane = lvl;
      }
    }
    int lvl = lv_lane;
    for (int o = 32; o > 0; o >>= 1) lvl = min(lvl, __shfl_xor(lvl, o, 64));
    if (lvl >= nrm) return kCanonMin;  // nothing left to grow: the pattern is covered
    unsigned long long fk = kNone;
    for (int j = lane; j < ncur; j += 64) {
      unsigned long long c = L.nb[src_of(j, lvl)] & ~E[j].vused;
      while (c) {
        const int w = __ffsll((long long)c) - 1;
        c &= c - 1;
        int32_t le, tl;
        if (fwd_ok(j, lvl, w, le, tl)) {
          const unsigned long long k = pair_key(le, tl);
          fk = k < fk ? k : fk;
        }
      }
    }
    fk = wave_min64(fk);
    const int newfrom = lvl < 0 ? maxtoc : g.from(L, L.rm[lvl]);
    const int32_t bel2 = key_hi(fk), btl2 = key_lo(fk);
    if (!(L.wfrom[mpos] == newfrom && L.wto[mpos] == maxtoc + 1 && L.wel[mpos] == bel2 && L.wtl[mpos] == btl2))
      return kCanonNotMin;
    // every embedding's images of the new edge: count, scan, write
    int nn = 0;
    for (int base = 0; base < ncur; base += 64) {
      const int j = base + lane;
      unsigned long long hits = 0;
      if (j < ncur) {
        unsigned long long c = L.nb[src_of(j, lvl)] & ~E[j].vused;
        while (c) {
          const int w = __ffsll((long long)c) - 1;
          c &= c - 1;
          int32_t le, tl;
          if (fwd_ok(j, lvl, w, le, tl) && le == bel2 && tl == btl2) hits |= 1ull << w;
        }
      }
      const int cnt = __popcll(hits);
      const int incl = wave_incl_scan(cnt, lane);
      const int total = __shfl(incl, 63, 64);
      if (nn + total > kCap) return kCanonHost;
      int pos = nn + incl - cnt;
      if (j < ncur) {
        const int u = src_of(j, lvl);
        while (hits) {
          const int w = __ffsll((long long)hits) - 1;
          hits &= hits - 1;
          Emb &y = X[pos++];
          copy_emb(y, E[j]);
          y.map[maxtoc + 1] = (uint8_t)w;
          y.used |= 1ull << L.eid[u * kCV + w];
          y.vused |= 1ull << w;
        }
      }
      nn += total;
    }
    ncur = nn;
    cur ^= 1;
  }
  return kCanonMin;
}

// Frequent keys of a later level: key -> (parent, new tuple) -> C.  The
// answer + 1 goes to bits 32-39 of the record's support word.
__global__ void __launch_bounds__(64) k_canon_level(CanonArgs c) {
  __shared__ CanonLds L;
  const int lane = threadIdx.x;
  const uint32_t nf = min(*c.freq_count, c.freq_cap);
  for (uint32_t slot = blockIdx.x; slot < nf; slot += gridDim.x) {
    unsigned long long *f = c.freq + 4ull * slot;
    const unsigned long long key = f[0];
    const int64_t p = (int64_t)(key / (unsigned long long)c.K) + c.p_base;
    const int64_t x = (int64_t)(key % (unsigned long long)c.K);
    const int nvp = meta_at(c.meta, c.S, p)->nv;
    const int len = c.len;
    __syncthreads();  // the previous key's LDS is no longer read
    for (int t = lane; t < len && t < kCE; t += 64) {
      const uint32_t w = c.pcode[p * len + t];
      L.wfrom[t] = (uint8_t)(w & 0xFF);
      L.wto[t] = (uint8_t)((w >> 8) & 0xFF);
      L.wel[t] = (int32_t)((w >> 16) & 0xFF);
      L.wtl[t] = (int32_t)(w >> 24);
    }
    for (int v = lane; v < nvp && v < kCV; v += 64) L.lab[v] = c.pvlab[p * c.S + v];
    int nv = nvp;
    if (lane == 0 && len < kCE) {
      if (x < (int64_t)c.S * c.NEL) {  // backward: rightmost -> to
        const int to = (int)(x / c.NEL);
        L.wfrom[len] = (uint8_t)(nvp - 1);
        L.wto[len] = (uint8_t)to;
        L.wel[len] = (int32_t)(x % c.NEL);
        L.wtl[len] = c.pvlab[p * c.S + to];
      } else {  // forward: from -> new vertex nvp
        const int64_t y = x - (int64_t)c.S * c.NEL;
        L.wfrom[len] = (uint8_t)(y / c.NVL / c.NEL);
        L.wto[len] = (uint8_t)nvp;
        L.wel[len] = (int32_t)((y / c.NVL) % c.NEL);
        L.wtl[len] = (int32_t)(y % c.NVL);
        if (nvp < kCV) L.lab[nvp] = (int32_t)(y % c.NVL);
      }
    }
    if (x >= (int64_t)c.S * c.NEL) nv = nvp + 1;
    __syncthreads();
    const int res = canon_core(L, len + 1, nv, lane);
    // answer + 1, so that a record the kernel did not reach reads as 0
    if (lane == 0) f[1] = (f[1] & 0xFFFFFFFFull) | ((unsigned long long)(res + 1) << 32);
  }
}

// Explicit codes: n codes of len tuples (from, to, fromlabel, elabel,
// tolabel); vertex labels as the code carries them (-1: none, 0 default),
// later tuples overriding earlier ones (DFSCode::toGraph,
// src/globals/dfs_code.cpp:245-265).
__global__ void __launch_bounds__(64) k_canon_list(CanonArgs c) {
  __shared__ CanonLds L;
  const int lane = threadIdx.x;
  const int len = c.len;
  for (int64_t q = blockIdx.x; q < c.n; q += gridDim.x) {
    const int32_t *code = c.codes + q * (int64_t)len * 5;
    __syncthreads();
    int nv = 0, bad = 0;
    for (int t = 0; t < len; ++t) {  // uniform loads
      const int a = code[5 * t], b = code[5 * t + 1];
      if (a < 0 || b < 0 || a == b) bad = 1;
      nv = max(nv, max(a, b) + 1);
    }
    if (bad) {
      if (lane == 0) c.out[q] = kCanonBad;
      continue;
    }
    if (len > kCE || nv > kCV - 1) {
      if (lane == 0) c.out[q] = kCanonHost;
      continue;
    }
    for (int t = lane; t < len; t += 64) {
      L.wfrom[t] = (uint8_t)code[5 * t];
      L.wto[t] = (uint8_t)code[5 * t + 1];
      L.wel[t] = code[5 * t + 3];
      L.wtl[t] = code[5 * t + 4];
    }
    if (lane == 0) {
      for (int v = 0; v < nv; ++v) L.lab[v] = 0;
      for (int t = 0; t < len; ++t) {
        if (code[5 * t + 2] != -1) L.lab[code[5 * t]] = code[5 * t + 2];
        if (code[5 * t + 4] != -1) L.lab[code[5 * t + 1]] = code[5 * t + 4];
      }
    }
    __syncthreads();
    const int res = canon_core(L, len, nv, lane);
    if (lane == 0) c.out[q] = (uint8_t)res;
  }
}

}  // namespace

hipError_t launch_canon_level(const CanonArgs &c, int grid, hipStream_t s) {
  if (grid <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_canon_level, dim3(grid), dim3(64), 0, s, c);
  return hipGetLastError();
}

hipError_t launch_canon_list(const CanonArgs &c, hipStream_t s) {
  if (c.n <= 0) return hipSuccess;
  const unsigned grid = (unsigned)(c.n < 16384 ? c.n : 16384);
  hipLaunchKernelGGL(k_canon_list, dim3(grid), dim3(64), 0, s, c);
  return hipGetLastError();
}

}  // namespace gm
