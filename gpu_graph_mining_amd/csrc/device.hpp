// device.hpp -- device-side layouts shared by the HIP kernels and the host
// driver of the miner.
//
// Database layout in HBM (one copy per gm_db):
//   gvbase[g]      first global vertex of graph g (graphs are contiguous)
//   vadj[v]        first adjacency entry of global vertex v; vadj[NV] = 2E
//   vlab[v]        dense vertex label (uint8)
//   adj[a]         packed adjacency entry: bits 0-7 neighbour's local id,
//                  8-15 dense edge label, 16-23 neighbour's dense label
//   gjob[g]        job index of graph g (0 unless several jobs are batched)
//   vgraph[v]      graph of global vertex v
// Local vertex ids fit a byte, so a graph has at most 256 vertices.  This is
// the padded-CSR graph_database_cuda of the reference
// (src/globals/cuda_graph_types.hpp:16-166) without the padding and with the
// neighbour label folded into the entry, so one 4-byte read answers every
// question the extension rules ask about a neighbour.
//
// Embedding store, one level at a time, pattern-major: the rows of a
// pattern are contiguous, and within them the rows of one graph (a segment):
//   pid[r]               pattern of row r
//   rgraph[r]            graph of row r, as the global id of the graph's
//                        first vertex (unique among graphs that have
//                        vertices; the row's vertices are found without a
//                        lookup of the graph's base)
//   vt[r*S .. r*S+S)     the row's vertex tuple: byte i = local id of pattern
//                        vertex i (DFS discovery order), S = stride in bytes
//   rbound[r]            the row's item count (adjacency entries of its
//                        rightmost vertex and rightmost-path sources): the
//                        size of its candidate-list region
// A row is one embedding of its pattern (|Proj(P)| rows per pattern), the
// counterpart of the reference's column store (embedding_lists.hpp:12-85)
// with whole tuples instead of back-linked columns.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gm {

constexpr int kMaxPatV = 64;     // vertices per pattern
constexpr int kNarrowPatV = 32;  // up to here tuple strides are multiples of 4 bytes; beyond, 64
constexpr int kMaxGraphV = 256;
constexpr uint32_t kPatDead = 1;  // PatMetaT::flags

// Tuple stride in bytes of a level whose largest pattern has nv vertices.
__host__ __device__ constexpr int tuple_stride(int nv) { return nv <= kNarrowPatV ? (nv + 3) & ~3 : kMaxPatV; }

// Per-pattern facts the extension rules need (src/globals/graph_types.cpp:
// 305-413 evaluated on the pattern side).  Rightmost-path edge i (i = 0 is
// the last forward edge, as DFSCode::buildRMPath orders them) is packed in
// rm[i]: bits 0-7 its from vertex, 8-15 its edge label, 16-23 the label of
// its to vertex.  Two layouts with the same first 16 bytes: V = 32 for the
// levels whose tuple stride is at most 32 bytes (every level of the common
// workloads, 192 bytes a pattern) and V = 64 for the levels of stride 64
// (patterns of 33-64 vertices, 352 bytes).  The vertex masks are split in
// two 32-bit halves; the high halves are zero in the narrow layout.
template <int V>
struct alignas(16) PatMetaT {
  int32_t job;
  uint8_t nv;        // vertex count; the rightmost vertex is nv-1
  uint8_t nrm;       // rightmost-path edges
  uint8_t minlabel;  // dense label of vertex 0
  uint8_t maxlabel;  // dense label of the rightmost vertex
  uint32_t bwd_mask; // bit t < 32: the pattern already has edge (rightmost, t)
  uint32_t flags;    // kPatDead: the code is not minimal (its rows are not extended)
  uint32_t rm[V];
  uint32_t bwd_from;     // bit f < 32: f is the from vertex of rightmost-path edge q >= 1 (a backward target)
  uint8_t rmq[V];        // that q, by vertex
  uint32_t bwd_mask_hi;  // bits 32-63 of the two masks
  uint32_t bwd_from_hi;
  uint32_t pad;
};
using PatMeta = PatMetaT<kNarrowPatV>;
using PatMetaWide = PatMetaT<kMaxPatV>;
static_assert(sizeof(PatMeta) == 192 && sizeof(PatMetaWide) == 352, "PatMeta layouts");
// bytes per pattern of the metadata of a level of tuple stride S
__host__ __device__ constexpr int meta_bytes(int S) { return S > kNarrowPatV ? (int)sizeof(PatMetaWide) : (int)sizeof(PatMeta); }
// The first 16 bytes (job, nv, nrm, labels, low mask half, flags) of
// pattern p's metadata in a level of tuple stride S, either layout.
__host__ __device__ inline const PatMeta *meta_at(const PatMeta *m, int S, int64_t p) {
  return reinterpret_cast<const PatMeta *>(reinterpret_cast<const char *>(m) + p * meta_bytes(S));
}

// Extension key of one pattern at a level whose tuple stride is S:
//   backward to vertex t with edge label e   : t*NEL + e
//   forward from vertex f, edge e, label l   : S*NEL + (f*NEL + e)*NVL + l
// K = S*NEL*(NVL+1) keys per pattern; (pattern, key) -> p*K + key.
struct KeySpace {
  int32_t S, NEL, NVL;
  int64_t K() const { return (int64_t)S * NEL * (NVL + 1); }
};

}  // namespace gm
