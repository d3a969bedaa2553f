/*
 * gspan_amd.h -- C ABI of the MI355X-native gSpan miner.
 *
 * One shared library (gpu_graph_mining_amd/_lib/libgspan_amd.so) exports
 * these symbols.  Plain C types only; no exceptions cross the boundary.
 * Every entry point returns GM_OK (0) or a negative error class; the text of
 * the last error on the calling thread is available from gm_last_error().
 *
 * The reference's drop-in boundary is the C++ class gspan_cuda
 * (src/gspan_cuda/gspan_cuda.hpp:53-293) plus the graph_output sink
 * (src/globals/graph_output.hpp:9-52) and the dbio readers
 * (src/globals/dbio.hpp:8-37).  Each declaration below names the reference
 * interface it replaces.
 *
 * Threading: a gm_db handle is thread-compatible (no concurrent calls on one
 * handle without external locking).  Pattern callbacks run synchronously on
 * the calling thread, as graph_output::output_graph does in the reference
 * (src/gspan_cuda/gspan_cuda_no_sort.cpp:107, 416).
 */
#ifndef GSPAN_AMD_H
#define GSPAN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error classes ---------------------------------------------------- */
#define GM_OK 0
#define GM_E_ARG (-1)       /* bad argument / malformed input            */
#define GM_E_IO (-2)        /* file could not be opened / read / written */
#define GM_E_OOM (-3)       /* host or device allocation failed          */
#define GM_E_DEVICE (-4)    /* HIP runtime / kernel launch failure       */
#define GM_E_UNSUPPORTED (-5) /* input outside the supported envelope    */
#define GM_E_CALLBACK (-6)  /* the pattern callback asked to stop        */

/* ---- label style of reported DFS codes -------------------------------- */
/* GM_LABELS_FULL: every DFS tuple fully labelled, as gspan_cuda reports
 *   (src/gspan_cuda/gspan_cuda_no_sort.cpp:409-416).
 * GM_LABELS_WILDCARD: fromlabel = -1 on every edge after the first and
 *   tolabel = -1 on backward edges, as the sequential gSpan reports
 *   (src/gspan/gspan.cpp:253, 266). */
#define GM_LABELS_FULL 0
#define GM_LABELS_WILDCARD 1

typedef struct gm_db gm_db;

/* Mining options.  Zero-initialise and set what you need; gm_cfg_init
 * fills the defaults.  Replaces cudapp::cuda_configurator
 * (src/cuda_pp/cuda_configurator.hpp) as the tuning surface. */
typedef struct gm_cfg {
  int32_t label_style;      /* GM_LABELS_FULL (default) or GM_LABELS_WILDCARD */
  int32_t max_edges;        /* stop growing patterns beyond this many edges; 0 = unbounded */
  int32_t canon_threads;    /* host threads for the canonical check; 0 = auto */
  int32_t block_threads;    /* accepted for compatibility (0, 64, 128 or 256) and otherwise unused: the
                               level kernels run one 64-lane wave per chunk of rows */
  int64_t mem_budget_bytes; /* device bytes for the embedding rows and candidate lists of the levels
                               held at once; a level whose children would pass it is written in
                               pieces by parent pattern, each piece's patterns mined to the end
                               before the next (more when one pattern's own children need more);
                               0 = unlimited (every level whole).  Ignored by gm_mine_noalloc */
  int32_t collect_proj;     /* 1 = also report |Proj(P)| per pattern (no extra pass) */
  int32_t candidate_lists;  /* accepted for compatibility and unused: the count pass always keeps each
                               embedding's candidates for the write pass */
  int32_t path_pruning;     /* 0 = drop extension candidates that have an infrequent 2-edge sub-path
                               (exact: such a child is infrequent; default); 1 = keep every candidate */
  int32_t speculate;        /* children written before their canonical check (the check overlaps the
                               write; non-minimal ones are dropped at the next level): 0 = when the
                               last level's non-minimal share was small (default), 1 = never, 2 = always */
  int32_t stage_events;     /* 1 = time each level's extension + support stage with HIP events on the
                               mining stream (gm_stats.seconds_stage_events); 0 = off (default) */
  int32_t chunk_rows;       /* launch shape: rows per wave of the level kernels (64..4096, a power of
                               two; smaller levels use shorter chunks); 0 = default (256) */
  int32_t table_bits;       /* launch shape: log2 entries of a wave's LDS count table (7..11); a chunk
                               that fills it is recounted; 0 = default (8) */
  int32_t canon_host;       /* where the minimum-code check of a level's frequent keys runs: 0 = on
                               the device before their download for the levels whose children are
                               not written speculatively, else on the host thread pool while the
                               device writes them (default); 1 = always on the host; 2 = always on
                               the device */
} gm_cfg;

/* Per-run statistics (filled when a non-NULL pointer is passed). */
typedef struct gm_stats {
  uint64_t n_patterns;     /* reported frequent, minimal DFS codes        */
  uint64_t e_child;        /* sum |Proj(P)| over reported P, >= 2 edges   */
  uint64_t e_root;         /* sum |Proj(P)| over reported 1-edge P        */
  uint64_t e_cand;         /* extension candidates enumerated             */
  uint64_t n_candidates;   /* frequent (pattern, extension) keys checked  */
  uint64_t n_levels;       /* level steps executed                        */
  double seconds_total;    /* wall time of gm_mine (device upload excluded) */
  double seconds_device;   /* time inside level kernels (HIP events)      */
  double seconds_canon;    /* host time in the canonical check            */
  double seconds_ext_kernel; /* time inside the extension/support kernel  */
  uint64_t ext_kernel_launches;
  uint64_t peak_device_bytes;
  uint64_t split_passes;   /* chunks recounted because their LDS count table filled */
  uint64_t max_graph_rows; /* unused (0) */
  uint64_t max_table_fill; /* most distinct keys one chunk's count table held */
  uint64_t pruned;         /* candidates dropped by the 2-edge path rule (counted in e_cand) */
  double seconds_stage_events; /* extension + support stages timed by HIP events (cfg.stage_events) */
  uint64_t canon_host_keys;    /* frequent keys the device check left to the host (beyond its capacity) */
} gm_stats;

/* Pattern callback: n_edges rows of (from, to, fromlabel, elabel, tolabel),
 * labels in the caller's original label values, plus the support (number
 * of distinct database graphs containing the pattern) and, when
 * cfg.collect_proj is set, |Proj(P)| (else 0).  Return 0 to continue,
 * non-zero to stop mining (gm_mine then returns GM_E_CALLBACK).
 * Replaces graph_output::output_graph(const DFSCode&, size_t support)
 * (src/globals/graph_output.hpp:12). */
typedef int (*gm_pattern_cb)(void *user, const int32_t *dfs5, int32_t n_edges, uint32_t support,
                             uint64_t n_proj);

/* Batch variant of the callback: also gets the job index. */
typedef int (*gm_job_pattern_cb)(void *user, int32_t job, const int32_t *dfs5, int32_t n_edges,
                                 uint32_t support, uint64_t n_proj);

/* ---- library ---------------------------------------------------------- */
const char *gm_version(void);
/* Text of the last error raised on this thread ("" if none). */
const char *gm_last_error(void);
/* Fill *cfg with defaults. */
void gm_cfg_init(gm_cfg *cfg);
/* Number of visible HIP devices (0 when no GPU); never fails. */
int gm_device_count(void);

/* ---- graph databases -------------------------------------------------- */
/* Create a database from per-graph vertex and edge lists (the in-memory
 * form of types::graph_database_t, src/globals/types.hpp:31).  Graph g owns
 * vertices [graph_vbegin[g], graph_vbegin[g+1]) with labels vlabels[], and
 * edges [graph_ebegin[g], graph_ebegin[g+1]) given as local vertex ids
 * efrom/eto plus elabel (undirected).  Replaces
 * gspan_cuda::set_database(types::graph_database_t&, bool)
 * (src/gspan_cuda/gspan_cuda.hpp:284; gspan_cuda.cpp:73-88).  The device
 * copy is made lazily by the first gm_mine / gm_db_upload. */
int gm_db_create(int32_t n_graphs, const int32_t *graph_vbegin, const int32_t *vlabels,
                 const int32_t *graph_ebegin, const int32_t *efrom, const int32_t *eto,
                 const int32_t *elabel, gm_db **out);

/* Read a database file.  filetype: "txt", "fsg", "bin" or "cudabin", each
 * optionally prefixed with '-', as dbio::ftype2str accepts
 * (src/globals/dbio.cpp:33-41).  Replaces dbio::read_database /
 * read_database_cudabin (src/globals/dbio.hpp:12, 21-22). */
int gm_db_load(const char *path, const char *filetype, gm_db **out);

/* Write the database as txt ("t # i" / "v id label" / "e from to label"),
 * the format Graph::read parses (src/globals/graph_types.cpp:66-113). */
int gm_db_write_txt(const gm_db *db, const char *path);

/* Write the database in `filetype` "txt", "bin" or "cudabin" (optionally
 * '-'-prefixed).  Replaces dbio::write_database_bin /
 * write_database_cudabin (src/globals/dbio.hpp; dbio.cpp:209-268,
 * 346-370); bin rejects an empty database as the reference does. */
int gm_db_write(const gm_db *db, const char *path, const char *filetype);

/* Read part `part` of `total_parts` of a BIN file: transactions
 * [n / total_parts * part, n / total_parts * (part + 1)), the last part up
 * to n.  Replaces dbio::read_database_bin(filename, gdb, total_parts, part)
 * (src/globals/dbio.cpp:275-329). */
int gm_db_load_part(const char *path, int32_t total_parts, int32_t part, gm_db **out);

/* Sizes of the database: counts[0..5] = n_graphs, n_vertices, n_edges,
 * max vertices per graph, distinct vertex labels, distinct edge labels. */
int gm_db_counts(const gm_db *db, int64_t counts[6]);

/* Copy the database back out in gm_db_create's array form.  Arrays must
 * hold n_graphs+1, n_vertices, n_graphs+1, n_edges (x3) entries. */
int gm_db_export(const gm_db *db, int32_t *graph_vbegin, int32_t *vlabels, int32_t *graph_ebegin,
                 int32_t *efrom, int32_t *eto, int32_t *elabel);

/* Copy the device layout to the GPU now (otherwise done by gm_mine).
 * Counterpart of graph_database_cuda::copy_to_device
 * (src/globals/cuda_graph_types.cpp:130-153). */
int gm_db_upload(gm_db *db);

/* Release host and device memory.  Replaces
 * gspan_cuda::delete_database_from_device (src/gspan_cuda/gspan_cuda.hpp:286)
 * and the destructor (gspan_cuda.cpp:52-64). */
void gm_db_destroy(gm_db *db);

/* ---- mining ----------------------------------------------------------- */
/* Mine every frequent, minimal DFS code with support >= abs_minsup on the
 * GPU and report each through cb.  Replaces the sequence
 * set_min_support / set_graph_output / run of gspan_cuda
 * (src/gspan_cuda/gspan_cuda.hpp:289-292; gspan_cuda_no_sort.cpp:20-23,
 * 68-136, 275-489).  abs_minsup < 1 is GM_E_ARG (the reference CLI exits
 * with 3, src/gspan_cuda/common_main.cpp:56-70). */
int gm_mine(gm_db *db, int32_t abs_minsup, const gm_cfg *cfg, gm_pattern_cb cb, void *user,
            gm_stats *stats);

/* gm_mine without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  Every buffer of the
 * level steps comes from a static device array of 64 MiB per GPU (held
 * until the process ends; host staging uses ordinary host memory and
 * explicit copies instead of pinned memory; no events).  The database must
 * already be on the device (gm_db_upload; GM_E_ARG otherwise).  A run whose
 * level steps need more than the 64 MiB fails with GM_E_OOM.  Calls on one
 * GPU share its array, which keeps the last call's contents: they take
 * turns (a lock per device), and calls that overlap on one GPU from
 * different processes each have their own array.  Same results as gm_mine;
 * cfg.stage_events is ignored. */
int gm_mine_noalloc(gm_db *db, int32_t abs_minsup, const gm_cfg *cfg, gm_pattern_cb cb, void *user,
                    gm_stats *stats);

/* Mine n_jobs independent databases in one batched schedule: the level
 * steps of every job run in the same kernel launches.  Each job has its own
 * database and absolute min support; results are identical to mining each
 * job alone.  No reference counterpart (the reference mines one database
 * per process); it batches many gm_mine calls. */
int gm_mine_batch(int32_t n_jobs, gm_db *const *dbs, const int32_t *abs_minsup, const gm_cfg *cfg,
                  gm_job_pattern_cb cb, void *user, gm_stats *stats);

/* A resident batch: the databases of n_jobs jobs packed into one device
 * layout once, so that repeated batched mining (or serving) does not upload
 * them again.  No reference counterpart (the reference mines one database
 * per process). */
typedef struct gm_batch gm_batch;
int gm_batch_create(int32_t n_jobs, gm_db *const *dbs, gm_batch **out);
/* A resident batch whose jobs are subsets of one resident database, built
 * on the device from db's own device copy (no host packing, no upload of
 * graph data): job j holds graphs graph_ids[job_begin[j] ..
 * job_begin[j+1]) of db, in that order (job_begin[0] = 0; subsets may
 * overlap).  Mining it gives each job the results of mining its subset on
 * its own.  No reference counterpart (serving). */
int gm_batch_create_subsets(gm_db *db, int32_t n_jobs, const int32_t *job_begin, const int32_t *graph_ids,
                            gm_batch **out);
/* Mine every job of the batch in one schedule with per-job min supports;
 * identical per-job results to gm_mine on each database. */
int gm_batch_mine(gm_batch *batch, const int32_t *abs_minsup, const gm_cfg *cfg, gm_job_pattern_cb cb, void *user,
                  gm_stats *stats);
void gm_batch_destroy(gm_batch *batch);

/* ---- pattern-support queries ------------------------------------------ */
/* Support (number of database graphs containing the pattern) of each of
 * n_queries patterns, given as DFS codes: query q has n_edges[q] rows of
 * (from, to, fromlabel, elabel, tolabel) in dfs5, the queries one after the
 * other.  A code may be any valid DFS code of a connected pattern in which
 * every vertex label is given (forward rows carry tolabel, the first row
 * fromlabel); it need not be minimal.  All queries run in one batched
 * projection over the resident database (level steps shared by all
 * queries).  supports[q] = 0 when the pattern does not occur.  No reference
 * counterpart in the mining path: the reference's support of a code is
 * gSpan::support over its projection (src/gspan/gspan.cpp:88-102), which
 * these queries reproduce for any code. */
int gm_support_queries(gm_db *db, int32_t n_queries, const int32_t *dfs5, const int32_t *n_edges,
                       const gm_cfg *cfg, uint32_t *supports, gm_stats *stats);
/* One query (gm_support_queries with n_queries = 1, default options). */
int gm_support_query(gm_db *db, const int32_t *dfs5, int32_t n_edges, uint32_t *support);

/* ---- canonical form ---------------------------------------------------- */
/* Minimum-DFS-code check of one code (n rows of from,to,fromlabel,elabel,
 * tolabel; either label style).  Returns 1 minimal, 0 not minimal, or a
 * negative error.  Replaces types::DFSCode::dfs_code_is_min
 * (src/globals/dfs_ismin.cpp:23-160). */
int gm_dfs_code_is_min(const int32_t *dfs5, int32_t n_edges);

/* The same check for n_codes codes of n_edges tuples each (dfs5 holds
 * n_codes * n_edges rows), on the device: one wave per code.  out[i] = 1
 * minimal, 0 not; each answer equals gm_dfs_code_is_min's.  Codes beyond
 * the device check's capacity (more than 64 edges or 39 vertices, or more
 * than 64 embeddings of a prefix in the code's own graph) are checked on
 * the host.  The batch form of the check gspan_cuda runs per candidate
 * (src/gspan_cuda/gspan_cuda_no_sort.cpp:408-414 calling
 * dfs_ismin.cpp:23-160); it allocates device memory for the codes. */
int gm_dfs_codes_are_min(int32_t n_codes, int32_t n_edges, const int32_t *dfs5, uint8_t *out);

/* ---- synthetic databases (seeded, deterministic) ---------------------- */
/* Random connected graphs: n_graphs, vertex count uniform in [vmin,vmax],
 * a random spanning tree (vertex i attaches to a uniform earlier vertex),
 * then uniform extra edges without self-loops or multi-edges until
 * edges = round(edge_factor * vertices); vertex labels uniform over
 * n_vlabels, edge labels uniform over n_elabels. */
int gm_gen_random(uint64_t seed, int32_t n_graphs, int32_t vmin, int32_t vmax, double edge_factor,
                  int32_t n_vlabels, int32_t n_elabels, gm_db **out);

/* Planted-pattern transactions: n_seed_patterns connected seed patterns with
 * max(1, Poisson(mean_pattern_edges)) edges and Exp(1) weights; each
 * transaction targets Poisson(mean_edges) edges, inserts weight-sampled
 * patterns (each pattern vertex merges into an existing same-label vertex
 * with probability merge_prob), then joins its components with random
 * labelled edges. */
int gm_gen_planted(uint64_t seed, int32_t n_graphs, double mean_edges, int32_t n_seed_patterns,
                   double mean_pattern_edges, int32_t n_vlabels, int32_t n_elabels, double merge_prob,
                   gm_db **out);

/* Molecule-like graphs: vertex count Poisson(mean_vertices) (at least 2),
 * degree <= max_degree, a random tree closed by 1-3 rings of size 5-6,
 * vertex labels from vlabel_probs[n_vlabels], edge labels from
 * elabel_probs[n_elabels]. */
int gm_gen_molecule(uint64_t seed, int32_t n_graphs, double mean_vertices, int32_t max_degree,
                    int32_t n_vlabels, const double *vlabel_probs, int32_t n_elabels,
                    const double *elabel_probs, gm_db **out);

/* Sub-database of the graphs listed in ids[0..n) (in that order). */
int gm_db_subset(const gm_db *db, const int32_t *ids, int32_t n, gm_db **out);

#ifdef __cplusplus
}
#endif

#endif /* GSPAN_AMD_H */
