/*
 * vertexsurge.h — C ABI of the MI355X (gfx950) variable-length expansion path.
 *
 * This is the drop-in boundary for VertexSurge's VExpand and Mintersect
 * operators.  The reference runs both operators on the host over
 * memory-mapped graph columns reached through a global `Graph*`
 * (src/include/operator/physical_operator.hh:58, src/objects/storage/
 * graph_store.cpp:70).  Here the CSR is uploaded to HBM once into a graph
 * handle, and the operators' `run_internal` bodies call the entry points
 * below.  Plain pointers and sizes only: no C++ or torch types cross the ABI,
 * and no exception leaves it.  Every entry point returns a vs_status; on
 * failure `vs_last_error()` returns a message for the calling thread.
 *
 * Bit-matrix conventions (src/include/common/common.hh:219-282, MASK64):
 *   bit (r, c) = "vertex c is reachable in the set of source row r";
 *   within a 64-bit word the bit position is r % 64, least significant first.
 *
 * Handles are not thread-safe, matching the reference's single-caller
 * operators (physical_operator.cpp:62-101).
 */
#ifndef VERTEXSURGE_H_
#define VERTEXSURGE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
typedef enum {
  VS_OK = 0,
  VS_ERR_INVALID = 1,     /* bad argument (replaces the reference's assert/exit) */
  VS_ERR_NOMEM = 2,       /* device or host allocation failed */
  VS_ERR_HIP = 3,         /* a HIP runtime call failed */
  VS_ERR_CAPACITY = 4,    /* caller buffer too small; size reported back */
  VS_ERR_UNSUPPORTED = 5, /* option not implemented */
  VS_ERR_NODEVICE = 6,    /* no GPU visible */
  VS_ERR_IO = 7,          /* a file could not be opened, read or written */
  VS_ERR_INTERNAL = 8     /* unexpected failure inside the library (no C++ exception
                           * ever crosses this interface) */
} vs_status;

/* Message of the last failing call on this thread ("" if none). */
const char *vs_last_error(void);
/* Library version string. */
const char *vs_version(void);

/* ------------------------------------------------------------- enums */
/* ExpandPathType, src/include/operator/expand.hh:25-29. */
typedef enum {
  VS_PATH_ANY = 0,           /* walks of length in [k_min, k_max] */
  VS_PATH_SHORTEST = 1,      /* BFS distance in [k_min, k_max] */
  VS_PATH_FIND_SHORTEST = 2  /* per-row distance to a given target */
} vs_path_type;

/* Direction, src/include/common/common.hh:154-159 (Undecide is resolved by
 * the caller from the edge schema before reaching the ABI). */
typedef enum {
  VS_DIR_FORWARD = 0,  /* follow CSR edges u -> adj[u] */
  VS_DIR_BACKWARD = 1, /* follow reversed edges */
  VS_DIR_BOTH = 2      /* union of both (== FORWARD on a symmetric graph) */
} vs_direction;

/* Per-hop expansion scheduling.  The reference's ExpandStrategy::Dense runs
 * every edge every hop (expand.cpp:263-311); Adaptive switches on a density
 * ratio (expand.cpp:346-414).  On the GPU every choice below yields the same
 * bits; only the time differs.  VS_SCHED_LEARNED also picks the storage tile
 * width of a query (or of a batch group) when none is forced (tile_words 0
 * in the call and in the graph tuning): the analogue of the reference's
 * stack-height choice decide_k0 (expand.cpp:416-449). */
typedef enum {
  VS_SCHED_HEURISTIC = 0, /* push/pull by frontier-edge density */
  VS_SCHED_PUSH = 1,      /* always sparse scatter */
  VS_SCHED_PULL = 2,      /* always dense gather */
  VS_SCHED_LEARNED = 3    /* MLP attached to the graph; heuristic fallback */
} vs_sched;

/* vs_vexpand_params.flags */
#define VS_VX_TIMING 0x1u /* record per-hop / per-launch HIP event timings */
/* Expand the rows in the caller's order.  By default a call with k_max >= 3
 * (>= 2 for vs_find_shortest) expands them sorted, within each tile of
 * 64 * tile_words rows, by the source's 2-hop volume and puts its result
 * matrices back into the caller's order before returning; results are the
 * same either way. */
#define VS_VX_KEEP_ORDER 0x2u
/* vs_batch_params.flags: pack rows in input order (default: by source degree;
 * outputs are in input order either way) */
#define VS_BATCH_KEEP_ORDER 0x2u

/* Result export layouts. */
typedef enum {
  VS_LAYOUT_NATIVE = 0,   /* [tile][vertex][tile_words] uint64 words */
  VS_LAYOUT_DVSM = 1,     /* reference DenseVarStackMatrix, 1024-row stacks
                             (src/include/storage/bitmap.hh:329-425) */
  VS_LAYOUT_ROWMAJOR = 2  /* row r = ceil(V/64) words, bit c%64 of word c/64 */
} vs_layout;

/* --------------------------------------------------------- host CSR data */
/* A host-side CSR (uint64 offsets[V+1], uint32 adj[E], each list sorted
 * ascending, no duplicates) produced by the synthetic generators below.
 * Mirrors the *_csr_src_dst / *_index columns of the reference
 * (src/include/meta/graph_loader.hh:163-182). */
typedef struct vs_csr_s *vs_csr_t;

/* Symmetrized R-MAT: 2^scale vertices, edge_factor * 2^scale generated pairs,
 * quadrant probabilities a, b, c (d = 1-a-b-c); self-loops and duplicates
 * removed.  Counter-based: the same (scale, ef, a, b, c, seed) gives the same
 * CSR on every machine and thread count. */
vs_status vs_rmat_csr(uint32_t scale, uint32_t edge_factor, double a, double b,
                      double c, uint64_t seed, vs_csr_t *out);
/* Symmetrized Erdos-Renyi G(V, M): M uniform pairs, same cleanup. */
vs_status vs_er_csr(uint32_t num_vertices, uint64_t num_pairs, uint64_t seed,
                    vs_csr_t *out);
/* CSR from an explicit edge list (src[i] -> dst[i]); optionally symmetrized;
 * self-loops kept unless drop_self_loops; duplicates removed. */
vs_status vs_csr_from_edges(uint32_t num_vertices, const uint32_t *src,
                            const uint32_t *dst, uint64_t n, int symmetrize,
                            int drop_self_loops, vs_csr_t *out);
vs_status vs_csr_info(vs_csr_t csr, uint32_t *num_vertices, uint64_t *num_edges,
                      const uint64_t **offsets, const uint32_t **adj);
vs_status vs_csr_free(vs_csr_t csr);

/* S distinct vertices drawn uniformly from [0, V), sorted ascending
 * (the contract of uniform_random_numbers, src/include/utils/utils.hh:74-99;
 * the draw sequence is this library's own counter-based one). */
vs_status vs_sample_sources(uint32_t num_vertices, uint32_t S, uint64_t seed,
                            uint32_t *out);
/* T random 2-hop paths a-b-c: a uniform over non-isolated vertices, b uniform
 * in N(a), c uniform in N(b).  out holds 3*T vertex ids, row-major. */
vs_status vs_sample_paths3(const uint64_t *offsets, const uint32_t *adj,
                           uint32_t num_vertices, uint64_t T, uint64_t seed,
                           uint32_t *out);

/* ----------------------------------------------------------------- device */
vs_status vs_device_count(int *count);
vs_status vs_set_device(int device);
vs_status vs_device_synchronize(void);
/* Raw device buffers for callers that keep inputs resident in HBM. */
vs_status vs_dev_alloc(size_t bytes, void **dptr);
vs_status vs_dev_free(void *dptr);
vs_status vs_memcpy_h2d(void *dst, const void *src, size_t bytes);
vs_status vs_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel (k_copy_words), a
 * known-size reference for HBM traffic counters.  bytes % 8 == 0. */
vs_status vs_device_copy(void *dst, const void *src, size_t bytes);

/* ------------------------------------------------------------------ graph */
/* Replaces the global PhysicalOperator::graph_store and its
 * get_edges_refs/get_csr_edge_refs views (graph_store.cpp:75-189): the CSR is
 * copied to HBM once and every later call reads it there. */
typedef struct vs_graph_s *vs_graph_t;

#define VS_GRAPH_SYMMETRIC 0x1u /* adj is symmetric: reverse CSR == CSR */

vs_status vs_graph_create(const uint64_t *offsets, const uint32_t *adj,
                          uint32_t num_vertices, uint64_t num_edges,
                          uint32_t flags, vs_graph_t *out);
vs_status vs_graph_destroy(vs_graph_t g);
vs_status vs_graph_info(vs_graph_t g, uint32_t *num_vertices,
                        uint64_t *num_edges, uint64_t *device_bytes);
/* Release cached device workspace (the per-graph buffer pool). */
vs_status vs_graph_trim(vs_graph_t g);

/* Push/pull tuning attached to a graph (written by vs_autotune_vexpand or
 * by the caller). */
typedef struct {
  uint32_t tile_words;  /* storage tile width in 64-bit words: 1..64, pow2 */
  uint32_t push_alpha;  /* push while frontier-edges * alpha < E */
  uint32_t sparse_pct;  /* pull reads the frontier's word masks (sparse-row
                           gathers) when its nonzero words, weighted by degree,
                           are below this percentage of all; 0 = 35 */
  uint32_t pull_hubs;   /* vertices whose frontier rows pull gathers first, with
                           cached loads (the most-listed ones); 0 = 131072 */
  uint32_t wpush_pml;   /* word push on pull hops that gather whole rows: the
                           frontier words cheapest to scatter (fewest out-edges of
                           the rows holding bits there), taken cheapest first
                           while their out-edges sum to at most this many per
                           mille of E, are scattered instead of gathered;
                           0 = 40; above 1000 = off */
} vs_tuning;
vs_status vs_graph_set_tuning(vs_graph_t g, const vs_tuning *t);
vs_status vs_graph_get_tuning(vs_graph_t g, vs_tuning *t);

/* ------------------------------------------------------------ graph store */
/* A graph directory in the reference's storage format, read the way
 * Graph::load (graph_store.cpp:47-73) and Graph::get_csr_edge_refs
 * (graph_store.cpp:75-105) read it: schema.json (GraphSchema,
 * schema.hh:10-141) plus one DiskArray file per variable (a 64-byte header
 * {int64 count, int64 element size, reserved}, then the elements;
 * disk_array.hh:267-354).  The relation's source-side CSR,
 * "<src>-<type>-<dst>_csr_src_dst[_index]" ("->" when directed;
 * schema.cpp:11-49), is mapped read-only; lists that are not ascending are
 * sorted and repeated ids dropped (the importer keeps multi-edges and
 * removes only adjacent duplicates, graph_loader.hh:71-77), which changes
 * no VExpand or Mintersect result. */
typedef struct vs_store_s *vs_store_t;
typedef struct {
  uint32_t num_vertices;        /* the vertex type's statistics.count */
  uint64_t stored_entries;      /* CSR entries in the files (edge statistics.count) */
  uint64_t num_edges;           /* entries kept */
  uint64_t duplicates;          /* repeated ids dropped */
  uint64_t unsorted_lists;      /* lists that were not ascending in the files */
  uint64_t self_loops;          /* entries u -> u in the files (kept) */
  int32_t directed;             /* EdgeSchema::directed */
  int32_t allow_multiple_edge;  /* EdgeSchema::allow_multiple_edge */
} vs_store_info;
/* edge: the relation's type name ("knows") or a simple name
 * "Person_knows_Person" (GraphSchema::get_edge_schema, schema.cpp:222-232);
 * NULL or "" picks the schema's only relation of a vertex type to itself.
 * Relations between two vertex types return VS_ERR_UNSUPPORTED (a device
 * graph is one vertex set; VExpand needs a self relation for k > 1,
 * vexpand.cpp:88-91). */
vs_status vs_store_open(const char *dir, const char *edge, vs_store_t *out);
vs_status vs_store_get_info(vs_store_t s, vs_store_info *info);
/* The CSR (num_vertices + 1 offsets, num_edges ids); valid until close. */
vs_status vs_store_csr(vs_store_t s, const uint64_t **offsets, const uint32_t **adj);
/* vs_graph_create over the store's CSR, VS_GRAPH_SYMMETRIC for an undirected
 * relation (stored with both directions, graph_concepts.hh:96-103). */
vs_status vs_store_graph(vs_store_t s, vs_graph_t *out);
vs_status vs_store_close(vs_store_t s);
/* Writes a CSR as a store the reference can load (gen_and_write_one_edge,
 * graph_loader.hh:152-276, for one relation of vertex_type to itself): the
 * forward CSR files, the reverse CSR files when directed, and schema.json.
 * With flat_partitions = P > 0 it also writes, for each stored direction, the
 * flat edge files the reference's importers list (store_info flats and
 * partitioned_flats {src_dst, hilbert}, partition_counts {P, P};
 * graph_loader.cpp:216-223): "<name>_src_dst" (edges by source, then
 * destination), "<name>_hilbert" (Peano-Hilbert order of (src, dst) on the
 * next power-of-two grid, peano.hh:73-89), and per destination partition i
 * (partition_by_dst, graph_loader.hh:111-139) "<name>_src_dst_<i>" and
 * "<name>_hilbert_<i>" -- the files the reference's Dense Hilbert and
 * BlockHilbert expansions and its ColMain expansion open (expand.cpp:197,
 * 221, 266; schema.cpp:382-389).  Backward files hold the pairs reversed.
 * With P = 0 no flat files are written or listed.  An undirected relation's
 * adj must hold both directions of every edge.  dir must exist. */
vs_status vs_store_write(const char *dir, const char *graph_name, const char *vertex_type,
                         const char *edge_type, int32_t directed, const uint64_t *offsets,
                         const uint32_t *adj, uint32_t num_vertices, uint64_t num_edges,
                         uint32_t flat_partitions);

/* ---------------------------------------------------------------- VExpand */
/* Parameters of VExpand (big_operators.hh:101-129, vexpand.cpp:7-101). */
typedef struct {
  uint32_t k_min;
  uint32_t k_max;
  int32_t path_type;     /* vs_path_type: ANY or SHORTEST */
  int32_t direction;     /* vs_direction */
  int32_t with_distance; /* export every layer L_i, i in [k_min, k_actual_max] */
  int32_t sched;         /* vs_sched */
  uint32_t tile_words;   /* 0 = graph tuning */
  uint32_t flags;        /* VS_VX_* */
} vs_vexpand_params;

#define VS_MAX_HOPS 64

typedef struct {
  uint32_t k_actual_max; /* VExpand::k_actual_max (vexpand.cpp:222-233) */
  uint32_t tile_words;
  uint32_t num_tiles;
  uint32_t reserved;
  uint64_t te;           /* traversed edges: sum_i sum_u deg(u)*popcount(F_i[.,u]) */
  uint64_t bytes_alg;    /* compulsory HBM bytes (words touched once), counted
                          * over the row order the call expands (see
                          * VS_VX_KEEP_ORDER) */
  uint64_t visit_count;  /* ExpandStatus::visit_count after the last hop */
  double ms_total;       /* device time of init + hops + finalize */
  double ms_hops;        /* device time of the hop loop only */
  /* per hop i = 1..k_actual_max at index i-1 */
  double ms_hop[VS_MAX_HOPS];
  int32_t mode_hop[VS_MAX_HOPS];   /* 1 = push, 2 = pull (last tile's) */
  uint64_t te_hop[VS_MAX_HOPS];
  uint64_t bytes_hop[VS_MAX_HOPS];
  uint64_t frontier_bits[VS_MAX_HOPS]; /* popcount of F_i (input of hop i) */
  uint64_t new_bits[VS_MAX_HOPS];      /* popcount of L_i */
  /* device time of the dominant kernel launches (pull / push) */
  double ms_pull;
  double ms_push;
  uint64_t n_pull;
  uint64_t n_push;
  uint64_t bytes_pull; /* algorithmic bytes of the hops run as pull */
  uint64_t bytes_push;
  /* scheduling features of hop i's frontier F_i (index i-1), summed over tiles */
  uint64_t frontier_act[VS_MAX_HOPS]; /* vertices with a nonzero frontier row */
  uint64_t frontier_deg[VS_MAX_HOPS]; /* their out-degree sum */
  double vis_frac[VS_MAX_HOPS];       /* nonzero visited words / all words */
  /* query features (filled when rows are packed, the learned scheduler picks
   * the tile width, or VS_VX_TIMING is set; else 0): out-degree sum of the
   * sources and their 2-hop volume sum_{s} sum_{u in N(s)} deg(u) */
  uint64_t src_degsum;
  uint64_t src_vol2;
  /* 1 when the learned scheduler picked tile_words */
  uint32_t tile_words_learned;
  uint32_t reserved2;
  /* words scattered by the word push at hop i (index i-1), summed over tiles */
  uint32_t wpush_words[VS_MAX_HOPS];
} vs_vexpand_stats;

typedef struct vs_result_s *vs_result_t;

/* Multi-source VExpand over the bit-matrix representation
 * (VExpand::before_execution/after_child/run_internal,
 *  vexpand.cpp:166-483; Expand<DenseVarStackMatrix>::expand_dense,
 *  expand.cpp:210-344).  S == 0 returns an empty result with VS_OK (the
 * reference warns and skips, vexpand.cpp:170-173).  Sources may repeat.
 * `stats` may be NULL. */
vs_status vs_vexpand(vs_graph_t g, const uint32_t *sources, uint32_t S,
                     const vs_vexpand_params *p, vs_result_t *out,
                     vs_vexpand_stats *stats);

/* FindShortest (vexpand.cpp:16-25, 186-205, 237-276): dist[r] = BFS distance
 * from sources[r] to targets[r], -1 if not found within k_max hops
 * (K_MAX_FIND_SHORTEST = 50 in the reference, expand.hh:37). */
vs_status vs_find_shortest(vs_graph_t g, const uint32_t *sources,
                           const uint32_t *targets, uint32_t S,
                           int32_t direction, uint32_t k_max, int32_t *dist,
                           vs_vexpand_stats *stats);

/* Result = VarRowMatrixTable (one merged matrix) or MultipleMatrixTable (one
 * matrix per distance) (vexpand.cpp:298-378, matrix_table.hh:44-149). */
vs_status vs_result_info(vs_result_t r, uint32_t *rows, uint32_t *cols,
                         uint32_t *num_matrices, uint32_t *k_actual_max);
/* distance tag of matrix i (with_distance), or UINT32_MAX for a merged one */
vs_status vs_result_distance(vs_result_t r, uint32_t i, uint32_t *distance);
/* tuple_count(): popcount of matrix i (bit_matrix_table.cpp:112) */
vs_status vs_result_count(vs_result_t r, uint32_t i, uint64_t *count);
/* per-source reachable counts: popcount of each row of matrix i */
vs_status vs_result_row_counts(vs_result_t r, uint32_t i, uint64_t *counts);
/* per-row 64-bit digest of matrix i (order-independent hash of set bits) */
vs_status vs_result_row_digests(vs_result_t r, uint32_t i, uint64_t *digests);
vs_status vs_result_export_size(vs_result_t r, uint32_t i, int32_t layout,
                                size_t *bytes);
vs_status vs_result_export(vs_result_t r, uint32_t i, int32_t layout,
                           void *host_buf, size_t bytes);
/* single bit (DenseVarStackMatrix::get, bitmap.cpp:953-956) */
vs_status vs_result_get(vs_result_t r, uint32_t i, uint32_t row, uint32_t col,
                        int *bit);
/* The result as the flat table of VExpand's CSR strategy (merge_to_Flat,
 * vexpand.cpp:380-446): one (start, target[, distance]) tuple per pair in
 * the union of the result's matrices; start = the row's source vertex,
 * target = the column's vertex.  With per-hop layers (with_distance) the
 * distance is that of the first layer holding the pair (Shortest layers are
 * disjoint; for Any it is the reference's dedup keeping the smallest
 * distance).  Order: by row, then target column ascending (the reference's
 * order depends on its threads).  Two-phase: with start == NULL only *n is
 * set; dist may be NULL, and needs a with_distance result. */
vs_status vs_result_flat(vs_result_t r, uint32_t *start, uint32_t *target, int32_t *dist,
                         uint64_t cap, uint64_t *n);
vs_status vs_result_free(vs_result_t r);

/* ------------------------------------------------------- batched VExpand */
typedef struct {
  const uint32_t *sources;
  uint32_t S;
  uint32_t k_min;
  uint32_t k_max;
  int32_t path_type; /* ANY or SHORTEST */
} vs_query;

typedef struct {
  uint32_t tile_words;  /* 0 = graph tuning */
  int32_t sched;        /* vs_sched for the per-(hop, tile) choice */
  uint32_t use_graph;   /* capture the launch sequence in a hipGraph */
  uint32_t replays;     /* >1: replay the captured graph this many times (timing) */
  size_t mem_budget;    /* device bytes for tile state; 0 = 70% of free */
  int32_t direction;
  uint32_t flags;
} vs_batch_params;

typedef struct {
  uint64_t te;
  double ms_total;      /* device time of all waves of one replay: the mean
                           over replays 2..n when replays > 1 */
  double ms_capture;
  uint32_t num_tiles;
  uint32_t num_waves;   /* groups of tiles resident at once */
  uint32_t launches;    /* kernel nodes per replay */
  uint32_t reserved;
  uint64_t push_decisions;
  uint64_t pull_decisions;
  /* with VS_VX_TIMING in flags (runs without graph capture): device time of
   * the persistent expand launches over dense rows (ms_expand, n_expand) and
   * over sparse rows read through word masks (ms_expand_sparse,
   * n_expand_sparse; one of each per hop), and the compulsory bytes of all
   * hops (bytes_alg) and of the sparse ones (bytes_sparse) */
  double ms_expand;
  uint64_t n_expand;
  uint64_t bytes_alg;
  /* host clock per replay (reset the wave counter and outputs, launch the
   * graph for every wave, synchronize), same averaging as ms_total; 0 when
   * the launch sequence is not captured */
  double ms_wall;
  double ms_expand_sparse;
  uint64_t n_expand_sparse;
  uint64_t bytes_sparse;
  uint64_t wpush_words; /* words scattered by the word push, summed over hops
                           and tiles */
} vs_batch_stats;

/* Runs nq independent VExpand queries packed into shared bit-matrix tiles.
 * Outputs, concatenated over queries in order (row offset of query q =
 * sum of S of the queries before it): per-row reachable counts and per-row
 * digests of the merged result.  Either output may be NULL. */
vs_status vs_vexpand_batch(vs_graph_t g, const vs_query *queries, uint32_t nq,
                           const vs_batch_params *p, uint64_t *row_counts,
                           uint64_t *row_digests, vs_batch_stats *stats);

/* The same batch, returning one result per query instead of per-row counts:
 * results[q] is what vs_vexpand returns for query q (VExpand::run_internal's
 * merge_to_MT, vexpand.cpp:298-354: a VarRowMatrixTable whose rows are the
 * query's sources in its order, the merged matrix of its hop range, and
 * k_actual_max as VExpand::after_child sets it, vexpand.cpp:222-233), to
 * be freed with vs_result_free.  Rows are packed per query: each query's rows
 * start at a 64-row word boundary and are sorted by 2-hop volume within each
 * tile of its own, so a query owns whole words; a one-query batch expands
 * straight into its result matrix, a larger batch copies every query's words
 * into its matrix at the end of each wave.  All results stay in device memory
 * (sum over queries of ceil(S_q / (64 tile_words)) * V * tile_words * 8
 * bytes): VS_ERR_NOMEM when they do not fit next to the batch's working
 * tiles.  On failure no result is returned. */
vs_status vs_vexpand_batch_results(vs_graph_t g, const vs_query *queries, uint32_t nq,
                                   const vs_batch_params *p, vs_result_t *results,
                                   vs_batch_stats *stats);

/* ------------------------------------------------------------- Mintersect */
typedef struct {
  double ms;            /* device time of the call's kernels (split to last copy) */
  uint64_t total;       /* sum of counts */
  uint64_t bytes_alg;
} vs_mintersect_stats;

/* Multi-way intersection (Mintersect::run_internal, mintersect.cpp:48-345):
 * for tuple t, the set  N_0(x_t0) ∩ N_1(x_t1) ∩ ... ∩ N_{arity-1}(x_t,arity-1)
 * where N_d is the adjacency of deps[d] (forward).  tuples is T*arity ids,
 * row-major.  counts[t] = |set|; when lists != NULL the sets are written in
 * ascending id order at lists[offsets[t] .. offsets[t+1]) (offsets has T+1
 * entries) and list_cap must be >= the total, else VS_ERR_CAPACITY with
 * *total set.  A vertex id >= V of its dependency gives the empty set (the
 * reference's missing-column case, mintersect.cpp:143-155). */
vs_status vs_mintersect(const vs_graph_t *deps, uint32_t arity,
                        const uint32_t *tuples, uint64_t T, uint32_t *counts,
                        uint64_t *offsets, uint32_t *lists, uint64_t list_cap,
                        uint64_t *total, vs_mintersect_stats *stats);
/* Same with every array already in device memory (from vs_dev_alloc).
 * With d_lists, one intersection pass stages every tuple's survivors
 * (4 bytes per entry of its shortest list, pooled device memory) and copies
 * them into place; on VS_ERR_CAPACITY d_counts, d_offsets and *total are
 * already written.  When the staging would exceed VS_MI_STAGE_MAX_MB
 * megabytes (environment, read per call; default 4096, 0 = never) the call
 * counts, scans and intersects a second time to fill instead. */
vs_status vs_mintersect_device(const vs_graph_t *deps, uint32_t arity,
                               const uint32_t *d_tuples, uint64_t T,
                               uint32_t *d_counts, uint64_t *d_offsets,
                               uint32_t *d_lists, uint64_t list_cap,
                               uint64_t *total, vs_mintersect_stats *stats);

/* vs_mintersect_device without any allocation.  No counterpart in the
 * reference: this entry point is for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  It always takes two passes (count -> scan -> fill) and
 * keeps its working memory (tuple buckets, counters, scan partials) in one
 * fixed device array of about 72 MiB per GPU, held from the library's first
 * use of it on that GPU until the process ends; it calls no allocator and
 * records no memset.  Limits and contract:
 *   T <= VS_NOALLOC_MAX_TUPLES, else VS_ERR_INVALID;
 *   d_offsets is required (T + 1 entries: the exclusive scan of d_counts and
 *   the total); d_lists may be NULL (counts and offsets only), and when the
 *   total exceeds list_cap the call returns VS_ERR_CAPACITY with d_counts,
 *   d_offsets and *total written;
 *   every call on one GPU shares that GPU's array: calls that overlap on one
 *   GPU (two host threads, or graphs with different streams) race, and the
 *   caller must serialise them.  The array keeps the last call's contents;
 *   each call clears what it needs. */
#define VS_NOALLOC_MAX_TUPLES 4718592u
vs_status vs_mintersect_device_noalloc(const vs_graph_t *deps, uint32_t arity,
                                       const uint32_t *d_tuples, uint64_t T,
                                       uint32_t *d_counts, uint64_t *d_offsets,
                                       uint32_t *d_lists, uint64_t list_cap,
                                       uint64_t *total);

/* ------------------------------------------------- learned scheduler (MLP) */
/* 3-layer FP32 MLP: in -> 64 -> 64 -> out, ReLU, trained and evaluated by
 * this library's own HIP kernels. */
typedef struct vs_sched_s *vs_sched_t;

#define VS_MLP_IN 8
#define VS_MLP_HIDDEN 64
#define VS_MLP_OUT 8

vs_status vs_sched_create(uint64_t seed, vs_sched_t *out);
vs_status vs_sched_free(vs_sched_t s);
vs_status vs_sched_num_params(size_t *n);
vs_status vs_sched_get_params(vs_sched_t s, float *params, size_t n);
vs_status vs_sched_set_params(vs_sched_t s, const float *params, size_t n);
/* Loss and gradient of masked MSE over a batch (one fwd + bwd on the GPU). */
vs_status vs_sched_loss_grad(vs_sched_t s, const float *X, const float *Y,
                             const float *M, uint32_t n, float *loss,
                             float *grad);
/* Full-batch Adam training for `epochs` steps; loss_curve may be NULL. */
vs_status vs_sched_train(vs_sched_t s, const float *X, const float *Y,
                         const float *M, uint32_t n, uint32_t epochs, float lr,
                         float *loss_curve);
vs_status vs_sched_predict(vs_sched_t s, const float *X, uint32_t n, float *Y);
/* The feature domain in which the scheduler may overrule the density
 * heuristic: n feature rows (VS_MLP_IN floats each; normally its training
 * samples, at most VS_SCHED_MAX_DOMAIN) and a radius.  A decision whose
 * feature vector lies farther than `radius` (Euclidean) from every row is the
 * heuristic's.  n == 0 removes the guard. */
#define VS_SCHED_MAX_DOMAIN 1024
vs_status vs_sched_set_domain(vs_sched_t s, const float *rows, uint32_t n, float radius);
/* Attach (or detach with NULL) a scheduler used by VS_SCHED_LEARNED.  The
 * graph holds its own reference: vs_sched_free may be called while the
 * scheduler is attached, and the scheduler lives until the last graph using
 * it detaches it or is destroyed. */
vs_status vs_graph_set_scheduler(vs_graph_t g, vs_sched_t s);

/* ------------------------------------------------ bit-matrix tables (ColMain) */
/* A device-resident DenseColMainMatrix with its table maps
 * (bitmap.hh:262-326, bitmap.cpp:580-595; ColMainMatrixTable,
 * matrix_table.hh:85-125): `rows` x `cols` bits, column c at words
 * [c*pitch, (c+1)*pitch), pitch = align8(ceil(rows/64)), row r at bit r%64 of
 * word r/64 -- the reference word layout, so export is a plain copy.  Bits of
 * rows >= `rows` are zero.  The optional row map gives the vertex id of each
 * row, the optional column map the vertex id of each column
 * (MatrixTable::row_map/col_map; a column lookup by id takes the first column
 * holding that id, as set_col_hash_map's insert does, bit_matrix_table.cpp:26-37).
 * The matrix lives on the device of `g` and is ordered on its stream. */
typedef struct vs_colmat_s *vs_colmat_t;

/* words == NULL: all zeros; else cols*pitch words in the layout above
 * (bits of rows >= `rows` are cleared on upload). */
vs_status vs_colmat_create(vs_graph_t g, uint32_t rows, uint32_t cols,
                           const uint64_t *words, vs_colmat_t *out);
/* row_map: `rows` ids or NULL (no map).  col_map: n_col_map <= cols ids or
 * NULL (n_col_map == 0 also means no map); columns past n_col_map are not
 * reachable by id (the reference test MintersectFlatMatMat maps 8 of 10
 * columns, mintersect.cpp:433-457). */
vs_status vs_colmat_set_maps(vs_colmat_t m, const uint32_t *row_map,
                             const uint32_t *col_map, uint32_t n_col_map);
/* *n_col_map = 0 when there is no column map */
vs_status vs_colmat_info(vs_colmat_t m, uint32_t *rows, uint32_t *cols,
                         uint64_t *pitch_words, int *has_row_map,
                         uint32_t *n_col_map);
/* row_map / col_map copies (NULL to skip; rows / n_col_map ids) */
vs_status vs_colmat_get_maps(vs_colmat_t m, uint32_t *row_map, uint32_t *col_map);
vs_status vs_colmat_export(vs_colmat_t m, uint64_t *words, size_t bytes);
vs_status vs_colmat_get(vs_colmat_t m, uint32_t row, uint32_t col, int *bit);
vs_status vs_colmat_free(vs_colmat_t m);

/* TableConvert VarRowMatrix -> ColMainMatrix (TableConvert::run_internal,
 * big_operators2.cpp:16-86) of matrix i of a VExpand result.  The maps come
 * along (take_maps_from): the result's row map is its source list
 * (vexpand.cpp:351-353).  transpose == 0: rows = sources, cols = vertices,
 * column v = the rows that reach v.  transpose != 0: rows = vertices,
 * cols = sources (row and column maps swapped), column r = the vertex set of
 * source row r. */
vs_status vs_table_convert(vs_result_t r, uint32_t i, int transpose,
                           vs_colmat_t *out);

/* MatrixTableSelect (big_operators2.cpp:138-213): a new single-matrix result
 * whose column k is column col_ids[k] of matrix i (a vertex id of the
 * result), with col_ids as its column map and the same row map. */
vs_status vs_result_select_cols(vs_result_t r, uint32_t i, const uint32_t *col_ids,
                                uint32_t n, vs_result_t *out);
/* column map of a result made by vs_result_select_cols (cols ids); a VExpand
 * result has none (VS_ERR_INVALID) */
vs_status vs_result_col_map(vs_result_t r, uint32_t *col_ids);

typedef enum { VS_ORDER_NONE = 0, VS_ORDER_ASC = 1, VS_ORDER_DESC = 2 } vs_order;

/* MatrixColumnCount (matrix_column_count.cpp:19-87): for every column, its id
 * (through the column map if any) and its number of set bits.  VS_ORDER_NONE
 * gives columns in index order; ASC / DESC sort by (count, id) as the
 * reference's pair sort does.  ids and counts hold `cols` entries. */
vs_status vs_result_col_counts(vs_result_t r, uint32_t i, int32_t order,
                               uint32_t *ids, int32_t *counts);
vs_status vs_colmat_col_counts(vs_colmat_t m, int32_t order, uint32_t *ids,
                               int32_t *counts);
/* MatrixRowOrSum (matrix_row_or_sum.cpp:14-55): ids of the columns with any
 * set bit, in column index order.  *n gets the number; ids (cap entries) may
 * be NULL for a size query; VS_ERR_CAPACITY if cap < *n. */
vs_status vs_result_row_or(vs_result_t r, uint32_t i, uint32_t *ids, uint32_t cap,
                           uint32_t *n);
vs_status vs_colmat_row_or(vs_colmat_t m, uint32_t *ids, uint32_t cap, uint32_t *n);

/* Mintersect with ColMainMatrix dependencies (Mintersect::run_internal
 * flat-base branch, mintersect.cpp:205-243, with tackle_di_for_base_vid's
 * ColMain case, mintersect.cpp:139-156): for base tuple t,
 *   I_t = AND_d column(deps[d], tuples[t*nd + d])
 * where the column is found through the dependency's column map (or is the
 * id itself when it has none), and a missing column makes I_t empty.  All
 * dependencies have the same row count.  counts[t] = |I_t|; with lists, the
 * rows of I_t in ascending row order, through deps[0]'s row map when it has
 * one (mintersect.cpp:115-122, 228-241), at lists[offsets[t] ..
 * offsets[t+1]).  Same two-phase contract as vs_mintersect. */
vs_status vs_mintersect_colmat(const vs_colmat_t *deps, uint32_t nd,
                               const uint32_t *tuples, uint64_t T, uint32_t *counts,
                               uint64_t *offsets, uint32_t *lists, uint64_t list_cap,
                               uint64_t *total, vs_mintersect_stats *stats);

/* Mintersect with a VarRowMatrixTable base (Mintersect::run_internal's
 * VarRow-base branch, mintersect.cpp:244-333): every set bit of matrix i of a
 * VExpand result is one base tuple (start vertex = the row's source, target
 * vertex = the column's vertex; the table's iterators,
 * bit_matrix_table.cpp:165-188), and dependency d intersects the column bound
 * to base variable var_idx[d] (0 = start, 1 = target).  The dependencies are
 * ColMain tables (mat_deps, as vs_mintersect_colmat: column lookup through
 * the column map, missing column = empty, new vertices through deps[0]'s row
 * map) or edge relations (graph_deps, as vs_mintersect: the bound vertex's
 * adjacency; the reference's flat dependencies), exactly one of the two
 * non-NULL.  Output: one (start, target, new vertex) tuple per element,
 * ordered by base tuple (row, then target ascending) and new vertex
 * ascending (the reference's order follows its threads).  The base tuples
 * and intersections stay on the device.  With start == NULL only *n is set
 * (the only_count total, mintersect.cpp:336-340); VS_ERR_CAPACITY when
 * cap < *n. */
vs_status vs_mintersect_varrow(vs_result_t base, uint32_t i, const vs_colmat_t *mat_deps,
                               const vs_graph_t *graph_deps, uint32_t nd, const uint32_t *var_idx,
                               uint32_t *start, uint32_t *target, uint32_t *nv, uint64_t cap,
                               uint64_t *n, vs_mintersect_stats *stats);

/* -------------------------------------------------------------- autotuner */
/* Tunes the batched engine on the caller's queries and stores the result in
 * the graph's tuning: coordinate descent over tile_words (the widest the
 * rows allow and two narrower), push_alpha (4, 16, 64), sparse_pct (35, 50,
 * 65), pull_hubs (32768, 131072, 524288, at most V) and wpush_pml (40, 200,
 * 2000 = off), each candidate timed
 * by vs_vexpand_batch (device time, the mean of two replays of the captured
 * graph after a warm one).  The starting tuning (the graph's, or
 * p->tile_words when set) is measured first (*start_ms) and a candidate is
 * kept only when at least 2% faster, so *best_ms <= *start_ms.  Scheduling knobs never
 * change a result bit.  best, best_ms and start_ms may be NULL. */
vs_status vs_autotune_batch(vs_graph_t g, const vs_query *queries, uint32_t nq,
                            const vs_batch_params *p, vs_tuning *best, double *best_ms,
                            double *start_ms);
/* vs_autotune_batch over one query: the given sources, hops 1..k_max. */
vs_status vs_autotune_vexpand(vs_graph_t g, const uint32_t *sources, uint32_t S,
                              uint32_t k_max, int32_t path_type,
                              vs_tuning *best, double *best_ms);

#ifdef __cplusplus
}
#endif

#endif /* VERTEXSURGE_H_ */
