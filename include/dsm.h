/*
 * dsm.h - C ABI of the MI355X continuous subgraph matching engine.
 *
 * Drop-in for the GPU matching path of the reference (qxk152/DSM, CDSM): same
 * input formats, same update rule, same match semantics (labelled, injective,
 * non-induced embeddings, counted per inserted edge), on one GPU.
 *
 * Conventions: the caller owns every host buffer it passes; the engine owns all
 * device memory. Every entry point returns an int error code (DSM_OK = 0,
 * negative on failure) and never exits the process; dsm_last_error() gives the
 * message. One engine is used by one host thread at a time.
 */
#ifndef DSM_H
#define DSM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ error codes */
#define DSM_OK 0
#define DSM_EINVAL -1    /* bad argument (out-of-range vertex, bad query ...) */
#define DSM_ENOMEM -2    /* device or host allocation failed */
#define DSM_EDEVICE -3   /* HIP runtime error */
#define DSM_EFORMAT -4   /* malformed DSM text file */
#define DSM_EIO -5       /* file could not be opened */
#define DSM_ELIMIT -6    /* a hard limit was exceeded (query size, labels ...) */
#define DSM_ETRUNC -7    /* tuple buffer too small (counts are still exact) */
#define DSM_EADMIT -8    /* served request rejected by admission control */

/* Hard limits (the reference's are MAX_VQ = 16 query vertices, defs.h:18). */
#define DSM_MAX_QUERY_VERTICES 16
#define DSM_MAX_QUERIES 1024
#define DSM_MAX_LABEL 254 /* query labels must be <= this */

typedef struct dsm_engine dsm_engine;

/* One update record: the reference's `change` (cpuGraph.h:13-18), read from a
 * line `a|d src dst label` (io.cu:232-242). */
typedef struct {
  uint32_t add;   /* 1 for 'a' (insert), 0 for 'd' (delete) */
  uint32_t src;
  uint32_t dst;
  uint32_t label; /* labelChange: an insert relabels BOTH endpoints with it */
} dsm_change;

/* Initial data graph: undirected edge list (each edge once, as the e-lines of
 * the data file) plus vertex labels. Duplicate edges collapse. */
typedef struct {
  uint32_t num_vertices;
  const uint32_t *labels;
  uint64_t num_edges;
  const uint32_t *edge_src;
  const uint32_t *edge_dst;
} dsm_graph_desc;

/* A connected query graph of at most DSM_MAX_QUERY_VERTICES vertices. */
typedef struct {
  uint32_t num_vertices;
  const uint32_t *labels;
  uint32_t num_edges;
  const uint32_t *edge_src;
  const uint32_t *edge_dst;
} dsm_query_desc;

typedef struct {
  int device;              /* HIP device ordinal (the reference's --gpu, main.cu:43) */
  int relabel;             /* 1: reference update rule (an insert relabels both
                              endpoints, cpuGraph.cpp:564); 0: label field ignored */
  int count_negatives;     /* 1: delete records are matched in the pre-batch graph */
  uint64_t tuple_capacity; /* tuples kept per batch (0: counts only) */
  int tuples_in_host_memory; /* 1: tuple buffer in pinned host memory (spill) */
  uint32_t waves_per_block;  /* DFS launch shape; 0 = default */
  uint32_t blocks_per_cu;    /* 0 = as many as occupancy allows */
  uint32_t donate_min;       /* smallest source range split off to an idle wave (0 = default) */
  uint32_t queue_capacity;   /* donated-task queue slots (0 = default) */
  uint32_t watchdog_ms;      /* an idle DFS wave that waits longer than this for
                                work aborts the round with DSM_EDEVICE (0 = 60 s) */
  uint32_t lane_work;        /* one level above the leaf: longest list a lane scans
                                alone before the wave takes it over (0 = 16) */
  int leaf_sets;             /* 1: one level above the leaf, candidates count their
                                leaf images against a leaf set shared by the parent;
                                0: one descent per candidate */
  uint32_t donate_every;     /* scan steps between a busy wave's checks for idle waves (0 = 8) */
  uint32_t donate_min_n2;    /* smallest range one level above the leaf that is split
                                off to an idle wave (0 = 64) */
  uint32_t poll_max;         /* a wave waiting for donated work sleeps at most this many
                                8-unit s_sleeps between polls (1..16; 0 = 16) */
  int leaf_segments;         /* 1: with leaf sets, the lanes' leaf lists are scanned as
                                one segmented array; 0: each lane scans its own */
  int graph_rounds;          /* 1: each DFS round (control reset, seeding, DFS, status
                                copy) is one HIP graph launch, built once and updated
                                per round; 0: separate launches */
  int overlap_apply;         /* 1: the update is applied to a second copy of the graph
                                while a round matches on the first: the next batch's
                                during pipelined rounds (dsm_process_batches, the served
                                stream), the batch's own during its negatives round
                                (dsm_apply_batch). Doubles the graph's device memory;
                                0: apply after the round */
  uint32_t leaf_budget;      /* one level above the leaf: inline leaf-count work (list
                                entries) of a 64-candidate window before a candidate
                                descends so that its scan can be shared with idle waves
                                (0 = 2048; 0xFFFFFFFF = never) */
  int cluster_filter;        /* 1: after every update the candidates are refined by the
                                cluster encoding of the reference's clusterFilter
                                (filter.cu:1525-1654: spanning tree, vertex cover,
                                clusters, encode/merge/combine passes); the seeding and
                                the DFS then use the refined sets. Counts are the same
                                either way; 0: LDF/NLF candidates only */
  int bit_probe_last;        /* 1: the DFS probes a source entry's candidate bit after its
                                membership tests instead of before them (searched by
                                dsm_autotune) */
} dsm_options;

/* Per-batch statistics of the last dsm_apply_batch + dsm_match_batch. */
typedef struct {
  uint64_t seeds_pos, seeds_neg;   /* DFS roots seeded from insert / delete records */
  uint64_t tasks_donated;          /* sub-ranges handed to idle waves */
  uint64_t tuples;                 /* tuples found (may exceed the capacity) */
  uint32_t touched_vertices;       /* vertices whose adjacency changed */
  uint32_t relabeled_vertices;     /* vertices whose label changed */
  uint32_t dirty_groups;           /* vertices re-filtered (endpoints, relabelled vertices
                                      and their neighbours); summed over the batches of
                                      dsm_process_batches, like filter_entries */
  uint32_t relocated_vertices;     /* adjacency lists moved to a larger slot */
  float ms_apply, ms_filter, ms_seed, ms_dfs_pos, ms_dfs_neg; /* device time */
  uint64_t pool_entries_used, pool_entries_capacity;
  /* DFS work: source-list entries tested (each a 4 B list read plus an 8 B
   * candidate-bitmap word) and 4 B reads of the membership binary searches */
  uint64_t dfs_entries, dfs_search_loads;
  uint32_t dfs_launches;
  /* 4 B x degree of every backward list of every partial match the DFS
   * expanded: the bytes merge intersections over whole lists would read */
  uint64_t dfs_merge_bytes;
  /* balance of the persistent DFS grid, summed over the launches: time the
   * waves spent on tasks (ms, all waves), waves launched, and per launch the
   * busiest block's task time over the mean block's */
  double dfs_busy_ms;
  uint64_t dfs_waves;
  double dfs_block_imbalance;
  /* where the waves' idle time lies, summed over the launches as fractions of
   * waves x (last task end - first task start): before a wave's first task,
   * and after its last one (the tail) */
  double dfs_head_idle;
  double dfs_tail_idle;
  float ms_refine;                 /* device time of the cluster-encoding refinement */
  uint64_t filter_entries;         /* list entries of the vertices whose candidacy was
                                      re-evaluated (the refilter's work measure) */
  uint64_t filter_bits;            /* candidate-bitmap bits those vertices were given (one
                                      per signature of their label) */
} dsm_stats;

/* Outcome of dsm_autotune: the launch parameters kept and the DFS time of the
 * sample round before and after. */
typedef struct {
  uint32_t waves_per_block, blocks_per_cu, donate_min, lane_work;
  int leaf_sets, leaf_segments;
  uint32_t donate_every;
  uint32_t trials;        /* configurations timed */
  float ms_initial, ms_best;
  uint32_t leaf_budget;
  int bit_probe_last;
  uint32_t order_changed; /* queries whose matching-order rule the tuner changed */
} dsm_tune_report;

/* ------------------------------------------------------------ engine life */

void dsm_options_default(dsm_options *opt);

/* Builds the device-resident graph, candidate bitmaps and per-query matching
 * plans. Replaces the reference's set-up in main.cu:59-145: readGraphToCPU-
 * Dynamic (io.cu:122), copyMeta (io.cu:332), decycle (decycle.cpp:6),
 * allocateMemGPU/copyGraphToGPU (io.cu:254-302) and the first clusterFilter
 * (filter.cu:1525). */
int dsm_create(const dsm_graph_desc *graph, const dsm_query_desc *queries, uint32_t num_queries,
               const dsm_options *opt, dsm_engine **out);

void dsm_destroy(dsm_engine *e);

/* Last error message of the engine (or, when e is NULL, of the last failed
 * call without an engine on this thread: dsm_create, dsm_intersect_sorted).
 * Never NULL. */
const char *dsm_last_error(const dsm_engine *e);

/* --------------------------------------------------------- per-batch path */

/* Applies one update batch. With count_negatives, delete records are first
 * matched against the pre-batch graph. Then the batch is applied in place on
 * the device under the reference's rule (updateHostGraph, cpuGraph.cpp:521-607:
 * symmetrise, stable (src,dst) order, set insert / erase, relabel on insert)
 * and the candidate bitmaps are refreshed for the vertices it affects
 * (replaces copyGraphToGPU io.cu:279, clusterFilter filter.cu:1525 and compact
 * filter.cu:1692). Records with an endpoint >= num_vertices fail the whole
 * batch with DSM_EINVAL before anything changes. */
int dsm_apply_batch(dsm_engine *e, const dsm_change *changes, uint32_t num_changes);

/* Matches the insert records of the last applied batch against the post-batch
 * graph and returns, per query, the positive count (sum over insert records of
 * the embeddings that map a query edge onto the record's edge) and the
 * negative count (the same over delete records, in the pre-batch graph).
 * Replaces the seeding loop of main.cu:259-329 and parallelMatch
 * (gpu_match.cu:1121), whose result it prints as `res: N` (:1338, :1429). */
int dsm_match_batch(dsm_engine *e, uint64_t *pos_counts, uint64_t *neg_counts);

/* dsm_apply_batch followed by dsm_match_batch. */
int dsm_process_batch(dsm_engine *e, const dsm_change *changes, uint32_t num_changes,
                      uint64_t *pos_counts, uint64_t *neg_counts);

/* A sequence of batches, each with the semantics of dsm_process_batch and in
 * order, pipelined: each DFS launch matches the positives of batch b together
 * with the negatives of batch b+1, which are evaluated on the same graph
 * state. The batches' host-side deltas are computed by a worker thread that
 * the call starts and joins, at most two batches ahead of the launches, while
 * the calling thread waits for the GPU; the records must stay unchanged until
 * the call returns. Batch b is changes[batch_offsets[b] .. batch_offsets[b+1]); the counts go
 * to pos_counts / neg_counts [num_batches][num_queries]; done_s (optional)
 * receives the steady-clock second at which each batch's counts were ready.
 * With overlap_apply, batch b+1's update is applied to the second copy of
 * the graph during that launch. Counts only (tuple_capacity must be 0). The
 * statistics cover the whole call. DSM_EINVAL (bad offsets or endpoints) is
 * returned before anything runs; after any other error the graph may be
 * partly updated and the engine should be destroyed. No counterpart in the
 * reference, which matches one batch per loop iteration (main.cu:152-348). */
int dsm_process_batches(dsm_engine *e, const dsm_change *changes, const uint32_t *batch_offsets,
                        uint32_t num_batches, uint64_t *pos_counts, uint64_t *neg_counts, double *done_s);

/* Tuples found by the last dsm_match_batch (needs tuple_capacity > 0). Each
 * tuple is (query, sign, f(u_0) .. f(u_{n-1})) in DSM_MAX_QUERY_VERTICES + 2
 * uint32 words (sign 1 = positive, 0 = negative; unused words are
 * 0xFFFFFFFF), in no particular order. Copies at most cap tuples; *num_out
 * receives the number found. Returns DSM_ETRUNC if more were found than the
 * engine's tuple_capacity could hold. */
int dsm_get_tuples(dsm_engine *e, uint32_t *out, uint64_t cap, uint64_t *num_out);

int dsm_get_stats(const dsm_engine *e, dsm_stats *out);

/* ------------------------------------------------------ state inspection */

/* Copies the device graph out as a sorted CSR: offsets[num_vertices + 1],
 * neighbours[cap], labels[num_vertices]. *num_entries receives the number of
 * neighbour entries (2x the undirected edges); returns DSM_ETRUNC if cap is too
 * small. The reference's copyGraphToCPU (io.cu:304) is its counterpart. */
int dsm_get_graph(dsm_engine *e, uint64_t *offsets, uint32_t *neighbours, uint64_t cap,
                  uint32_t *labels, uint64_t *num_entries);

/* Candidate bitmap of query vertex u of query q: ceil(num_vertices/64) words,
 * bit v set iff v passes the LDF/NLF filter (filter.cu:394-566) for u. The
 * reference keeps the equivalent in d_compact_encodings_ (main.cu:141-142). */
int dsm_get_candidates(dsm_engine *e, uint32_t q, uint32_t u, uint64_t *words);

/* Candidates the seeding and the DFS use for query vertex u of query q, in
 * the layout of dsm_get_candidates: with cluster_filter, the vertex's bits at
 * its final position of the cluster encoding (enc_pos_of_u_, filter.cu:
 * 1589-1609, read out as compact does, filter.cu:1692-1750); without it, the
 * LDF/NLF set. */
int dsm_get_refined_candidates(dsm_engine *e, uint32_t q, uint32_t u, uint64_t *words);

/* The cluster plan the refinement uses for a query (pure host logic): the
 * spanning tree's vertex cover (decycle.cpp:6-135, filter.cu:23-116), the
 * clusters with their merges and combines (filter.cu:131-392), the passes in
 * the order encode() runs them (filter.cu:936-1177) and each vertex's final
 * position (filter.cu:1589-1609), as words: n, ncover, cover[ncover],
 * nclusters, per cluster (valid, size, members[size]), final cluster,
 * position[n], npasses, per pass (kind 0 encode / 1 merge / 2 combine, cluster
 * or combine index). Edge lines in file order. Returns the number of words
 * (written up to cap), or 0 for an invalid query or one past the reference's
 * limits (150 clusters, 32 small clusters per combine). No counterpart in
 * the reference, which keeps the plan in host structures (structure.cuh). */
uint32_t dsm_cluster_plan(uint32_t num_vertices, uint32_t num_edges, const uint32_t *edge_src,
                          const uint32_t *edge_dst, uint32_t *out, uint32_t cap);

/* DFS task time per query (ms, summed over the waves, so it can exceed the
 * wall time) since the engine was created or this was last called; the call
 * resets it. No counterpart in the reference, which runs one query per
 * process; it tells which standing queries cost the most. */
int dsm_get_query_time(dsm_engine *e, double *ms, uint32_t num_queries);

/* Sorted-set intersection a ∩ b on device `device`, with the membership test
 * the DFS uses for every backward list (a cooperative wave search for the
 * window the 64 lanes' values span, then LDS-staged binary searches): the
 * reference's intersect (gpu_match.cu:21-94), which binary-searches each
 * element of the first set in the second. a and b are host arrays, strictly
 * ascending, with no value 0xFFFFFFFF (else DSM_EINVAL); out (room for
 * min(na, nb) values) receives the intersection in ascending order and
 * *num_out its size. Allocates and frees its device buffers; errors are read
 * with dsm_last_error(NULL). */
int dsm_intersect_sorted(int device, const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *out,
                         uint32_t *num_out);

/* dsm_intersect_sorted without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Its device
 * memory is a fixed array of 4,194,304 words per GPU, held until the process
 * ends; a call needs na + nb + 66 words per 64 entries of a (rounded up) +
 * min(na, nb) + 1 of them, and a larger one fails with DSM_ELIMIT before any
 * device work. Calls that overlap on one GPU share that array and race. */
int dsm_intersect_sorted_noalloc(int device, const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb,
                                 uint32_t *out, uint32_t *num_out);

/* Number of vertices and queries of the engine. */
uint32_t dsm_num_vertices(const dsm_engine *e);
uint32_t dsm_num_queries(const dsm_engine *e);

/* Chooses the DFS launch parameters for this engine's queries and graph:
 * waves per block, blocks per CU, the donation threshold, the lane work limit
 * and leaf sets on or off, by coordinate descent over timed DFS rounds that
 * match the sample's records in the current graph (read-only, as negatives
 * are), and then each query's matching-order rule from the queries' own task
 * times under every rule. Every configuration must give the same counts (else
 * DSM_EDEVICE).
 * Keeps the fastest; changes no graph state but overwrites the counts and
 * statistics of the last batch. Call between batches. No counterpart in the
 * reference, whose launch shapes are constants (gpu_match.cu:1121-1431). */
int dsm_autotune(dsm_engine *e, const dsm_change *sample, uint32_t num_changes, dsm_tune_report *report);

/* --------------------------------------------------------- served stream */

/* A server runs a worker thread over one engine and serves update batches
 * ("requests") submitted from any thread, in submission order, each with the
 * semantics of dsm_process_batch. It keeps a rolling pipeline: the positives
 * of the request in flight (already applied) and the negatives of the next
 * one are matched in one DFS launch over the same graph, the next request's
 * delta is computed on the host while that launch runs, and each request's
 * counts are published as soon as its positives are done. A request is
 * admitted when it fits the memory budget beside the request in flight
 * (dsm_admit); one that does not fit even alone is rejected with DSM_EADMIT.
 * No counterpart in the reference, which runs one batch at a time
 * (main.cu:152-348). While a server exists, the engine must not be used
 * directly. */
typedef struct dsm_server dsm_server;

typedef struct {
  uint64_t budget_bytes;          /* admission budget (0: free HBM minus the headroom) */
  double hbm_headroom;            /* fraction of total HBM kept free (0: 0.30) */
} dsm_server_options;

typedef struct {
  uint64_t id;
  int status;                     /* DSM_OK, DSM_EADMIT or another error */
  double t_submit, t_start, t_done; /* steady-clock seconds: submitted, admitted, counts published */
  uint32_t queue_depth;           /* requests still pending when it was admitted */
  uint32_t fused;                 /* 1: its negatives shared a launch with the previous request's positives */
  uint32_t num_changes;
} dsm_request_info;

int dsm_server_create(dsm_engine *e, const dsm_server_options *opt, dsm_server **out);
/* Copies the changes and queues the request; *id identifies it. An
 * out-of-range endpoint rejects it here with DSM_EINVAL. */
int dsm_server_submit(dsm_server *s, const dsm_change *changes, uint32_t num_changes, uint64_t *id);
/* Blocks until request id is done; fills its info and per-query counts (either
 * may be NULL) and returns its status. Each id can be waited for once. */
int dsm_server_wait(dsm_server *s, uint64_t id, dsm_request_info *info, uint64_t *pos_counts,
                    uint64_t *neg_counts);
const char *dsm_server_error(const dsm_server *s);
/* Serves what is pending, then stops the worker. */
void dsm_server_destroy(dsm_server *s);

/* Admission rule of the server, pure host logic: 1 = admit now (it fits
 * beside the in-flight bytes), 0 = wait for the request in flight,
 * DSM_EADMIT = larger than the whole budget. */
int dsm_admit(uint64_t request_bytes, uint64_t in_flight_bytes, uint64_t budget);
/* Estimated intermediate device memory of a request of num_changes records
 * (seed table, device records, update arrays). */
uint64_t dsm_estimate_request_bytes(const dsm_engine *e, uint32_t num_changes);

/* ------------------------------------------------------- DSM text files */

typedef struct dsm_file dsm_file;

/* Parses a DSM text file (io.cu:122-252): `t n m`, n `v id label deg` lines in
 * id order, m `e a b` lines (each may carry an edge label as a third number,
 * which the reference's format lacks; it is kept, see dsm_file_edge_labels,
 * and the matching ignores it), then for data files `-` and batches `b bid k`
 * with k `a|d src dst label` lines, bid = 0,1,... in order. A v-line degree
 * that differs from the vertex's e-line count is DSM_EFORMAT (the reference
 * would build a corrupt CSR, io.cu:174). A data file that ends after its e
 * lines, without `-`, is the reference's static format (readGraphToCPU2,
 * io.cu:27-120): a graph with no batches. */
int dsm_file_read(const char *path, int is_query, dsm_file **out);
void dsm_file_free(dsm_file *f);
/* Views into the parsed file; valid until dsm_file_free. */
int dsm_file_graph(const dsm_file *f, dsm_graph_desc *graph);
int dsm_file_query(const dsm_file *f, dsm_query_desc *query);
/* Edge labels of the e lines, in file order (0 on lines without one): returns
 * 1 if any e line carried one, else 0. No counterpart in the reference, whose
 * reader has no edge labels; the engine's matching ignores them. */
int dsm_file_edge_labels(const dsm_file *f, const uint32_t **labels);
uint32_t dsm_file_num_batches(const dsm_file *f);
int dsm_file_batch(const dsm_file *f, uint32_t b, const dsm_change **changes, uint32_t *num_changes);
const char *dsm_file_error(void);

#ifdef __cplusplus
}
#endif
#endif /* DSM_H */
