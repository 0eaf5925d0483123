/* gspan_amd.h -- C-ABI of the MI355X (gfx950) gSpan single-large-graph mining path.
 *
 * The reference (haozixu/gSpan) has no FFI; its seams are C++ functions selected by
 * editing `mine` (gspan.cu:1118-1169).  Every entry point below replaces one of them:
 *
 *   gs_create / gs_destroy   init_gpu_context + init_cpu_context (gspan.cu:97-136): the
 *                            CSR graph (graph.h:51-65 layout) is copied to the device once
 *   gs_mine                  fsm(const CSRGraph&, int k, int min_support) (gspan.h:41,
 *                            gspan.cu:1171-1199) with mine() (gspan.cu:1118-1169); patterns
 *                            are delivered through a callback in the reference's order
 *                            (the lines gspan.cu:67-77 prints)
 *   gs_support               support_gpu (gspan.cu:200-289) / support_omp (gspan.cu:321-374)
 *   gs_extend                extend_omp (gspan.cu:674-780) for one pattern
 *   gs_is_canonical_batch    is_canonical (gspan.cu:586-606) for many codes at once
 *
 * Conventions: every function returns an int status (GS_OK = 0) and nothing throws
 * across the ABI; the message of the last failure is available from gs_last_error.
 * Pointers are plain host pointers; the inputs are only borrowed for the call.  One
 * context = one device + one HIP stream, used from one host thread at a time.
 * Pattern codes are arrays of 5 ints per edge: (from, to, from_label, edge_label,
 * to_label) exactly as PatternEdge (graph.h:28-47), with -1 where the reference stores no
 * label (from_label of every edge after the first, to_label of backward edges).
 * Embedding layers are arrays of 2 ints per row: (prev_idx, edge_id) as EmbeddingRel
 * (gspan.h:18-24), prev_idx = -1 in layer 0.
 */
#ifndef GSPAN_AMD_H
#define GSPAN_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define GS_OK 0
#define GS_ERR_INVALID 1  /* bad argument or malformed graph                         */
#define GS_ERR_HIP 2      /* a HIP runtime call failed                               */
#define GS_ERR_BUDGET 3   /* a single pattern's children exceed the device budget    */
#define GS_ERR_LIMIT 4    /* pattern larger than the kernels support (k > 7, ...)    */
#define GS_ERR_CALLBACK 5 /* the report callback asked to stop                       */
#define GS_ERR_INTERNAL 6 /* kernel-side consistency check failed                    */

#define GS_MODE_PLAIN 0   /* reference DFS order, one parent pattern per launch set   */
#define GS_MODE_BATCHED 1 /* level-synchronous: every pattern of a DFS level at once   */
#define GS_MODE_PREDICTED 2 /* batched; below the last level child buffers are sized by
                               the learned predictor and filled in one pass (exact two-pass
                               fallback on any misprediction)                          */

#define GS_MNI_GPU 0         /* support_gpu semantics (gspan.cu:165-190), the shipped fsm */
#define GS_MNI_OMP_LITERAL 1 /* support_omp semantics (gspan.cu:321-374)                  */

#define GS_MAX_PATTERN_EDGES 7

typedef struct gs_ctx gs_ctx;

typedef struct gs_opts {
    int mode;                        /* GS_MODE_*                                       */
    int mni_mode;                    /* GS_MNI_*                                        */
    int device;                      /* HIP device ordinal                              */
    int timing;                      /* 1: time every extension kernel with HIP events */
    unsigned long long budget_bytes; /* device bytes for mining buffers; 0 = 85% free  */
    long long chunk_rows;            /* parent rows per extension pass; 0 = 16M        */
    int two_pass;                    /* 1: always count, then fill (no single-pass last
                                        level)                                          */
    int set_kind;                    /* MNI image sets: GS_SETS_AUTO (byte per vertex
                                        where the budget allows, else bit per vertex),
                                        GS_SETS_BITS or GS_SETS_BYTES                  */
    int autotune;                    /* 1 (default): choose the extension kernels' tile
                                        size and workgroups per CU by timing launches  */
    int tile;                        /* 0, or a fixed tile size (256, 512, 1024) with
                                        full occupancy (turns the tuner off)           */
    int groups;                      /* last level: GS_GROUPS_AUTO counts by neighbour
                                        runs and inserts new vertices once per group of
                                        rows sharing a segment vertex where the graph's
                                        runs are long, else scans every neighbour of
                                        every row; GS_GROUPS_ON / GS_GROUPS_OFF force it */
    int two_edge;                    /* last level of two-edge children (k = 2): on a
                                        graph with no self-loops or repeated neighbours,
                                        GS_TWO_EDGE_AUTO takes child rows and image counts
                                        from per-vertex neighbour runs (batched and
                                        predicted methods); GS_TWO_EDGE_OFF extends the
                                        parent rows as every other level                */
} gs_opts;

#define GS_SETS_AUTO 0
#define GS_SETS_BITS 1
#define GS_SETS_BYTES 2
#define GS_GROUPS_AUTO 0
#define GS_GROUPS_ON 1
#define GS_GROUPS_OFF 2
#define GS_TWO_EDGE_AUTO 0
#define GS_TWO_EDGE_OFF 1

typedef struct gs_stats {
    long long extended;     /* child rows emitted by extension (gspan.cu:708,717,730);   */
                            /* a child with more than 2^31-1 rows adds 2^31 (lower bound) */
    long long candidates;   /* candidate codes examined (mine() calls)                   */
    long long canonical;    /* candidates passing the canonicality check                 */
    long long reported;     /* frequent patterns reported                                */
    long long batches;      /* extension batches run                                     */
    long long items;        /* (row, candidate entry) work items scanned                 */
    long long parent_rows;  /* parent rows extended                                      */
    long long fill_launches;
    long long count_launches;
    double alg_bytes_fill;  /* algorithmic bytes of all fill launches                    */
    double alg_bytes_count; /* algorithmic bytes of all count launches                   */
    double t_fill_ms;       /* summed fill-kernel time (timing = 1)                      */
    double t_count_ms;      /* summed count-kernel time (timing = 1)                     */
    double t_mine_ms;       /* wall time of the last gs_mine                             */
    unsigned long long peak_device_bytes;
    long long predicted_steps; /* steps run in one pass from predicted sizes               */
    long long fallbacks;       /* predicted steps redone exactly after a misprediction     */
    long long spilled_rows;    /* level rows kept in pinned host memory (budget exceeded)  */
    double t_sync_ms;          /* host time waiting for the device at synchronisations    */
    double t_order_ms;         /* host time putting the reports into the reference order  */
    double t_plan_ms;          /* host time turning candidate keys into children and sets */
    long long two_edge_levels; /* last levels of two-edge children taken from neighbour
                                  runs (their launches count as fill launches)           */
} gs_stats;

/* Receives one frequent pattern.  Return 0 to continue, non-zero to stop mining. */
typedef int (*gs_report_cb)(const int* code5, int len, int support, void* user);

/* Receives one child of gs_extend: its new edge (5 ints), its rows (2 ints each),
 * whether the child code is canonical, and its MNI support. */
typedef void (*gs_child_cb)(const int* edge5, const int* rows, long long count, int canonical,
                            int support, void* user);

/* Fills *opts with the defaults (batched, GPU MNI semantics, device 0). */
int gs_default_opts(gs_opts* opts);

/* Copies the CSR graph (graph.cc:178-214 layout: offsets[n+1], vlabels[n], edges as
 * nnz x {from, to, label}) to the device.  Each adjacency must be sorted by (to, label),
 * as CSRGraph(Graph&) (graph.cc:66-112) writes it. */
int gs_create(const int* offsets, const int* vlabels, const int* edges3, int n, int nnz,
              const gs_opts* opts, gs_ctx** out);

/* Mines every frequent connected pattern with at most k edges (fsm, gspan.cu:1171). */
int gs_mine(gs_ctx* ctx, int k, int min_support, gs_report_cb cb, void* user);

/* MNI support of one pattern given its layered rows. */
int gs_support(gs_ctx* ctx, const int* code5, int m, const int* const* layers,
               const long long* counts, int mni_mode, int* support_out);

/* One rightmost-path extension of one pattern.  Children arrive in the reference's
 * recursion order (backward: ascending (v_i, el); forward: descending v_i, then el, tl).
 * *total = child rows emitted. */
int gs_extend(gs_ctx* ctx, const int* code5, int m, const int* const* layers,
              const long long* counts, int mni_mode, gs_child_cb cb, void* user,
              long long* total);

/* Canonicality of ncodes codes; code i has lens[i] edges and starts at 5*sum(lens[:i]). */
int gs_is_canonical_batch(gs_ctx* ctx, const int* codes5, const int* lens, int ncodes,
                          int* flags);

/* gs_is_canonical_batch without any allocation: no hipMalloc, hipFree or pinned-host
 * allocation in the call.  No counterpart in the reference; for callers that cannot
 * allocate (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Inputs and outputs are the caller's host arrays, as for
 * gs_is_canonical_batch; the device staging lives in a static device array of 2.5 MiB that
 * each GPU the library runs on holds until the process ends.  At most
 * GS_NOALLOC_MAX_CODES codes per call (GS_ERR_LIMIT beyond).  Calls on one GPU share that
 * array: two calls that overlap on one GPU (from two threads, or two contexts) race. */
#define GS_NOALLOC_MAX_CODES 8192
int gs_is_canonical_batch_noalloc(gs_ctx* ctx, const int* codes5, const int* lens, int ncodes,
                                  int* flags);

/* Child-count predictor (MLP 32 -> 128 -> 128 -> 1, ReLU, bf16 weights, fp32 accumulation;
 * no counterpart in the reference, whose count pass is gspan.cu:737-755).
 * gs_predictor_log: while on, every exact count pass appends (features, log1p(count)) of its
 * candidate child keys to a device buffer of at most `capacity` samples (at most 2^24, as
 * for n of gs_mlp_forward / gs_mlp_gradients).
 * gs_predictor_train: full-batch Adam on the logged samples with the library's own forward
 * and backward kernels; *loss = final mean squared error (log space).
 * Weights are GS_MLP_PARAMS floats: W1[32][128], b1[128], W2[128][128], b2[128], W3[128], b3. */
#define GS_MLP_FEATURES 32
#define GS_MLP_PARAMS 20865
int gs_predictor_log(gs_ctx* ctx, int on, long long capacity);
int gs_predictor_samples(const gs_ctx* ctx, long long* count);
int gs_predictor_train(gs_ctx* ctx, int epochs, float lr, float* loss);
int gs_predictor_get_weights(const gs_ctx* ctx, float* weights);
int gs_predictor_set_weights(gs_ctx* ctx, const float* weights);
/* The network on caller data: y[n] for x[n][32]; gradients (GS_MLP_PARAMS floats) of the mean
 * squared error against t[n] at the current weights, without updating them. */
int gs_mlp_forward(gs_ctx* ctx, const float* x, long long n, float* y);
int gs_mlp_gradients(gs_ctx* ctx, const float* x, const float* t, long long n, float* grads,
                     float* loss);

/* prune_graph (graph.cc:216-266) of the context's graph on the device: vertices whose label
 * occurs fewer than min_sup times are dropped, the rest renumbered in ascending id order,
 * edges between kept vertices kept in CSR order.  Outputs must hold n+1, n and 3*nnz ints
 * (the input sizes); *n_out / *nnz_out receive the pruned sizes. */
int gs_prune_graph(gs_ctx* ctx, int min_sup, int* n_out, int* nnz_out, int* offsets_out,
                   int* vlabels_out, int* edges3_out);

/* The launch-shape tuner's measurements as a JSON object: per bucket (pass, single-pass
 * last level, parent pattern size) the chosen [tile, workgroups] and the time per work item
 * of every shape tried.  Writes at most cap bytes including the terminating NUL; *len
 * receives the full length.  No counterpart in the reference (its extension is OpenMP). */
int gs_tuner_report(const gs_ctx* ctx, char* buf, int cap, int* len);

/* Graph::from_txt (graph.cc:38-63) followed by CSRGraph(Graph&) (graph.cc:66-112), on the
 * GPU: text holds the bytes of a graph file (len bytes), one record per line ("t # id",
 * "v id label", "e u v label"; lines of other ops are skipped).  A line the reference would
 * read across a line boundary (wrong token count, non-integer argument, an op letter after
 * an unknown op) is rejected: GS_ERR_INVALID with its 0-based index in *bad_line.  On
 * success *out holds the CSR arrays on the device and *n_out / *nnz_out their sizes. */
typedef struct gs_text_graph gs_text_graph;
int gs_text_parse(const char* text, long long len, int device, gs_text_graph** out, int* n_out,
                  int* nnz_out, long long* bad_line);
/* Copies the parsed graph: offsets n+1, vlabels n, edges3 3*nnz ints (from, to, label). */
int gs_text_copy(const gs_text_graph* g, int* offsets, int* vlabels, int* edges3);
void gs_text_free(gs_text_graph* g);

/* Results of the last gs_mine, in report order, as flat arrays (for callers that want
 * them in bulk instead of through a callback; gs_mine may then be called with cb = NULL).
 * gs_result_count: number of patterns and total code edges.  gs_result_copy fills
 * codes5[5 * total_edges], lens[count] and supports[count]. */
int gs_result_count(const gs_ctx* ctx, long long* count, long long* total_edges);
int gs_result_copy(const gs_ctx* ctx, int* codes5, int* lens, int* supports);

int gs_get_stats(const gs_ctx* ctx, gs_stats* out);
int gs_reset_stats(gs_ctx* ctx);
const char* gs_last_error(const gs_ctx* ctx);
int gs_destroy(gs_ctx* ctx);

/* Library build string (architecture and version). */
const char* gs_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GSPAN_AMD_H */
