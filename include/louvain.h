/*
 * louvain.h — C-ABI of the MI355X-native Louvain engine (gfx950, HIP).
 *
 * The reference (olearczuk/gpu-louvain) has no plugin or FFI layer: its
 * phases are C++ functions over structs of raw device pointers, linked into
 * one executable.  Each entry point below replaces one of them; the comment
 * names the reference interface (file:line) it stands in for.  Plain
 * pointers and sizes only; every call returns an lv_status and never exits
 * the process (the reference's HANDLE_ERROR exits, utils.cuh:92-104).
 *
 * Threading: one lv_graph owns one HIP stream and is not thread-safe;
 * separate handles may be used from separate threads.
 *
 * Numerics: M, vertex strengths, community totals, per-candidate weight
 * sums, merged aggregation weights and modularity are FP64; gains are fp32
 * in the reference's operation order (modularity_optimisation.cu:96-102);
 * aggregated weights are stored fp32 (the reference's CSR type).  Weights
 * are rounded to a grid first (lv_weight_grid) on which every FP64 sum of
 * weights is exact, so results do not depend on the order of atomics and
 * repeat bit for bit.
 */
#ifndef LOUVAIN_H
#define LOUVAIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t lv_status;
#define LV_OK 0
#define LV_ERR_ARG (-1)          /* bad argument / malformed input        */
#define LV_ERR_OOM (-2)          /* device or host allocation failed      */
#define LV_ERR_HIP (-3)          /* HIP runtime / kernel launch error     */
#define LV_ERR_IO (-4)           /* file cannot be read or written        */
#define LV_ERR_STATE (-5)        /* call not valid in the handle's state  */
#define LV_ERR_NODEVICE (-6)     /* no HIP device visible                 */

typedef struct lv_graph lv_graph;

/* Creation options.  Zero-initialise, then set what you need. */
typedef struct {
    int32_t device;          /* HIP device ordinal (default 0)                       */
    int32_t timing;          /* 1: record per-kernel HIP events (lv_kernel_time)      */
    int32_t heavy_wg_max;    /* degree bound of the one-workgroup LDS tiers (0 = 4096) */
    int32_t chunk_edges;     /* edges per chunk of the chunked tier (0 = 1024)        */
    int32_t agg_lds_max;     /* summed-degree bound of LDS aggregation (0 = auto)     */
    int32_t part_target;     /* pairs per merge partition of the chunked tier (0=1024) */
    int32_t merge_limit;     /* pairs an LDS merge table may take (0 = 1536)          */
    int32_t q_incremental;   /* track modularity incrementally after a sweep whose
                                movers' rows hold at most E/q_incremental entries
                                (0 = 2, -1 = never)                                    */
    int32_t xcd_rows;        /* XCD-local passes over heavy rows when rows are sorted:
                                0 = modularity pass on levels with V >= 2^21 whose
                                move kernels read original columns (col_order off),
                                1 = it on every level, 2 = it and the chunked tier,
                                -1 = off                                           */
    int32_t agg_mode;        /* aggregation: 0 = radix sort of (row, column) community
                                keys, 1 = per-community hash merges                 */
    int32_t compact;         /* 0: every level is stored without its vertices of
                                degree 0 (results and every returned array stay in
                                the reference's numbering; batches always store
                                every vertex); -1: store every vertex               */
    int32_t candidate_bounds; /* with weights >= 0 the move kernels may read a
                                candidate community's record only when a bound on
                                its gain shows it can change the decision (same
                                decisions): 0 = on levels of >= 2^20 stored
                                vertices, 1 = on every level, -1 = off           */
    int32_t heavy_lds;        /* rows above the workgroup tiers are first tried in
                                one LDS table when their last evaluation saw at
                                most 4096 distinct communities (same decisions):
                                0 = on, -1 = off, n > 0 = on, but a row gives up
                                at n communities (exercises the fallback)       */
    int32_t exact_sums;       /* 0: every weight is first rounded to the grid of
                                lv_weight_grid, so that every FP64 sum of weights
                                is exact and no result depends on the order of
                                atomics (weights already on the grid, e.g.
                                integers, are unchanged); -1: weights are used as
                                given (sums of weights off the grid may then
                                round differently from run to run)             */
    /* Degree cuts and shapes of the tiers inside degree groups 5 and 6
     * (autotuning; results do not depend on them): */
    int32_t wave_split;       /* group 5 (85..319): degrees up to this run in
                                256-slot wave tables, above in 448-slot ones
                                (85..255; 0 = 128)                              */
    int32_t wg_bound_a;       /* group 6 one-workgroup tiers: degree bounds of   */
    int32_t wg_bound_b;       /* the first three (ascending, 320..heavy_wg_max;  */
    int32_t wg_bound_c;       /* 0 = 512, 1024, 2048); the fourth ends at
                                heavy_wg_max                                    */
    int32_t wg_lanes[4];      /* lanes per row of the four one-workgroup tiers:
                                128, 256, 512 or 1024 (0 = 128, 256, 256, 1024) */
    int32_t table_pct;        /* LDS hash slots of those tiers per degree bound,
                                percent (120..200; 0 = 150, 125 for the third)  */
    int32_t sub_lanes[3];     /* lanes per vertex of the sub-wave buckets of
                                degrees 5..8, 9..16 and 17..32: a power of two
                                from 4 to the bucket's bound; each lane then
                                holds bound / lanes row entries (0 = 4, 4, 4)    */
    int32_t col_order;        /* the move kernels read row columns renumbered in
                                degree-bucket order, with a copy of the
                                communities in that order (same decisions;
                                4 bytes per entry and 8 per vertex more): 0 = on
                                single-graph levels of >= 2^20 vertices that all
                                have entries, 1 = on every such level, -1 = off */
    int32_t hinted;           /* rows of the one-workgroup tiers whose last
                                evaluation saw at most 256 distinct communities
                                are first tried one wave per row in a 512-slot
                                table (plain sweeps; same decisions): 0 = on,
                                -1 = off                                        */
    int32_t nmv_pull;         /* the neighbour-move counts of tracked sweeps:
                                0 = from the rows' side when the last sweep's
                                movers hold >= E/16 row entries, else from the
                                movers' side; 1 = always the rows' side, -1 =
                                always the movers' (same counts)               */
} lv_options;

/* Run statistics (lv_run).  Times are device times from HIP events. */
typedef struct {
    int32_t levels;          /* phases run                                          */
    int32_t sweeps;          /* sweeps over all levels                              */
    int32_t final_V;         /* vertices of the last level                          */
    int32_t level0_sweeps;   /* sweeps of the first phase                           */
    int64_t final_E;         /* CSR entries of the last level                       */
    int64_t edges_swept;     /* sum over sweeps of the level's CSR entries          */
    double modularity;       /* final Q, FP64                                       */
    double algo_ms;          /* device time of the whole run (after upload)         */
    double sweep_ms;         /* device time inside sweeps (moves+apply+Q)           */
    double level0_sweep_ms;  /* device time of the first phase's sweeps             */
    double aggregate_ms;     /* device time inside aggregation                      */
    double memory_ms;        /* upload time at lv_create                            */
    int64_t evaluated;       /* vertex evaluations over all sweeps (pruned: frontier) */
    int32_t full_sweeps;     /* sweeps that evaluated every vertex                  */
    int32_t pad_;
    double reserved[4];
} lv_stats;

/* Per-kernel timing (options.timing = 1): kernel ids below. */
typedef struct {
    int64_t launches;
    double total_ms;         /* summed HIP-event durations                          */
    double bytes;            /* algorithmic bytes moved by those launches           */
} lv_kernel_time;

#define LV_K_MOVE_SUB 0      /* sub-wavefront move kernels (degree <= 32)           */
#define LV_K_MOVE_WAVE 1     /* one-wavefront move kernels (33..319)                */
#define LV_K_MOVE_WG 2       /* one-workgroup LDS hash move kernel (320..2048)      */
#define LV_K_MOVE_GLOBAL 3   /* chunked tier: chunk aggregation + partition merge   */
#define LV_K_APPLY 4         /* apply moves + incremental totals                    */
#define LV_K_MODULARITY 5    /* fused modularity reduction                          */
#define LV_K_AGGREGATE 6     /* whole aggregation (phase 2)                         */
#define LV_K_COUNT 7

/* Single-kernel timing ids (lv_kernel_timing_detail): one HIP kernel each. */
#define LV_KD_MOVE_BUCKET0 0     /* +b, b = 0..7: move kernel of execution bucket b
                                    (0-3 sub-wave, 4-5 wave, 6-7 workgroup)          */
#define LV_KD_CHUNK_AGGREGATE 8  /* chunked tier stage A (bytes: 12 per heavy edge)      */
#define LV_KD_CHUNK_MERGE 9      /* chunked tier stage B (bytes: 12 per heavy edge, the
                                    pairs it reads; a row has at most one per edge)    */
#define LV_KD_CHUNK_FINAL 10     /* chunked tier decisions (bytes not counted)           */
#define LV_KD_MODULARITY_ROWS 11 /* modularity row pass (12 B per edge it reads + 16 V)  */
#define LV_KD_FRONTIER 12        /* frontier model inference (43 bytes per vertex)       */
#define LV_KD_MODULARITY_XCD 13  /* heavy-row modularity pass by column block (12 B/edge)  */
#define LV_KD_MOVE_BUCKET8 14    /* +b-8, b = 8..10: move kernel of execution bucket b  */
#define LV_KD_HEAVY_LDS 17       /* LDS attempt of the chunked tier's rows (bytes not
                                    counted)                                           */
#define LV_KD_HINTED 18          /* small-table attempt of the one-workgroup tiers' rows
                                    (lv_options.hinted; bytes not counted)             */
#define LV_KD_COUNT 19

/* ---- graph handle (utils.cu:64-106 copyStructures + initM mo.cu:420-422) ---- */
/* Uploads the CSR (rowptr[V+1], col[E], w[E]; host pointers) and sets up the
 * singleton state comm = iota, size = 1 (utils.cu:84-89).  M = (sum w)/2 of
 * the weights after the grid rounding (lv_options.exact_sums).
 * Precondition (the reference's reader guarantees it, utils.cu:32-45): the
 * CSR is symmetric, every entry (v, u, w) matched by an entry (u, v, w) with
 * the same weight bits (duplicates counted as a multiset).  An asymmetric CSR
 * or a non-finite weight is rejected with LV_ERR_ARG. */
lv_status lv_create(const int32_t *rowptr, const int32_t *col, const float *w, int32_t V,
                    int64_t E, const lv_options *opt, lv_graph **out);
/* utils.cu:108-141 deleteStructures. */
void lv_destroy(lv_graph *g);
/* Back to the uploaded level-0 graph with the singleton state (no re-upload;
 * when level 0 is current only the community state is rewritten). */
lv_status lv_reset(lv_graph *g);

/* ---- phases ---- */
/* One modularity-optimisation phase (modularity_optimisation.cu:284-363,
 * declared modularity_optimisation.cuh:28): sweeps while gain >= min_gain
 * and gain > 0 (the second condition only matters for min_gain <= 0, where
 * the reference can loop forever); *changed = any sweep had gain > 0. */
lv_status lv_optimise(lv_graph *g, float min_gain, int32_t *changed);
/* Community aggregation (community_aggregation.cu:248-347, declared
 * community_aggregation.cuh:11-12).  Rows come out in ascending new id. */
lv_status lv_aggregate(lv_graph *g);
/* Modularity of the current state, FP64 (modularity_optimisation.cu:395-401,
 * declared modularity_optimisation.cuh:30). */
lv_status lv_modularity(lv_graph *g, double *q);
/* Whole multi-level run (gpulouvain.cu:30-44). */
lv_status lv_run(lv_graph *g, float min_gain, lv_stats *stats);

/* Modularity without any allocation.  No counterpart in the reference: a
 * variant of lv_modularity for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Every array is a caller-owned DEVICE pointer on `device`:
 * rowptr[V+1], col[E], w[E], comm[V] (community of each vertex, in [0, V)),
 * and q (one double, written by the last kernel).  Work is enqueued on
 * `stream` (a hipStream_t; null = the default stream) and not waited for.
 * Q = own/2M - sum_c S_c^2/(2M)^2 in FP64 with 2M = the sum of w (the weights
 * as given; lv_modularity uses the handle's rounded weights and run M).  A
 * row, column or community outside its range makes *q NaN (nothing outside
 * the arrays is touched).  Scratch is one static device array per GPU (about
 * 8.4 MB for LV_NOALLOC_MAX_V), held until the process ends: V above
 * LV_NOALLOC_MAX_V is rejected with LV_ERR_ARG, and calls that overlap on one
 * GPU race on it. */
#define LV_NOALLOC_MAX_V (1 << 20)
lv_status lv_modularity_noalloc(int32_t device, void *stream, const int32_t *rowptr,
                                const int32_t *col, const float *w, const int32_t *comm,
                                int32_t V, int64_t E, double *q);

/* ---- state access (for -v output and parity checks) ---- */
/* Final/current original-vertex -> community map in the current level's ids
 * (data printed by printOriginalToCommunity, modularity_optimisation.cu:403-418). */
lv_status lv_assignment(lv_graph *g, int32_t *orig2comm, int32_t *level_V);
lv_status lv_level_size(lv_graph *g, int32_t *V, int64_t *E);
lv_status lv_level_graph(lv_graph *g, int32_t *rowptr, int32_t *col, float *w);
/* Level dump (no reference counterpart; for resuming a run and for feeding a
 * level to the CPU oracle): <prefix>.lvcsr holds the current level's CSR
 * (lv_write_csr), <prefix>.o2c the map of the original vertices into it
 * (8-byte magic "LVO2C01\0", int64 original V, int64 level V, fp64 total
 * weight M of the run, int32 map[original V]), and with with_mtx
 * <prefix>.mtx the CSR as MatrixMarket.  lv_create_from_dump builds a handle
 * whose level 0 is the dumped level, with the run's M, and whose
 * lv_assignment maps the original vertices (levels count from the dump). */
lv_status lv_dump_level(lv_graph *g, const char *prefix, int32_t with_mtx);
lv_status lv_create_from_dump(const char *prefix, const lv_options *opt, lv_graph **out);
lv_status lv_get_communities(lv_graph *g, int32_t *comm);
/* Replace the current level's community state (sizes and totals are rebuilt). */
lv_status lv_set_communities(lv_graph *g, const int32_t *comm);
/* Move decisions of degree group `group` (0..6, bounds {0,4,8,16,32,84,319,inf},
 * modularity_optimisation.cuh:7-8) from the current state, without applying
 * them (computeMove, modularity_optimisation.cu:111-228).  new_comm[V]
 * receives the decision for the group's vertices and comm for the others. */
lv_status lv_decide(lv_graph *g, int32_t group, int32_t *new_comm);
/* One sweep over groups 0..6 with the modularity after it (the body of the
 * loop at modularity_optimisation.cu:308-354); *moved = vertices moved. */
lv_status lv_sweep(lv_graph *g, double *q_after, int64_t *moved);
/* New tier options for an existing handle (no counterpart in the reference;
 * for autotuning).  The device, the timing flag, compact and exact_sums are
 * kept from creation; the current level is re-bucketed and keeps its
 * community state.  Results do not depend on these options. */
lv_status lv_set_options(lv_graph *g, const lv_options *opt);
/* Per-kernel timing, id in [0, LV_K_COUNT). */
lv_status lv_kernel_timing(lv_graph *g, int32_t kernel, lv_kernel_time *out);
lv_status lv_kernel_timing_reset(lv_graph *g);
/* Per-kernel timing of one HIP kernel, id in [0, LV_KD_COUNT). */
lv_status lv_kernel_timing_detail(lv_graph *g, int32_t kernel, lv_kernel_time *out);

const char *lv_last_error(void);

/* ---- batches of independent graphs (no counterpart in the reference) ----
 * G graphs given as one block-diagonal CSR: graph g owns vertices
 * vptr[g] .. vptr[g+1] and its rows may only reference them.  lv_batch_run
 * solves each graph exactly as lv_run would alone (its own M, phases,
 * termination and levels); all graphs share every launch (one kernel per
 * degree bucket covers the vertices of all graphs in that bucket).  Each
 * graph's weights are rounded to its own grid, as lv_create would round that
 * graph alone; the union CSR must be symmetric. */
typedef struct lv_batch lv_batch;
typedef struct {
    double modularity;       /* final Q of the graph (FP64)                          */
    int32_t levels;          /* phases run on the graph                               */
    int32_t sweeps;          /* sweeps run on the graph                               */
    int32_t final_V;         /* vertices of the graph's last level                    */
    int32_t reserved;
} lv_batch_result;
lv_status lv_batch_create(int32_t G, const int32_t *vptr, const int32_t *rowptr, const int32_t *col,
                          const float *w, int64_t E, const lv_options *opt, lv_batch **out);
void lv_batch_destroy(lv_batch *b);
/* per_graph[G] (may be null); total: summed sweeps/levels, device times, mean Q. */
lv_status lv_batch_run(lv_batch *b, float min_gain, lv_batch_result *per_graph, lv_stats *total);
/* comm[V]: final community of each vertex, numbered inside its graph as
 * lv_assignment numbers a single graph's communities. */
lv_status lv_batch_assignment(lv_batch *b, int32_t *comm);
lv_status lv_batch_kernel_timing(lv_batch *b, int32_t kernel, lv_kernel_time *out);
/* New tier options for a batch between runs (as lv_set_options; for tuning
 * the batch's launches; results do not depend on them). */
lv_status lv_batch_set_options(lv_batch *b, const lv_options *opt);

/* ---- learned frontier (no counterpart in the reference) ----
 * A move predictor (MLP 8 -> 32 -> 32 -> 1, ReLU, sigmoid; 1377 fp32
 * parameters W1[32][8], b1[32], W2[32][32], b2[32], w3[32], b3) chooses
 * which vertices a pruned sweep re-evaluates.  Its 8 features per vertex
 * are read from the state after the last sweep: log-degree, own-community
 * weight fraction and best-gain margin of the last evaluation, moves in
 * the last 4 sweeps, log community size, log community total, share of
 * neighbours that moved in the last sweep, sweeps since the last move.
 * Inference uses bf16 MFMA with fp32 accumulation; training is fp32. */
#define LV_MLP_PARAMS 1377
#define LV_FEATURES 8

#define LV_FRONTIER_MODEL 0      /* p(move) > threshold                         */
#define LV_FRONTIER_NEIGHBOR 1   /* moved in the last sweep or a neighbour did  */
#define LV_FRONTIER_RANDOM 2     /* seeded Bernoulli(random_p) per vertex/sweep  */
#define LV_FRONTIER_ALL 3        /* every vertex (a plain sweep)                 */

typedef struct {
    int32_t kind;            /* LV_FRONTIER_*                                        */
    int32_t full_every;      /* a full sweep every k sweeps of a phase (0 = 4)       */
    float threshold;         /* model: evaluate when p(move) > threshold             */
    float random_p;          /* random: evaluation probability                       */
    uint64_t seed;           /* random: seed                                         */
    float heavy_threshold;   /* model: threshold for vertices of degree above        */
    int32_t heavy_degree;    /*   heavy_degree (0 = one threshold for every degree)  */
    int32_t and_neighbor;    /* model: 1 = a vertex is evaluated only when the model
                                selects it and it or a neighbour moved in the last
                                sweep (the model prunes the neighbour rule's
                                frontier)                                           */
    int32_t reserved[5];
} lv_frontier;

typedef struct {
    int32_t epochs;
    int32_t batch;           /* samples per Adam step (contiguous ranges, caller shuffles) */
    float lr;
    float pos_weight;        /* BCE weight of the positive class (0 = 1)             */
    float beta1, beta2, eps; /* Adam (0 = 0.9, 0.999, 1e-8)                           */
    int32_t reserved[9];
} lv_train_options;

/* Levels like lv_run, with pruned phases: sweep i of a phase is full
 * (every vertex) when i % full_every == 0 and pruned (the frontier only)
 * otherwise.  Each sweep is judged by the termination rule of
 * modularity_optimisation.cu:284-363 against the sweep before it.  A pruned
 * sweep that lowered Q ends the phase (as a full sweep would); one that
 * converged (no move, or a gain in [0, min_gain)) is followed by a full
 * sweep judged against the state before that pruned sweep, which ends the
 * phase or continues it.  Without pruning this is lv_run. */
lv_status lv_run_pruned(lv_graph *g, float min_gain, const lv_frontier *fr, lv_stats *stats);
/* One pruned phase (as lv_optimise). */
lv_status lv_optimise_pruned(lv_graph *g, float min_gain, const lv_frontier *fr,
                             int32_t *changed, int32_t *sweeps, int64_t *evaluated);
/* One sweep over the frontier fr selects (LV_FRONTIER_ALL: a plain sweep);
 * track = 1 also counts neighbour moves for the next sweep's features. */
lv_status lv_sweep_pruned(lv_graph *g, const lv_frontier *fr, int32_t track, double *q_after,
                          int64_t *moved, int64_t *evaluated);
/* Learned-pruned batched run: each graph follows lv_run_pruned's rule with
 * the model frontier fr (kind LV_FRONTIER_MODEL) on its own sweeps; the
 * model's features use each graph's own M and vertex count, so every graph
 * ends as its own lv_run_pruned would.  total->full_sweeps sums the graphs'
 * full sweeps. */
lv_status lv_batch_set_model(lv_batch *b, const float *params);
lv_status lv_batch_run_pruned(lv_batch *b, float min_gain, const lv_frontier *fr,
                              lv_batch_result *per_graph, lv_stats *total);
/* One plain phase (as lv_optimise) that records, for up to max_sweeps
 * sweeps, the features before the sweep (feat[s][V][8]) and whether each
 * vertex moved in it (moved[s][V], 0/1). */
lv_status lv_trace_phase(lv_graph *g, float min_gain, float *feat, uint8_t *moved,
                         int32_t max_sweeps, int32_t *traced, int32_t *changed);
/* Features of the current level and state, feat[V][8]. */
lv_status lv_features(lv_graph *g, float *feat);
/* Raw frontier observations of the current level (any may be null):
 * margin[V], kown[V] (last evaluation), hist[V] (bit 0 = moved in the last
 * sweep), nmv[V] (row entries whose neighbour moved in the last sweep).
 * margin is exact after evaluations that record observations (lv_decide,
 * tracked pruned sweeps, the sweeps of pruned runs and traces that a frontier
 * reads); after other sweeps on levels using candidate bounds it may be lower
 * for a vertex that stayed (a candidate that could not move it is not read). */
lv_status lv_move_state(lv_graph *g, float *margin, float *kown, uint8_t *hist, int32_t *nmv);
/* Frontier model of the handle (LV_MLP_PARAMS floats). */
lv_status lv_set_model(lv_graph *g, const float *params);
/* Model output for the current level: prob[V] and active[V] (p > threshold
 * and degree > 0); either may be null. */
lv_status lv_predict(lv_graph *g, float threshold, float *prob, uint8_t *active);
/* Standalone model kernels on `device`: MFMA inference prob[n] from feat[n][8];
 * mean gradient and loss of the weighted BCE; Adam training (params in/out,
 * loss_per_epoch[epochs] may be null). */
lv_status lv_mlp_forward(int32_t device, const float *params, const float *feat, int64_t n,
                         float *prob);
lv_status lv_mlp_grad(int32_t device, const float *params, const float *feat, const float *labels,
                      int64_t n, float pos_weight, float *grad, float *loss);
lv_status lv_mlp_train(int32_t device, const float *feat, const float *labels, int64_t n,
                       const lv_train_options *opt, float *params, float *loss_per_epoch);

/* ---- weight grid (no counterpart in the reference, whose fp32 atomic sums
 * depend on thread order: modularity_optimisation.cu:144-153) ----
 * T = sum of |w| (FP64, in chunks of 65536 entries added in order),
 * c = the smallest integer with 2^c >= T * (1 + 2^-20), shift = 50 - c
 * (0 when T = 0); a weight w is rounded to q(w) = rint(w * 2^shift) *
 * 2^-shift (ties to even), which fp32 holds exactly.  Every partial sum the
 * engine forms stays below 8T, i.e. below 2^53 grid steps, so every FP64 sum
 * of rounded weights is exact.  These run on the host (no device needed). */
/* shift for w[n]; *changed (may be null) = how many weights the rounding changes. */
lv_status lv_weight_grid(const float *w, int64_t n, int32_t *shift, int64_t *changed);
/* out[i] = q(w[i]) for the given shift (out may alias w). */
lv_status lv_quantize_weights(const float *w, int64_t n, int32_t shift, float *out);

/* ---- host-side input (utils.cu:9-62 readInputData) and synthetic graphs ---- */
/* MatrixMarket coordinate reader with the reference's conventions (skip '%'
 * lines, V := cols, entries are whitespace-separated "i j w" tokens, each
 * non-loop triple stored in both rows, insertion order kept).  The file is
 * memory-mapped and parsed by all OpenMP threads; the CSR is the same for
 * any thread count.  Arrays are malloc'd; free with lv_free_host. */
lv_status lv_read_mtx(const char *path, int32_t **rowptr, int32_t **col, float **w,
                      int32_t *V, int64_t *E);
/* Binary CSR file (no reference counterpart; a cache for large inputs):
 * 8-byte magic "LVCSR01\0", int64 V, int64 E, then rowptr[V+1] (int32),
 * col[E] (int32), w[E] (fp32), little-endian. */
lv_status lv_write_csr(const char *path, const int32_t *rowptr, const int32_t *col,
                       const float *w, int32_t V, int64_t E);
lv_status lv_read_csr(const char *path, int32_t **rowptr, int32_t **col, float **w,
                      int32_t *V, int64_t *E);
/* Reads a binary CSR file (recognised by its magic) or a MatrixMarket file. */
lv_status lv_read_graph(const char *path, int32_t **rowptr, int32_t **col, float **w,
                        int32_t *V, int64_t *E);
/* Writes the lower triangle (i >= j) of a symmetric CSR as a symmetric real
 * (or integer-valued real) MatrixMarket file. */
lv_status lv_write_mtx(const char *path, const int32_t *rowptr, const int32_t *col,
                       const float *w, int32_t V);
/* RMAT graph: 2^scale vertices, edge_factor*2^scale samples with
 * probabilities (a,b,c,1-a-b-c), self-loops and duplicates removed,
 * symmetrised, seeded vertex permutation; weights 1 (unit_weights) or
 * (1 + floor(r*2^24))*2^-24 in (0,1].  Arrays malloc'd (lv_free_host). */
lv_status lv_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                      uint64_t seed, int32_t unit_weights, int32_t **rowptr, int32_t **col,
                      float **w, int32_t *V, int64_t *E);
/* Planted partition: n vertices in `blocks` near-equal blocks, each pair
 * an edge with probability p_in (same block) or p_out. */
lv_status lv_gen_planted(int32_t n, int32_t blocks, double p_in, double p_out, uint64_t seed,
                         int32_t unit_weights, int32_t **rowptr, int32_t **col, float **w,
                         int32_t *V, int64_t *E);
void lv_free_host(void *p);

#ifdef __cplusplus
}
#endif
#endif
