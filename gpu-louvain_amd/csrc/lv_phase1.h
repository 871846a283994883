// Phase 1 (modularity optimisation) kernels: launch declarations.
#pragma once

#include "lv_common.h"

namespace lv {

// Chunked tier (rows too long for one workgroup).  A heavy vertex's row is
// cut into chunks of at most kChunkEdges edges; stage A aggregates each
// chunk in LDS and writes its distinct (community, weight) pairs bucketed
// by partition; stage B merges one partition of one vertex in LDS and
// evaluates its gains.  With sorted rows the chunks are cut at the 8
// column-block boundaries and dispatched to the XCD of their block, and the
// partitions are community-id ranges dispatched to the XCD of their range,
// so that both stages gather from one eighth of comm / crec.  Chunks also
// drive the heavy rows of the strength and modularity reductions.
constexpr int kChunkEdges = 1024;    // max edges per chunk (stage A LDS: 2048 slots)
constexpr int kChunkCapLog2 = 11;
constexpr int kMergeCapLog2 = 11;    // stage B LDS: 2048 slots
constexpr int kMergeLimit = 1536;    // pairs a stage-B table may take (else global)
constexpr int kPartTarget = 1024;    // expected pairs per partition
constexpr int kMaxParts = 1024;
constexpr int kMergeTile = 1024;     // chunks staged per pass in stage B

struct Chunk {
    int32_t hv;         // index of the heavy vertex
    int32_t v;          // vertex id
    int64_t lo, hi;     // edge range
    int64_t out;        // first output pair of this chunk
    int32_t pbase;      // offsets of its partitions in poff[pbase .. pbase+P]
    int32_t P;          // partitions of the vertex
    int32_t ptot_base;  // per-partition totals of the vertex in ptot[]
    int32_t first;      // 1 for the vertex's first chunk
};

struct HeavyVertex {
    int32_t v;
    int32_t P;
    int32_t chunk0, nchunks;
    int32_t ptot_base;  // also the index of the vertex's first merge item
    int32_t tab_lg;     // overflow global table: 2^tab_lg slots at tab_off
    int64_t tab_off;
};

struct MergeItem {
    int32_t hv;
    int32_t p;
};

struct ChunkedTier {
    const Chunk *chunks;
    const HeavyVertex *hvs;
    const MergeItem *items;
    int32_t *out_key;   // pairs written by stage A
    double *out_val;
    int32_t *poff;      // per chunk: P+1 partition offsets
    int32_t *ptot;      // per vertex and partition: pairs (zeroed by stage B)
    int32_t *gkeys;     // overflow tables (kept all-empty between uses)
    double *gvals;
    int32_t *ovf;       // per heavy vertex: its overflow table was used (cleared after the sweep step)
    uint64_t *best;     // per vertex packed argmax, 0 = none
    double *kcur;       // per vertex weight to its own community
    uint64_t *item_best;  // per merge item: its packed argmax
    double *item_K;       // ... and the weight to that candidate
    int32_t merge_limit;
    // XCD-aligned dispatch (null = identity): workgroup b of stage A takes
    // chunk chunk_order[b], of stage B item item_order[b]; -1 = padding.
    const int32_t *chunk_order;
    const int32_t *item_order;
    int32_t n_chunk_order, n_item_order;
    int32_t range_blk;    // > 0: partitions by community-id range (see tier_part)
    int32_t chunk_cap_log2;  // stage A table slots (log2): twice the chunk size
    int32_t direct;          // stage A may write rows of mostly distinct
                             // communities without merging a chunk's duplicates
    // LDS attempt (launch_heavy_try), per heavy vertex: decided in LDS this
    // sweep step; distinct communities of its last evaluation (the attempt
    // is made when they fit); their count summed by stage B.
    int32_t *done;
    int32_t *hint;
    int32_t *hdist;
};

// Distinct communities up to which a heavy row is tried in LDS; the table
// (launch_heavy_try's slots) needs more than this plus 1024 slots.
constexpr int32_t kHeavyTryLimit = 4096;

// Segment of the vertex list processed with a fixed lane count per vertex
// (row reductions: strength and modularity).
struct Seg {
    int32_t list_off;
    int32_t count;
    int32_t block_start;
    int32_t lanes_log2;  // 2..8; 8 = whole 256-lane workgroup per vertex
};
constexpr int kMaxSegs = 12;
struct SegTable {
    Seg s[kMaxSegs];
    int32_t nseg;
    int32_t chunk_block_start;  // blocks >= this handle heavy chunks
    int32_t nchunks;
    int32_t sq_block_start;     // blocks >= this handle sum of S_c^2 (modularity only)
    int32_t total_blocks;
};
constexpr int kRowThreads = 256;
constexpr int kSqItems = 2048;  // communities per block in the S_c^2 pass

// --- level setup ---
void launch_degree_keys(const int32_t *rowptr, int32_t V, BucketBounds bb, int8_t *keys,
                        hipStream_t s);
// With loops_out: also *loops_out = the weight of the self-loop entries
// (col[e] == row), FP64, from per-block partials (loop_partials: one per
// block of st, then kSelfLoopBlocks folds) in a fixed order.
void launch_strength(const int32_t *rowptr, const float *w, const int32_t *order,
                     const SegTable &st, const Chunk *chunks, double *k, hipStream_t s,
                     const int32_t *col = nullptr, double *loop_partials = nullptr,
                     double *loops_out = nullptr);
void launch_level_state(int32_t V, const double *k, int32_t *comm, int32_t *size, double *sigma,
                        uint64_t *crec, hipStream_t s);
void launch_records_all(int32_t V, const double *sigma, const int32_t *size, uint64_t *crec,
                        hipStream_t s);

// Column numbering of the move kernels (MoveArgs::ccol, cself): order lists
// all V vertices; cpos = its inverse, ccol[e] = cpos[col[e]].  launch_col_comm
// copies comm into that numbering (after any change of comm but the applies').
void launch_col_numbering(const int32_t *order, int32_t V, const int32_t *col, int64_t E,
                          int32_t *cpos, int32_t *ccol, hipStream_t s);
void launch_col_comm(const int32_t *order, int32_t V, const int32_t *comm, int32_t *ccomm,
                     hipStream_t s);

// --- moves ---
// X lanes per vertex, E row entries per lane (degrees <= X * E); the shapes
// move_sub_shape accepts: X * E = 4 (4x1), 8 (8x1, 4x2), 16 (16x1, 8x2, 4x4)
// or 32 (32x1, 16x2, 8x4, 4x8).
bool move_sub_shape(int X, int E);
void launch_move_sub(int X, int E, const MoveArgs &a, const int32_t *list, int32_t n,
                     hipStream_t s);
// cap: LDS slots per row: 128 (degrees <= 84) or 192 (degrees <= 128) with
// two rows per wave, 256 (degrees <= 255) or 448 (degrees <= 319) with one
// row per wave.
void launch_move_wave(int cap, const MoveArgs &a, const int32_t *list, int32_t n, hipStream_t s);
// tier 0..3: the workgroup buckets in degree order (separate kernel
// instances), lanes (128, 256, 512 or 1024) per vertex.
// slots: LDS table entries allocated per workgroup (12 bytes each, more than
// the tier's degree bound); pct: a row's table is pct % of its degree
// (at least 512 slots, at most the allocation).  hint/done (both or neither,
// indexed like list): rows with done[i] set were decided by
// launch_move_hinted and are skipped; hint[i] gets the distinct communities
// of each row evaluated.
void launch_move_wg(int tier, int lanes, int slots, int pct, const MoveArgs &a,
                    const int32_t *list, int32_t n, int32_t *hint, int32_t *done, hipStream_t s);
// Small-table attempt over the one-workgroup tiers' rows (list, n): rows with
// hint[i] <= 256 are tried one wave per row in 512 slots; done[i] = 1 for
// those decided (0 for the others), hint[i] updated.  The next n2 rows of
// list are the chunked tier's (hints and flags hint2/done2, its own), tried
// the same way; done2 is only set, for rows decided here.
void launch_move_hinted(const MoveArgs &a, const int32_t *list, int32_t n, int32_t *hint,
                        int32_t *done, int32_t n2, int32_t *hint2, int32_t *done2,
                        hipStream_t s);
void set_wg_lds_limits(int bytes);
// LDS attempt of the chunked tier's rows (list = its heavy vertices, in the
// order of t.hvs), before its stages; fills t.done and t.hint.  A row is
// given up at `limit` distinct communities (at most kHeavyTryLimit).
void launch_heavy_try(int slots, int32_t limit, const MoveArgs &a, const int32_t *list,
                      int32_t n, const ChunkedTier &t, hipStream_t s);
void launch_move_chunked(const MoveArgs &a, const ChunkedTier &t, int32_t nchunks,
                         int32_t nitems, const int32_t *heavy_list, int32_t nh, hipStream_t s);
// The three kernels of launch_move_chunked, for callers that time them apart.
void launch_chunk_aggregate(const MoveArgs &a, const ChunkedTier &t, int32_t nchunks,
                            hipStream_t s);
void launch_chunk_merge(const MoveArgs &a, const ChunkedTier &t, int32_t nitems, hipStream_t s);
void launch_heavy_final(const MoveArgs &a, const ChunkedTier &t, int32_t nh, hipStream_t s);
void launch_clear_tables(int32_t *keys, double *vals, int64_t n, hipStream_t s);

// Bitmaps of the movers of a group step (two, alternating by step, so that
// one is cleared while the other is read).
constexpr int kMoveSlots = 64;
constexpr int kMoveSlotStride = 16;  // 128 bytes between slots

struct GroupMarks {
    uint32_t *movers;        // this step's bitmap over vertices
    uint32_t *clear_movers;  // the other step's, cleared by this step
    int32_t words;
    int32_t rounds;  // correction: rows walked per lane group of a block (see corr_items)
    // batches: the correction of each mover's row goes to gin[gid[v]] (the
    // weight inside communities of v's graph) instead of own
    const int32_t *gid;
    double *gin;
};

// Items one correction block checks: rounds rows per X-lane group, at most
// kCorrItems.  rounds = 1 gives one row per lane group (most vertices
// moved); late sweeps, where few move, use more so that blocks stay few.
constexpr int32_t kCorrItems = 1024;
// Lanes per row of the correction passes: 4 for rows of at most 32 entries
// (of a row table's segments with lanes_log2 <= 5; their gathers go four a
// step), else the segment's lanes, at most a workgroup.
__host__ __device__ inline int corr_lanes(int lanes_log2) {
    return lanes_log2 <= 5 ? 4 : 1 << (lanes_log2 < 8 ? lanes_log2 : 8);
}
__host__ __device__ inline int32_t corr_items(int X, int32_t rounds) {
    const int64_t c = (int64_t)(kRowThreads / X) * rounds;
    return (int32_t)(c < kCorrItems ? c : kCorrItems);
}

// --- apply ---
// Applies the decisions of the vertices list[0..n) (a superset of the
// group's movers; newc[v] then holds the community v left) and refreshes the
// records of the touched communities.  With gm (incremental modularity) the
// movers are marked and the group's change of the weight inside communities
// is added to own (kMoveSlots slots kMoveSlotStride apart; gst: the group's
// segments over order, then its heavy-row chunks, corr_items(X, gm->rounds)
// items per block).
// Without gm, sweep_movers (may be null) is a bitmap over vertices in which
// the movers are marked (the sweep's movers, for launch_sweep_delta).
// hist[v] gets bit 0 set for a mover (the move history of the frontier model);
// moved holds kMoveSlots counter pairs kMoveSlotStride apart: summed over
// the slots, [0] counts the movers and [1] their row entries.
void launch_apply(const MoveArgs &a, const int32_t *list, int32_t n, const int32_t *order,
                  const SegTable *gst, const Chunk *chunks, int32_t *comm, int32_t *size,
                  double *sigma, uint64_t *crec, const GroupMarks *gm, uint8_t *hist,
                  unsigned long long *moved, double *own, hipStream_t s,
                  uint32_t *sweep_movers = nullptr);

// --- modularity ---
// *out = the weight of the self-loop entries (col[e] == row), FP64, over a
// fixed grid of kSelfLoopBlocks blocks with per-block partials, for rows
// sorted by column without duplicates (a binary search per row; unsorted
// levels sum it with their strengths, launch_strength).
constexpr int kSelfLoopBlocks = 1024;
void launch_selfloop_weight(const int32_t *rowptr, const int32_t *col, const float *w, int32_t V,
                            double *partials, double *out, hipStream_t s);
// Full pass: per-block partials (own, sq), then q[0] = Q, q[1] = own, q[2] = sq.
void launch_modularity(const int32_t *rowptr, const int32_t *col, const float *w,
                       const int32_t *comm, const double *sigma, int32_t V, const int32_t *order,
                       const SegTable &st, const Chunk *chunks, double M, double *partials,
                       double *q, hipStream_t s);

// The two halves of launch_modularity: the row/square pass (per-block
// partials) and the fixed-order reduction into q.  With ccol, the rows'
// columns are read as ccol (a renumbering of col) and their communities as
// ccomm (comm in that numbering).
void launch_modularity_rows(const int32_t *rowptr, const int32_t *col, const float *w,
                            const int32_t *comm, const double *sigma, int32_t V,
                            const int32_t *order, const SegTable &st, const Chunk *chunks,
                            double *partials, hipStream_t s, const int32_t *ccol = nullptr,
                            const int32_t *ccomm = nullptr);
void launch_modularity_finish(const SegTable &st, double M, double *partials, double *q,
                              hipStream_t s);

// nmv[v] = entries (v, u) of v's row whose u moved in the last sweep (hist
// bit 0) and has degree 1..max_degree, for the rows of st (its segments and
// chunks: nmv must be zero before for the chunked rows, and for vertices in
// no row).  marks: (V + 31) / 32 words of scratch.
void launch_neighbor_count(const int32_t *rowptr, const int32_t *col, const uint8_t *hist,
                           int32_t V, int32_t max_degree, const int32_t *order,
                           const SegTable &st, const Chunk *chunks, uint32_t *marks,
                           int32_t *nmv, hipStream_t s);

// XCD-local row passes over the heavy rows (degree > 32, bucket-ordered
// list): split[i*9 .. i*9+9) = the row of list[i] cut at the column-block
// boundaries x*ceil(V/8); *unsorted |= 1 when a listed row is not sorted
// by column (the split is then unusable).
void launch_xcd_splits(const int32_t *rowptr, const int32_t *col, const int32_t *list, int32_t n,
                       int32_t V, int32_t *split, int32_t *unsorted, hipStream_t s);
// Weight inside communities of the heavy rows, one column block per XCD;
// writes (own, 0) partials at blocks [base, base + 8*xst.total_blocks).
void launch_modularity_xcd(const int32_t *col, const float *w, const int32_t *comm,
                           const int32_t *list, const SegTable &xst, const int32_t *split,
                           double *partials, int32_t base, hipStream_t s);
// Fixed-order reduction of nb partial pairs into q (as launch_modularity_finish);
// S_c^2 is summed over the nb_sq pairs from sq_lo.
void launch_modularity_reduce(int32_t nb, int32_t sq_lo, int32_t nb_sq, double M,
                              double *partials, double *q, hipStream_t s);

// Change over one sweep of the weight inside communities, added to the
// kMoveSlots slots of own_slots: the rows of the vertices whose community
// changed (comm0 = communities before the sweep; movers = the sweep's mover
// bitmap; dk = each mover's weight change with its neighbours as its
// decision saw them) walked by the blocks of the correction table ct (the
// level's row table with corr_items block starts for rounds).
void launch_sweep_delta(const int32_t *rowptr, const int32_t *col, const float *w,
                        const int32_t *comm, const int32_t *comm0, const double *dk,
                        const uint32_t *movers, const int32_t *order, const SegTable &ct,
                        const Chunk *chunks, int32_t rounds, double *own_slots, hipStream_t s);

// Incremental: one pass for sum of S_c^2; q[0] = Q from q[1] plus the
// applies' corrections in own_slots (kMoveSlots slots kMoveSlotStride apart,
// folded into q[1] and cleared), q[2] = sq.
void launch_modularity_from_own(const double *sigma, int32_t V, double M, double *partials,
                                double *q, double *own_slots, hipStream_t s);

// --- original-vertex map ---
void launch_compose(int32_t *orig2comm, int32_t V0, const int32_t *map, hipStream_t s);

}  // namespace lv
