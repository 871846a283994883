// Phase 1 (modularity optimisation) kernels: launch declarations.
#pragma once

#include "lv_common.h"

namespace lv {

// Work table for rows too long for one workgroup: one entry per chunk of a
// heavy vertex's row (edge range) or of its hash table (slot range).
struct Chunk {
    int32_t hv;  // index into the heavy-vertex list
    int32_t v;   // vertex id
    int64_t lo;  // first edge / slot (absolute)
    int64_t hi;  // one past the last
};

// Heavy-vertex global hash tables (one region per heavy vertex).
struct HeavyTables {
    int32_t *keys;
    double *vals;
    const int64_t *off;       // [nh] slot offset of each vertex's table
    const int32_t *cap_log2;  // [nh]
    uint64_t *best;           // [nh] packed argmax, 0 = none
    double *kcur;             // [nh] weight to the own community
};

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
void launch_strength(const int32_t *rowptr, const float *w, const int32_t *order,
                     const SegTable &st, const Chunk *chunks, double *k, hipStream_t s);
void launch_level_state(int32_t V, const double *k, int32_t *comm, int32_t *size, double *sigma,
                        hipStream_t s);

// --- moves ---
void launch_move_sub(int X, const MoveArgs &a, const int32_t *list, int32_t n, hipStream_t s);
void launch_move_wave(int cap, const MoveArgs &a, const int32_t *list, int32_t n, hipStream_t s);
void launch_move_wg(int threads, int cap_log2, const MoveArgs &a, const int32_t *list, int32_t n,
                    hipStream_t s);
int move_wg_max_cap_log2(int max_lds_bytes);
void set_wg_lds_limits(int bytes);
void launch_move_heavy(const MoveArgs &a, const int32_t *heavy_list, int32_t nh,
                       const Chunk *ins, int32_t nins, const Chunk *eval, int32_t neval,
                       const HeavyTables &t, hipStream_t s);
void launch_clear_tables(int32_t *keys, double *vals, int64_t n, hipStream_t s);

// --- apply ---
void launch_apply(const int32_t *list, int32_t n, const int32_t *newc, int32_t *comm,
                  int32_t *size, double *sigma, const double *k, unsigned long long *moved,
                  hipStream_t s);

// --- modularity ---
// Writes per-block partials (own, sq) and then q[0] = Q, q[1] = own, q[2] = sq.
void launch_modularity(const int32_t *rowptr, const int32_t *col, const float *w,
                       const int32_t *comm, const double *sigma, int32_t V, const int32_t *order,
                       const SegTable &st, const Chunk *chunks, double M, double *partials,
                       double *q, hipStream_t s);

// --- original-vertex map ---
void launch_compose(int32_t *orig2comm, int32_t V0, const int32_t *map, hipStream_t s);

}  // namespace lv
