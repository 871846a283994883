// Sort-based community aggregation: every CSR entry becomes a key
// (new id of its row's community << 32 | new id of its column's community)
// with its weight, the keys are radix-sorted (stable LSD, 8-bit digits,
// wave-ballot ranks), equal keys are summed in FP64 and the unique keys form
// the next level's CSR, rows and columns in ascending new id.  Bandwidth-
// bound passes instead of per-community hash merges.
#pragma once

#include "lv_common.h"
#include "lv_phase1.h"

namespace lv {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;                          // per thread
constexpr int kSortTile = kSortThreads * kSortItems;   // 2048 items per workgroup

// Scratch (int32 elements) of one radix pass over n items.
int64_t radix_scratch_elems(int64_t n);

// Stable LSD sort of (keys, vals) by the bits [shift, shift+8) of each listed
// digit shift, ping-ponging between (keys, vals) and (tk, tv); returns true
// when the result ends in (tk, tv).  Digits are taken from the key with its
// high word moved down to bit lo_bits, (key >> 32) << lo_bits | (uint32)key:
// when both words are below 2^lo_bits that is one contiguous 2*lo_bits-bit
// key with the same order, so no digit is spent on the low word's zero bits
// (lo_bits = 32 sorts the key as it is).
bool radix_sort_pairs(uint64_t *keys, float *vals, uint64_t *tk, float *tv, int64_t n,
                      const int *shifts, int npasses, int lo_bits, int32_t *scratch,
                      hipStream_t s);

// keys[e] = newid[comm[v]] << 32 | newid[comm[col[e]]], vals[e] = w[e] for
// every CSR entry e of row v (rows walked by the level's segment table);
// cn (V entries) receives newid[comm[v]] first.  With ccol (col renumbered
// by cpos: ccol[e] = cpos[col[e]]) the columns are read through ccol and
// ccn (V entries, cn in that order), whose gathers share lines more often.
void launch_agg_pairs(const int32_t *rowptr, const int32_t *col, const float *w,
                      const int32_t *comm, const int32_t *newid, int32_t V, const int32_t *order,
                      const SegTable &st, const Chunk *chunks, int32_t *cn, uint64_t *keys,
                      float *vals, hipStream_t s, const int32_t *ccol = nullptr,
                      const int32_t *cpos = nullptr, int32_t *ccn = nullptr);

// From sorted keys: head[i] = 1 where a key run starts.
void launch_run_heads(const uint64_t *keys, int64_t n, int32_t *head, hipStream_t s);
// With pos = exclusive scan of head (U = unique keys): ukey[pos[i]] = key of
// each run head, acc[pos] = FP64 sum of the run's values (acc zeroed before).
void launch_run_reduce(const uint64_t *keys, const float *vals, const int32_t *head,
                       const int32_t *pos, int64_t n, uint64_t *ukey, double *acc, hipStream_t s);
// Next level's CSR from the U unique keys: cnt[src] (zeroed before) = entries
// of row src; col[k], w[k] = dst and fp32(acc[k]).
void launch_unique_rows(const uint64_t *ukey, const double *acc, const int32_t *U, int32_t *cnt,
                        int32_t *col, float *w, int64_t n_max, hipStream_t s);

}  // namespace lv
