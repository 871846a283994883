// Batches of independent graphs solved together as one block-diagonal union
// graph: every bucket kernel of a sweep covers the vertices of all graphs of
// that bucket, each vertex using its own graph's M (MoveArgs::gid), and the
// per-graph modularity, move counts and termination are reduced by graph
// range.  Graphs occupy contiguous vertex ranges vptr[g] .. vptr[g+1] at
// every level (community ids stay inside their graph's range and the
// order-preserving renumbering keeps the ranges contiguous).
#pragma once

#include "lv_common.h"
#include "lv_phase1.h"

namespace lv {

// own[v] = weight of v's row entries inside comm[v] (FP64; heavy rows by chunk
// with atomics, exact for grid weights) for the rows st lists over order
// (chunks of vertices with active[v] == 0 are skipped; active may be null).
// own must be zero before.
void launch_vertex_own(const int32_t *rowptr, const int32_t *col, const float *w,
                       const int32_t *comm, const int32_t *order, const SegTable &st,
                       const Chunk *chunks, const uint8_t *active, double *own, hipStream_t s);
// Per graph g (one workgroup each): out[3g] = the weight inside communities,
// out[3g+1] = sum of sigma^2 over its communities, out[3g+2] = vertices that
// moved in the last sweep (hist bit 0).  Fixed order.  With own, the weight
// inside is the sum of own over the graph's vertices and is also stored in
// gin[g]; with own null it is read from gin[g] (kept up to date by the
// applies' corrections).  Graphs with gact[g] == 0 are skipped (gact may be
// null).
void launch_graph_sums(const int32_t *vptr, int32_t G, const double *own, const double *sigma,
                       const uint8_t *hist, const uint8_t *gact, double *gin, double *out,
                       hipStream_t s);
// Frontier of the active graphs: active[v] = gact[gid[v]] and deg > 0;
// keys[v] = bucket or -1.
void launch_graph_frontier(const int32_t *rowptr, int32_t V, const int32_t *gid,
                           const uint8_t *gact, BucketBounds bb, int8_t *keys, uint8_t *active,
                           hipStream_t s);
// gid[v] for v in [vptr[g], vptr[g+1]).
void launch_graph_ids(const int32_t *vptr, int32_t G, int32_t *gid, hipStream_t s);
// After a renumbering newid (exclusive scan of present flags, total newV):
// vptr_new[g] = newid[vptr[g]] (newV for g = G and for empty tails).
void launch_graph_offsets(const int32_t *vptr, int32_t G, int32_t V, const int32_t *newid,
                          const int32_t *newV, int32_t *vptr_new, hipStream_t s);
// final[x] = orig2comm[x] - vptr[g0[x]] for original vertices x whose graph
// g0[x] has done[g0[x]] set.
void launch_graph_final(const int32_t *orig2comm, int32_t V0, const int32_t *gid0,
                        const uint8_t *done, const int32_t *vptr, int32_t *final_comm,
                        hipStream_t s);

}  // namespace lv
