// Phase 2 (community aggregation) kernels: launch declarations.
#pragma once

#include "lv_common.h"

namespace lv {

// Operands of the merge kernels.  Communities are addressed by their
// current-level id c.  Members of c are members[mstart[c] .. mstart[c]+size[c]),
// and mscan is the exclusive scan of member degrees in that order, so the
// row entries of all members of c form the flat range
// [edgepos[c], edgepos[c] + D[c]) with edgepos[c] = mscan[mstart[c]].
struct AggArgs {
    const int32_t *rowptr;
    const int32_t *col;
    const float *w;
    const int32_t *comm;
    const int32_t *size;
    const int32_t *newid;
    const int32_t *mstart;
    const int32_t *members;
    const int32_t *mscan;  // [V+1]
    int32_t *cnt;          // distinct neighbour communities per community
    int32_t *tmp_col;      // merged rows at edgepos[c] (upper-bound layout)
    float *tmp_w;
};

// Aggregation tiers by summed member degree D (LDS tiers are exact-fit
// hash + bitonic sort inside one workgroup; the last tier is global).
constexpr int kAggTiers = 4;
constexpr int32_t kAggTierMax[kAggTiers] = {128, 1024, 4096, 0x7fffffff};

struct AggChunk {
    int32_t c;      // community
    int32_t ti;     // index in the global-tier list
    int64_t lo;     // flat member-edge offset (insert) or absolute slot (extract)
    int64_t hi;
};

void launch_present(const int32_t *size, int32_t V, int32_t *flag, hipStream_t s);

// --- levels without their isolated vertices ---
// A level may store only its vertices of positive degree ("compact" ids,
// ascending in the reference's ids); ref_of maps compact -> reference id and
// cmp_of reference -> compact id (-1 for a vertex without edges, which is
// always a singleton community that no vertex can join).
//
// flag[c] = 1 for every community c holding a vertex of positive degree
// (flag zero before).
void launch_has_degree(const int32_t *rowptr, const int32_t *comm, int32_t V, int32_t *flag,
                       hipStream_t s);
// flag[c] = size[c] > 0 && flag[c]: the communities that become vertices of
// the next compact level.
void launch_present_compact(const int32_t *size, int32_t V, int32_t *flag, hipStream_t s);
// present_ref (all ones over the reference ids of the level before) gets
// present_ref[ref_of[c]] = size[c] > 0 for every compact community c.
void launch_ref_present(const int32_t *size, const int32_t *ref_of, int32_t V,
                        int32_t *present_ref, hipStream_t s);
// ref_of_next[newid[c]] = newid_ref[ref_of[c]] for the communities c with
// keep[c] != 0.
void launch_next_ref_of(const int32_t *keep, const int32_t *newid, const int32_t *ref_of,
                        const int32_t *newid_ref, int32_t V, int32_t *ref_of_next, hipStream_t s);
// cmp_of[ref_of[i]] = i for i < V (cmp_of filled with -1 before).
void launch_cmp_of(const int32_t *ref_of, int32_t V, int32_t *cmp_of, hipStream_t s);
// End of a phase at a compact level: orig2comm[x] (a reference id r) becomes
// the reference id of r's community, ref_of[comm[cmp_of[r]]] (r itself when
// it has no edges).
void launch_compose_ref(int32_t *orig2comm, int32_t Vorig, const int32_t *cmp_of,
                        const int32_t *ref_of, const int32_t *comm, hipStream_t s);
void launch_members(const int32_t *comm, int32_t V, const int32_t *mstart, int32_t *fillc,
                    int32_t *members, hipStream_t s);
void launch_member_degrees(const int32_t *rowptr, const int32_t *members, int32_t V,
                           int32_t *mdeg, hipStream_t s);
// Tier key per community (0..3, -1 = nothing to merge).  bounds[t] caps
// tier t (a tier the device cannot hold in LDS gets its predecessor's bound,
// leaving it empty); larger communities go to the global tier.
struct AggBounds {
    int32_t hi[kAggTiers - 1];
};
void launch_agg_keys(const int32_t *size, const int32_t *mstart, const int32_t *mscan, int32_t V,
                     AggBounds bounds, int8_t *keys, hipStream_t s);
void launch_merge_lds(int tier, const AggArgs &a, const int32_t *list, int32_t n, hipStream_t s);
int merge_lds_bytes(int tier);
void set_merge_lds_limits(int bytes);
void launch_agg_global(const AggArgs &a, const AggChunk *ins, int32_t nins,
                       const AggChunk *ext, int32_t next, const int32_t *glist, int32_t ng,
                       const int64_t *tab_off, const int32_t *cap_log2, int32_t *keys,
                       double *vals, uint64_t *sort_a, uint64_t *sort_b,
                       const int64_t *sort_off, hipStream_t s);
// rowptr2[newid[c]] = cstart[c] for present c; rowptr2[newV] = *total.
void launch_new_rowptr(const int32_t *size, const int32_t *newid, const int32_t *cstart,
                       int32_t V, const int32_t *newV, const int32_t *total, int32_t *rowptr2,
                       hipStream_t s);
// Copy merged rows of the listed communities to their final positions.
void launch_copy_rows(const int32_t *list, int32_t n, bool wide, const int32_t *mstart,
                      const int32_t *mscan, const int32_t *cnt, const int32_t *cstart,
                      const int32_t *tmp_col, const float *tmp_w, int32_t *col2, float *w2,
                      hipStream_t s);

}  // namespace lv
