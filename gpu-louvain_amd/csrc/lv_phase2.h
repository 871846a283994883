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
