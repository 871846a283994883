// Engine and C-ABI of the Louvain library (include/louvain.h).
//
// Host orchestration of one graph on one GPU stream.  Per level:
//   setup   degree buckets (stable partition, no PCIe traffic), strengths,
//           singleton state, heavy-vertex work tables (planned and written on
//           the device; the host reads back bucket offsets and table totals);
//   phase   sweeps of 7 degree groups (moves of a group are simultaneous,
//           groups run in order), each group = move kernels + apply; one
//           modularity per sweep; the termination test (optimiseModularity,
//           modularity_optimisation.cu:284-363) runs on the device and the
//           host enqueues the next sweep ahead of it;
//   merge   aggregation into the other CSR buffer (aggregateCommunities,
//           community_aggregation.cu:248-347).
// The level-0 CSR is kept so lv_reset can restart without re-uploading.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/louvain.h"
#include "lv_common.h"
#include "lv_error.h"
#include "lv_aggsort.h"
#include "lv_batch.h"
#include "lv_learn.h"
#include "lv_phase1.h"
#include "lv_phase2.h"
#include "lv_prims.h"
#include "lv_weights.h"

namespace {

thread_local std::string g_last_error;

lv_status fail(lv_status st, const std::string &msg) {
    g_last_error = msg;
    return st;
}

#define LV_HIP(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? LV_ERR_OOM : LV_ERR_HIP,                 \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

#define LV_TRY(expr)                                                                         \
    do {                                                                                     \
        lv_status s_ = (expr);                                                               \
        if (s_ != LV_OK) return s_;                                                          \
    } while (0)

#define LV_LAUNCHED(what)                                                                    \
    do {                                                                                     \
        hipError_t e_ = hipGetLastError();                                                   \
        if (e_ != hipSuccess)                                                                \
            return fail(LV_ERR_HIP, std::string("launch ") + what + ": " + hipGetErrorString(e_)); \
    } while (0)

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

__global__ void k_zero_u64_n(unsigned long long *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// flag = 1 when some weight is negative or NaN (grid-stride).
__global__ void k_weight_sign(const float *__restrict__ w, int64_t n, int32_t *flag) {
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        bad |= !(w[i] >= 0.0f);
    if (__ballot(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

// ------------------------------------------------ chunked tier work tables
// Built on the device once per level.  k_tier_plan (one workgroup) walks the
// heavy vertices in list order and gives each its partition count, chunks,
// merge items and table slots by exclusive scans; k_tier_fill (one workgroup
// per heavy vertex) writes the chunks and merge items.  The host reads back
// only the totals (to size the tables).  With xcd = true the chunks of a row
// are cut at its column-block boundaries (split, 9 per row) and both lists
// get XCD-aligned dispatch orders: workgroup 8j + x takes the j-th chunk
// (item) of XCD x's list.

// Per heavy vertex, what k_tier_fill needs beyond its HeavyVertex.
struct HeavyPlan {
    int64_t lo, hi;  // row range
    int64_t out0;    // first pair slot of its first chunk
    int32_t pbase0;  // partition offsets of its first chunk
    int32_t xc[8];   // rank of its first chunk of column block x in XCD x's list
    int32_t xi[8];   // ... of its first merge item of block x
};

struct TierTotals {
    int64_t outs, slots, nchunks, nitems, pbase;
    int64_t xc[8];  // chunks in XCD x's list
    int64_t xi[8];  // merge items in XCD x's list
};

constexpr int kPlanThreads = 256;

// Exclusive scan of x over the workgroup (kPlanThreads threads) plus the
// running carry; returns the thread's offset and adds the tile total to carry.
__device__ __forceinline__ int64_t plan_scan(int64_t x, int64_t &carry, int64_t *lds) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int64_t inc = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int64_t y = __shfl_up(inc, off, 64);
        if (lane >= off) inc += y;
    }
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    int64_t woff = 0, tot = 0;
#pragma unroll
    for (int w2 = 0; w2 < kPlanThreads / 64; ++w2) {
        woff += w2 < wid ? lds[w2] : 0;
        tot += lds[w2];
    }
    __syncthreads();
    const int64_t r = carry + woff + inc - x;
    carry += tot;
    return r;
}

template <bool XCD>
__global__ __launch_bounds__(kPlanThreads) void k_tier_plan(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ list, int32_t nh, int32_t ce,
    int32_t pt, const int32_t *__restrict__ split, lv::HeavyVertex *hvs, HeavyPlan *plan,
    TierTotals *tot) {
    __shared__ int64_t lds[kPlanThreads / 64];
    constexpr bool xcd = XCD;
    int64_t c_chunk = 0, c_item = 0, c_pb = 0, c_slot = 0, c_out = 0;
    int64_t c_xc[8] = {0, 0, 0, 0, 0, 0, 0, 0}, c_xi[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int32_t base = 0; base < nh; base += kPlanThreads) {
        const int32_t i = base + (int32_t)threadIdx.x;
        const bool ok = i < nh;
        int32_t v = 0;
        int64_t lo = 0, hi = 0, deg = 0, P = 0, n = 0;
        int64_t nx[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int tab_lg = 6;
        if (ok) {
            v = list[i];
            lo = rowptr[v];
            hi = rowptr[v + 1];
            deg = hi - lo;
            if constexpr (XCD) {
                const int64_t m = max((int64_t)1, (deg + 8 * (int64_t)pt - 1) / (8 * (int64_t)pt));
                P = min((int64_t)lv::kMaxParts, 8 * m);
#pragma unroll
                for (int x = 0; x < 8; ++x) {
                    const int64_t s0 = split[9 * (int64_t)i + x], s1 = split[9 * (int64_t)i + x + 1];
                    nx[x] = (s1 - s0 + ce - 1) / ce;
                    n += nx[x];
                }
            } else {
                P = max((int64_t)1, min((int64_t)lv::kMaxParts, (deg + pt - 1) / pt));
                n = (deg + ce - 1) / ce;
            }
            while (((int64_t)1 << tab_lg) < 2 * deg) ++tab_lg;  // >= 2 deg, at least 2^6
        }
        const int64_t chunk0 = plan_scan(n, c_chunk, lds);
        const int64_t tb = plan_scan(P, c_item, lds);
        const int64_t pb = plan_scan(n * (P + 1), c_pb, lds);
        const int64_t toff = plan_scan(ok ? ((int64_t)1 << tab_lg) : 0, c_slot, lds);
        const int64_t out0 = plan_scan(deg, c_out, lds);
        HeavyPlan hp{};
        if constexpr (XCD) {
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                hp.xc[x] = (int32_t)plan_scan(nx[x], c_xc[x], lds);
                hp.xi[x] = (int32_t)plan_scan(P / 8, c_xi[x], lds);
            }
        }
        if (ok) {
            lv::HeavyVertex h;
            h.v = v;
            h.P = (int32_t)P;
            h.chunk0 = (int32_t)chunk0;
            h.nchunks = (int32_t)n;
            h.ptot_base = (int32_t)tb;
            h.tab_lg = tab_lg;
            h.tab_off = toff;
            hvs[i] = h;
            hp.lo = lo;
            hp.hi = hi;
            hp.out0 = out0;
            hp.pbase0 = (int32_t)pb;
            plan[i] = hp;
        }
    }
    if (threadIdx.x == 0) {
        tot->outs = c_out;
        tot->slots = c_slot;
        tot->nchunks = c_chunk;
        tot->nitems = c_item;
        tot->pbase = c_pb;
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            tot->xc[x] = c_xc[x];
            tot->xi[x] = c_xi[x];
        }
    }
}

__global__ __launch_bounds__(256) void k_tier_fill(const lv::HeavyVertex *__restrict__ hvs,
                                                   const HeavyPlan *__restrict__ plan, int32_t ce,
                                                   const int32_t *__restrict__ split,
                                                   lv::Chunk *chunks, lv::MergeItem *items,
                                                   int32_t *corder, int32_t *iorder) {
    const int32_t i = blockIdx.x;
    const lv::HeavyVertex h = hvs[i];
    const HeavyPlan hp = plan[i];
    auto put = [&](int32_t k, int64_t clo, int64_t chi) {
        lv::Chunk c;
        c.hv = i;
        c.v = h.v;
        c.lo = clo;
        c.hi = chi;
        c.out = hp.out0 + (clo - hp.lo);  // chunks tile the row in order
        c.pbase = hp.pbase0 + k * (h.P + 1);
        c.P = h.P;
        c.ptot_base = h.ptot_base;
        c.first = k == 0 ? 1 : 0;
        chunks[h.chunk0 + k] = c;
    };
    if (!split) {
        for (int32_t k = threadIdx.x; k < h.nchunks; k += blockDim.x) {
            const int64_t clo = hp.lo + (int64_t)k * ce;
            put(k, clo, min(hp.hi, clo + ce));
        }
    } else {
        int32_t kb = 0;
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            const int64_t s0 = split[9 * (int64_t)i + x], s1 = split[9 * (int64_t)i + x + 1];
            const int32_t cnt = (int32_t)((s1 - s0 + ce - 1) / ce);
            for (int32_t j = threadIdx.x; j < cnt; j += blockDim.x) {
                const int64_t clo = s0 + (int64_t)j * ce;
                put(kb + j, clo, min(s1, clo + ce));
                corder[8 * (hp.xc[x] + j) + x] = h.chunk0 + kb + j;
            }
            kb += cnt;
        }
    }
    const int32_t m = h.P / 8;
    for (int32_t p = threadIdx.x; p < h.P; p += blockDim.x) {
        items[h.ptot_base + p] = lv::MergeItem{i, p};
        if (split) iorder[8 * (plan[i].xi[p / m] + p % m) + p / m] = h.ptot_base + p;
    }
}

// Level 0 without its vertices of degree 0.  flag[v] = deg(v) > 0.
__global__ void k_degree_flags(const int32_t *__restrict__ rowptr, int32_t V, int32_t *flag) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) flag[v] = rowptr[v + 1] > rowptr[v] ? 1 : 0;
}

// From the exclusive scan of those flags: ref_of[compact] = vertex,
// cmp_of[vertex] = compact id or -1.
__global__ void k_level0_maps(const int32_t *__restrict__ rowptr, int32_t V,
                              const int32_t *__restrict__ scan, int32_t *ref_of, int32_t *cmp_of) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const bool keep = rowptr[v + 1] > rowptr[v];
    if (keep) ref_of[scan[v]] = v;
    cmp_of[v] = keep ? scan[v] : -1;
}

// rowptr of the kept vertices (their rows are contiguous already: the
// dropped rows are empty); rp_c[Vc] = E.
__global__ void k_compact_rowptr(const int32_t *__restrict__ rowptr,
                                 const int32_t *__restrict__ ref_of, int32_t Vc, int32_t E,
                                 int32_t *rp_c) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < Vc) rp_c[i] = rowptr[ref_of[i]];
    if (i == Vc) rp_c[i] = E;
}

__global__ void k_relabel(int32_t *col, int64_t E, const int32_t *__restrict__ cmp_of) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x)
        col[e] = cmp_of[col[e]];
}

// Edge counts per bucket (algorithmic-byte accounting): per-block LDS
// counters, one global atomic per block and bucket.
__global__ __launch_bounds__(256) void k_bucket_edges(const int32_t *__restrict__ rowptr,
                                                      const int32_t *__restrict__ order,
                                                      const int32_t *__restrict__ off, int nb,
                                                      unsigned long long *edges) {
    __shared__ unsigned long long cnt[16];
    if (threadIdx.x < 16) cnt[threadIdx.x] = 0ull;
    __syncthreads();
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < off[nb]) {
        int b = 0;
        while (b + 1 < nb && i >= off[b + 1]) ++b;
        const int32_t v = order[i];
        atomicAdd(&cnt[b], (unsigned long long)(rowptr[v + 1] - rowptr[v]));
    }
    __syncthreads();
    if (threadIdx.x < nb && cnt[threadIdx.x]) atomicAdd(&edges[threadIdx.x], cnt[threadIdx.x]);
}

// Sizes and totals of an externally supplied state.
__global__ void k_state_from_comm(int32_t V, const int32_t *__restrict__ comm,
                                  const double *__restrict__ k, int32_t *size, double *sigma) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) {
        const int32_t c = comm[v];
        atomicAdd(&size[c], 1);
        atomicAdd(&sigma[c], k[v]);
    }
}

// Summed member degree of each global-tier community (host builds its chunks).
// Work tables of the aggregation's global hash tier (agg_mode 1), built on
// the device like the chunked tier's: per listed community i, its summed
// member degree D (insert chunks of `ic` flat member edges), a table of
// 2^lg >= 2 min(D, newV) slots (extract chunks of `xc` slots) and its sort
// range.  k_agg_plan scans them (one workgroup) into off / soff / lg and the
// first chunk of each kind; k_agg_fill writes the chunks.
struct AggPlan {
    int32_t D;
    int32_t ins0, ext0;
    int32_t pad;
};

struct AggTotals {
    int64_t nins, next, slots, sorts;
};

__global__ __launch_bounds__(kPlanThreads) void k_agg_plan(
    const int32_t *__restrict__ list, int32_t n, const int32_t *__restrict__ size,
    const int32_t *__restrict__ mstart, const int32_t *__restrict__ mscan, int32_t newV, int32_t ic,
    int32_t xc, int64_t *off, int64_t *soff, int32_t *lg, AggPlan *plan, AggTotals *tot) {
    __shared__ int64_t lds[kPlanThreads / 64];
    int64_t c_ins = 0, c_ext = 0, c_slot = 0, c_sort = 0;
    for (int32_t base = 0; base < n; base += kPlanThreads) {
        const int32_t i = base + (int32_t)threadIdx.x;
        const bool ok = i < n;
        int64_t D = 0, cap = 0;
        int l = 6;
        if (ok) {
            const int32_t c = list[i];
            D = mscan[mstart[c] + size[c]] - mscan[mstart[c]];
            const int64_t distinct_max = D < (int64_t)newV ? D : (int64_t)newV;
            while (((int64_t)1 << l) < 2 * distinct_max) ++l;
            cap = (int64_t)1 << l;
        }
        const int64_t i0 = plan_scan((D + ic - 1) / ic, c_ins, lds);
        const int64_t x0 = plan_scan((cap + xc - 1) / xc, c_ext, lds);
        const int64_t t0 = plan_scan(cap, c_slot, lds);
        const int64_t s0 = plan_scan(D, c_sort, lds);
        if (ok) {
            off[i] = t0;
            soff[i] = s0;
            lg[i] = l;
            plan[i] = AggPlan{(int32_t)D, (int32_t)i0, (int32_t)x0, 0};
        }
    }
    if (threadIdx.x == 0) *tot = AggTotals{c_ins, c_ext, c_slot, c_sort};
}

__global__ __launch_bounds__(256) void k_agg_fill(const int32_t *__restrict__ list,
                                                  const AggPlan *__restrict__ plan,
                                                  const int64_t *__restrict__ off,
                                                  const int32_t *__restrict__ lg, int32_t ic,
                                                  int32_t xc, lv::AggChunk *ins, lv::AggChunk *ext) {
    const int32_t i = blockIdx.x;
    const int32_t c = list[i];
    const AggPlan p = plan[i];
    const int64_t D = p.D;
    const int32_t ni = (int32_t)((D + ic - 1) / ic);
    for (int32_t j = threadIdx.x; j < ni; j += blockDim.x) {
        const int64_t f = (int64_t)j * ic;
        ins[p.ins0 + j] = lv::AggChunk{c, i, f, f + ic < D ? f + ic : D};
    }
    const int64_t cap = (int64_t)1 << lg[i];
    const int32_t nx = (int32_t)((cap + xc - 1) / xc);
    for (int32_t j = threadIdx.x; j < nx; j += blockDim.x) {
        const int64_t q = (int64_t)j * xc;
        ext[p.ext0 + j] = lv::AggChunk{c, i, off[i] + q, off[i] + (q + xc < cap ? q + xc : cap)};
    }
}

template <typename T>
lv_status dev_alloc(T **p, int64_t n) {
    *p = nullptr;
    if (n <= 0) n = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * (size_t)n);
    if (e != hipSuccess)
        return fail(LV_ERR_OOM, "hipMalloc of " + std::to_string(sizeof(T) * (size_t)n) +
                                    " bytes failed: " + hipGetErrorString(e));
    return LV_OK;
}

template <typename T>
void dev_free(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Growable device array.
template <typename T>
struct DevVec {
    T *p = nullptr;
    int64_t cap = 0;
    lv_status reserve(int64_t n) {
        if (n <= cap) return LV_OK;
        dev_free(p);
        cap = 0;
        int64_t c = std::max<int64_t>(n, 1024);
        LV_TRY(dev_alloc(&p, c));
        cap = c;
        return LV_OK;
    }
    void release() {
        dev_free(p);
        cap = 0;
    }
};

unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

// Control block of a device-resident phase (lv_graph::optimise_resident):
// the last kernel of every sweep applies the reference's termination test
// (modularity_optimisation.cu:308-354) to it, and the sweeps enqueued after
// the phase has ended find `done` set and do nothing.
struct PhaseCtl {
    int32_t done;     // the phase has ended (the gate of its sweeps' kernels)
    int32_t sweeps;   // sweeps run
    int32_t changed;  // some sweep had a gain > 0
    int32_t pad;
    double q_prev;    // modularity before the next sweep
    double q;         // after the last sweep run
    unsigned long long moved, moved_edges;  // movers of the last sweep run, their row entries
};
constexpr int kCtlRing = 4;  // pinned copies of the block, one per sweep in flight

// Weight inside communities of a singleton state: the level's self-loop
// weight (sown), into q[1]; the correction slots start at zero.
__global__ void k_singleton_own(double *q, const double *__restrict__ sown, double *own_slots) {
    const int i = threadIdx.x;
    if (i == 0) q[1] = *sown;
    if (i < lv::kMoveSlots) own_slots[(int64_t)i * lv::kMoveSlotStride] = 0.0;
}

__global__ void k_phase_start(PhaseCtl *c, const double *__restrict__ q) {
    c->done = 0;
    c->sweeps = 0;
    c->changed = 0;
    c->pad = 0;
    c->q_prev = q[0];
    c->q = q[0];
    c->moved = 0;
    c->moved_edges = 0;
}

// After sweep i: gain = Q after - Q before; a gain > 0 marks the phase as
// changing the state; the phase ends when nothing moved, the gain is not
// positive, or it is below min_gain (for min_gain > 0 the second test adds
// nothing; it ends the phase for min_gain <= 0, where the reference can loop
// forever).  The moves of the last sweep are kept, as the reference keeps them.
__global__ void k_phase_step(PhaseCtl *c, const double *__restrict__ q,
                             const unsigned long long *__restrict__ slots, float min_gain) {
    if (c->done) return;
    unsigned long long mv = 0, me = 0;
    for (int i = 0; i < lv::kMoveSlots; ++i) {
        mv += slots[(int64_t)i * lv::kMoveSlotStride];
        me += slots[(int64_t)i * lv::kMoveSlotStride + 1];
    }
    const double q1 = q[0];
    const double gain = q1 - c->q_prev;
    if (gain > 0.0) c->changed = 1;
    c->sweeps += 1;
    c->moved = mv;
    c->moved_edges = me;
    c->q = q1;
    c->q_prev = q1;
    if (mv == 0 || !(gain > 0.0) || !(gain >= (double)min_gain)) c->done = 1;
}

// Two HIP events, destroyed on every return path.
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() = default;
    EventPair(const EventPair &) = delete;
    EventPair &operator=(const EventPair &) = delete;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    hipError_t create() {
        const hipError_t e = hipEventCreate(&a);
        return e != hipSuccess ? e : hipEventCreate(&b);
    }
};

constexpr int kSlotChunk = 4096;  // hash slots per evaluation chunk
constexpr int kAggChunk = 4096;   // flat member edges per aggregation chunk
// Degree bounds of the one-workgroup buckets (tables of 1.5 x the bound):
constexpr int32_t kWgA = 512;       // 128 lanes, 768 slots
constexpr int32_t kWgSplit = 1024;  // 256 lanes, 1536 slots
constexpr int32_t kWgMid = 2048;    // 256 lanes, 2560 slots (1.25 x)
constexpr int32_t kWgMax = 4096;    // 1024 lanes, 6144 slots
// LDS attempt of the chunked tier's rows: 1024 lanes, 72 KB table
constexpr int kHeavyTrySlots = 6144;
static_assert(kHeavyTrySlots > lv::kHeavyTryLimit + 1024, "the attempt's table never fills");
// Group 5 (degrees 85..319) runs 256-slot wave tables up to this degree.
constexpr int32_t kWave256Hi = 128;
// levels from which the move kernels read bucket-ordered columns: comm
// (4 bytes a vertex) well beyond one XCD's 4 MB L2
constexpr int32_t kColOrderMinVertices = 1 << 20;
constexpr int32_t kXcdMinVertices = 1 << 21;  // XCD-local row passes from this level size
// candidate bounds from this level size; the small-table attempt skips input
// levels of this size
constexpr int32_t kBoundsMinVertices = 1 << 20;
// Neighbour moves are counted from the rows' side when the movers' row
// entries reach E / kPullDen (from the movers' side below that).
constexpr int64_t kPullDen = 16;
// A phase with a frontier sweeps only full after this many sweeps (bounds a
// phase whose pruned and full sweeps keep undoing each other).
constexpr int64_t kMaxPrunedSweeps = 4096;
// Internal frontier kind: the vertices of the active graphs of a batch.
constexpr int32_t kFrontierBatchGraphs = 100;
// Internal frontier kind: batches with the model; a graph's vertices follow
// the model when its sweep is pruned, all of them when it is full.
constexpr int32_t kFrontierBatchModel = 101;

int ceil_log2(int64_t x) {
    int lg = 0;
    while (((int64_t)1 << lg) < x) ++lg;
    return lg;
}

}  // namespace

void lv::set_error(const std::string &msg) { g_last_error = msg; }

struct lv_graph {
    int device = 0;
    hipStream_t stream = nullptr;
    lv_options opt{};
    int max_lds = 65536;

    // level 0 (kept) and two buffers for aggregated levels
    int32_t V0 = 0;
    // original vertices (length of orig2comm): V0, or those of the run a
    // level dump came from, whose map into level 0 of this handle is o2c0
    int32_t Vorig = 0;
    int32_t *o2c0 = nullptr;
    int64_t E0 = 0;
    int32_t *rp0 = nullptr, *col0 = nullptr;
    float *w0 = nullptr;
    int32_t *rpb[2] = {nullptr, nullptr}, *colb[2] = {nullptr, nullptr};
    float *wb[2] = {nullptr, nullptr};
    int cur_buf = -1;

    int32_t V = 0;
    int64_t E = 0;
    double M = 0.0;
    float M32 = 0.0f, den = 0.0f;
    bool level_ready = false;

    // community state
    double *k = nullptr, *sigma = nullptr;
    int32_t *comm = nullptr, *size = nullptr, *newc = nullptr, *orig2comm = nullptr;
    uint64_t *crec = nullptr;
    double *dk = nullptr;        // per mover: change of weight to own community
    float *margin = nullptr, *kown = nullptr;  // last evaluation of each vertex
    uint8_t *hist = nullptr;     // move history (bit 0 = last sweep)
    uint8_t *active = nullptr;   // frontier flags of a pruned sweep
    int32_t *nmv = nullptr;      // row entries whose neighbour moved in the last sweep
    int8_t *fkeys = nullptr;     // frontier bucket keys
    int32_t *forder = nullptr, *fd_off = nullptr;  // frontier lists by bucket
    int32_t h_foff[lv::kNumBuckets + 1] = {0};
    float *mlp = nullptr;        // frontier model parameters
    bool mlp_set = false;
    uint64_t sweep_count = 0;
    DevVec<float> tfeat;         // feature staging of traced sweeps

    // batch of independent graphs (block-diagonal union); G = 0: one graph
    int32_t G = 0;
    std::vector<int32_t> h_vptr0;
    std::vector<double> gM;
    int32_t *vptr = nullptr, *vptr_next = nullptr, *vptr0 = nullptr;
    int32_t *gid = nullptr, *gid0 = nullptr, *gfinal = nullptr;
    float *gM32 = nullptr, *gden = nullptr;
    uint8_t *gact = nullptr, *gdone = nullptr;
    double *gsums = nullptr, *own_v = nullptr;
    double *gin = nullptr;  // per graph: weight inside communities (incremental sweeps)
    uint8_t *gfull = nullptr;  // per graph: this sweep is full (batched pruned runs)
    // batched pruned runs: the vertices of positive degree of the active
    // graphs by bucket (rebuilt when the active set changes; the model and
    // the neighbour-move count run over it instead of the whole union)
    int32_t *glist = nullptr, *gl_off = nullptr;
    int32_t glist_n = 0;
    bool glist_ok = false;
    float *gvscale = nullptr;  // per graph: V_g / (2 M_g) at the current level (features)
    int32_t gmaxV = 0;         // vertices of the batch's largest graph (level 0)
    std::vector<double> h_gsums;     // per graph: own, sq, moved of the last sweep
    // pinned staging of per-graph flag uploads, one slot per destination: an
    // upload may still be pending when the next one (to another array) is
    // staged, since none waits for its copy
    uint8_t *h_up = nullptr;
    enum UpSlot { kUpAct = 0, kUpFull = 1, kUpDone = 2, kUpSlots = 3 };
    size_t up_stride() const { return ((size_t)G + 64 + 63) & ~(size_t)63; }

    // Per-graph flags to the device through pinned memory; the caller syncs
    // the stream before the next upload.
    lv_status upload_flags(uint8_t *dst, const std::vector<uint8_t> &src, UpSlot slot) {
        uint8_t *h = h_up + (size_t)slot * up_stride();
        std::memcpy(h, src.data(), src.size());
        LV_HIP(hipMemcpyAsync(dst, h, src.size(), hipMemcpyHostToDevice, stream));
        return LV_OK;
    }
    uint32_t *bits = nullptr;    // two mover bitmaps (alternating group steps), then the
                                 // marks of the neighbour-move count
    int32_t bit_words = 0;
    uint64_t step = 0;           // group steps applied with marks
    bool own_valid = false;      // d_q[1] holds the weight inside communities
    // Singleton state (every vertex alone, after a level's set-up or a
    // reset): its weight inside communities is the level's self-loop weight,
    // summed at the level's set-up into d_sown ([0] level 0, [1] the current
    // aggregated level; then partials), so that Q needs only the sum of
    // S_c^2.  Only with weights on the grid, where that sum is exact in any
    // order and so equals the full pass's.  Single graphs only.
    bool singleton_state = false;
    bool sown0_ok = false, sown_next_ok = false;
    double *d_sown = nullptr;
    // the current aggregated level's rows are sorted by column without
    // duplicates (the radix aggregation writes them so)
    bool rows_sorted = false;
    bool weights_nonneg = false; // every weight >= 0: candidate bounds are valid
    int64_t last_moved = -1;     // movers of the previous sweep (-1: unknown)
    int64_t last_moved_edges = -1;  // row entries of those movers
    int32_t incr_den = 2;        // incremental modularity when their rows hold <= E/incr_den entries

    // bucketing
    lv::BucketBounds bb{};
    // LDS slots allocated per launch of each one-workgroup bucket
    int wg_slots[4] = {768, 1536, 2560, 6144};
    int wg_lanes[4] = {128, 256, 256, 1024};  // lanes per row of each tier
    int sub_x[4] = {4, 4, 4, 4};              // lanes per vertex of buckets 0..3
    int table_pct = 0;                        // per-row table size, percent (0 = 150)
    // small-table attempt of the one-workgroup tiers' rows (plain sweeps):
    // per row of buckets 7..10 (from h_off[7]) its distinct communities at
    // its last evaluation (INT32_MAX: none yet) and whether the attempt
    // decided it this sweep
    bool hinted_on = true;
    int32_t nmv_pull = 0;  // lv_options.nmv_pull
    DevVec<int32_t> ghint, hdone_w;
    bool heavy_try = false;  // LDS attempt of the chunked tier's rows
    int32_t heavy_limit = lv::kHeavyTryLimit;  // ... given up at this many communities
    int chunk_edges = lv::kChunkEdges, part_target = lv::kPartTarget, merge_limit = lv::kMergeLimit;
    lv::AggBounds agg_bounds{};
    int8_t *keys8 = nullptr;
    int32_t *order = nullptr, *d_off = nullptr;
    int32_t h_off[lv::kNumBuckets + 1] = {0};
    int64_t h_bedges[lv::kNumBuckets] = {0};
    unsigned long long *d_bedges = nullptr;
    int32_t *scratch = nullptr;
    int64_t scratch_elems = 0;

    // chunked tier (last bucket)
    int32_t nh = 0, nins = 0, nitems = 0;
    DevVec<lv::Chunk> ins;
    DevVec<lv::HeavyVertex> hvs;
    DevVec<lv::MergeItem> items;
    DevVec<HeavyPlan> hplan;      // per heavy vertex: k_tier_fill's inputs
    DevVec<TierTotals> tier_tot;  // totals of the tables (read back to size them)
    DevVec<int32_t> out_key, poff, ptot;
    DevVec<double> out_val;
    DevVec<uint64_t> item_best;
    DevVec<double> item_K;
    DevVec<uint64_t> hbest;
    DevVec<double> hkcur;
    DevVec<int32_t> hovf;
    DevVec<int32_t> hdone, hhint, hdist;  // LDS attempt of the heavy rows
    DevVec<int32_t> corder, iorder;  // XCD-aligned dispatch orders of the chunked tier
    int32_t n_corder = 0, n_iorder = 0;
    bool tier_xcd = false;
    DevVec<int32_t> gkeys;  // global hash arena, kept all-empty between uses
    DevVec<double> gvals;

    // column numbering of the move kernels (lv::MoveArgs::ccol): positions
    // in `order`, the rows' columns in that numbering and comm copied into it
    bool col_shadow = false;
    DevVec<int32_t> c_pos, c_col, c_comm;
    DevVec<int32_t> c_cn;  // the aggregation's new ids in that column order

    // row-reduction segments
    lv::SegTable st{};
    // XCD-local modularity: light rows (buckets 0..3) by st_light, heavy rows
    // (buckets 4..7) by column block with the split table
    bool xcd_ok = false;
    lv::SegTable st_light{}, xst{};
    int32_t xh_lo = 0, xh_n = 0;
    DevVec<int32_t> xsplit;
    lv::SegTable gst[lv::kNumGroups]{};
    DevVec<double> qpart;
    double *d_q = nullptr;
    unsigned long long *d_moved = nullptr;
    double *d_own_slots = nullptr;   // the applies' modularity corrections (kept zero between sweeps)
    // device-resident phases: the control block, its pinned host copies, and
    // the gate the sweeps' kernels read (null outside such a phase)
    PhaseCtl *d_ctl = nullptr;
    PhaseCtl *h_ctl = nullptr;
    const int32_t *gate_ptr = nullptr;

    // Levels after the first may keep only their vertices of positive degree
    // (compact ids, ascending in the reference's ids).  ref_of/cmp_of map
    // compact <-> reference ids of the current level (map_cur); Vref counts
    // the reference's vertices.  With maps_identity both numberings agree.
    bool compact_on = true;
    bool maps_identity = true;
    int32_t Vref = 0;
    int32_t *ref_of[2] = {nullptr, nullptr}, *cmp_of[2] = {nullptr, nullptr};
    int map_cur = 0;
    int32_t *newid_ref = nullptr;  // reference renumbering at an aggregation
    // level 0 stored without its vertices of degree 0 (when it has some):
    // its maps, kept for reset()
    bool compact0 = false;
    int32_t Vref0 = 0;
    int32_t *ref_of0 = nullptr, *cmp_of0 = nullptr;
    int32_t *a_keep = nullptr;     // communities that become vertices of the next level

    // aggregation scratch
    int32_t *a_newid = nullptr, *a_mstart = nullptr, *a_fill = nullptr, *a_members = nullptr,
            *a_mscan = nullptr, *a_cnt = nullptr, *a_cstart = nullptr, *a_order = nullptr,
            *a_off = nullptr, *a_scalars = nullptr;
    int8_t *a_keys = nullptr;
    int32_t *tmp_col = nullptr;
    float *tmp_w = nullptr;
    DevVec<lv::AggChunk> a_ins, a_ext;
    DevVec<int64_t> a_toff, a_soff;
    DevVec<int32_t> a_lg;
    DevVec<AggPlan> a_plan;
    DevVec<AggTotals> a_tot;
    DevVec<uint64_t> sortA, sortB;
    // sort-based aggregation: key/value ping-pong, run heads and slots, sums
    DevVec<uint64_t> skA, skB;
    DevVec<float> svA, svB;
    DevVec<int32_t> shead, spos, sscratch;
    DevVec<double> sacc;

    // pinned host staging
    void *h_stage = nullptr;
    size_t h_stage_bytes = 0;

    // timing
    struct Rec {
        hipEvent_t a, b;
        int kid;       // LV_K_* (or -1)
        int kd;        // LV_KD_* (or -1)
        double bytes;
    };
    std::vector<hipEvent_t> ev_pool;
    std::vector<Rec> pending;
    lv_kernel_time ktime[LV_K_COUNT]{};
    lv_kernel_time ktd[LV_KD_COUNT]{};
    double memory_ms = 0.0;

    // ---------------------------------------------------------------- util
    lv_status stage(size_t bytes) {
        if (bytes <= h_stage_bytes) return LV_OK;
        if (h_stage) (void)hipHostFree(h_stage);
        h_stage = nullptr;
        h_stage_bytes = 0;
        size_t b = std::max<size_t>(bytes, 1 << 16);
        LV_HIP(hipHostMalloc(&h_stage, b, hipHostMallocDefault));
        h_stage_bytes = b;
        return LV_OK;
    }
    lv_status sync() {
        LV_HIP(hipStreamSynchronize(stream));
        return flush_timers();
    }
    hipEvent_t take_event() {
        if (!ev_pool.empty()) {
            hipEvent_t e = ev_pool.back();
            ev_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    hipEvent_t tbegin() {
        if (!opt.timing) return nullptr;
        hipEvent_t e = take_event();
        if (e) (void)hipEventRecord(e, stream);
        return e;
    }
    // Closes a timed region opened by tbegin: counted under kid (coarse
    // family, -1 = none) and kd (single kernel, -1 = none).
    void tend(hipEvent_t a, int kid, double bytes, int kd = -1) {
        if (!opt.timing || !a) return;
        hipEvent_t b = take_event();
        if (!b) return;
        (void)hipEventRecord(b, stream);
        pending.push_back({a, b, kid, kd, bytes});
    }
    lv_status flush_timers() {
        for (auto &r : pending) {
            float ms = 0.0f;
            LV_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            if (r.kid >= 0) {
                ktime[r.kid].launches++;
                ktime[r.kid].total_ms += ms;
                ktime[r.kid].bytes += r.bytes;
            }
            if (r.kd >= 0) {
                ktd[r.kd].launches++;
                ktd[r.kd].total_ms += ms;
                ktd[r.kd].bytes += r.bytes;
            }
            ev_pool.push_back(r.a);
            ev_pool.push_back(r.b);
        }
        pending.clear();
        return LV_OK;
    }

    const int32_t *rowptr() const { return cur_buf < 0 ? rp0 : rpb[cur_buf]; }
    const int32_t *col() const { return cur_buf < 0 ? col0 : colb[cur_buf]; }
    const float *w() const { return cur_buf < 0 ? w0 : wb[cur_buf]; }

    lv::MoveArgs move_args() const {
        lv::MoveArgs a;
        a.rowptr = rowptr();
        a.col = col();
        a.w = w();
        a.comm = comm;
        a.size = size;
        a.k = k;
        a.sigma = sigma;
        a.crec = crec;
        a.newc = newc;
        a.dk = dk;
        a.margin = margin;
        a.kown = kown;
        a.active = nullptr;
        a.M32 = M32;
        a.den = den;
        a.gid = G > 0 ? gid : nullptr;
        a.gM32 = G > 0 ? gM32 : nullptr;
        a.gden = G > 0 ? gden : nullptr;
        // bounds pay where the records' gathers miss L2: large levels (a
        // batch's gathers stay inside one graph's vertex range)
        const int32_t span = G > 0 ? gmaxV : V;
        a.bounds = weights_nonneg && (opt.candidate_bounds > 0 ||
                                      (opt.candidate_bounds == 0 && span >= kBoundsMinVertices))
                       ? 1
                       : 0;
        a.obs = 1;
        a.gate = gate_ptr;
        a.ccol = col_shadow ? c_col.p : a.col;
        a.ccomm = col_shadow ? c_comm.p : comm;
        a.cself = col_shadow ? c_pos.p : nullptr;
        return a;
    }

    // The state changed other than by a sweep: the tracked weight inside
    // communities, the mover bitmaps and the frontier observations start over.
    void state_changed(bool singletons = false) {
        if (col_shadow) lv::launch_col_comm(order, V, comm, c_comm.p, stream);
        lv::fill_i32(reinterpret_cast<int32_t *>(bits), 0, 2 * (int64_t)bit_words, stream);
        own_valid = false;
        singleton_state = singletons;
        last_moved = -1;
        last_moved_edges = -1;
        const int64_t Vp = (int64_t)V + 1;
        lv::fill_i32(reinterpret_cast<int32_t *>(margin), 0, Vp, stream);
        lv::fill_i32(reinterpret_cast<int32_t *>(kown), 0, Vp, stream);
        lv::fill_i32(nmv, 0, Vp, stream);
        lv::fill_i32(reinterpret_cast<int32_t *>(hist), 0, (Vp + 3) / 4, stream);
        // heavy rows: no distinct count known for the new state, so the next
        // sweep takes them through the chunked tier
        if (nh > 0) lv::fill_i32(hhint.p, 0x7fffffff, nh, stream);
    }

    lv::FeatureArgs feature_args() const {
        lv::FeatureArgs f;
        f.rowptr = rowptr();
        f.k = k;
        f.comm = comm;
        f.crec = crec;
        f.margin = margin;
        f.kown = kown;
        f.hist = hist;
        f.nmv = nmv;
        f.V = V;
        f.M = M32;
        // V of the reference's numbering (vertices without edges included)
        f.vscale = M > 0.0 ? (float)((double)Vref / (2.0 * M)) : 0.0f;
        f.ref_of = maps_identity ? nullptr : ref_of[map_cur];
        f.gid = G > 0 ? gid : nullptr;
        f.gM = G > 0 ? gM32 : nullptr;
        f.gvscale = G > 0 ? gvscale : nullptr;
        return f;
    }

    lv::GroupMarks marks() const {
        lv::GroupMarks gm;
        const int cur = (int)(step & 1);
        gm.movers = bits + (int64_t)cur * bit_words;
        gm.clear_movers = bits + (int64_t)(cur ^ 1) * bit_words;
        gm.words = bit_words;
        gm.rounds = 1;
        gm.gid = nullptr;
        gm.gin = nullptr;
        return gm;
    }

    // The group's correction table for `rounds` (block starts by
    // lv::corr_items).
    lv::SegTable corr_table(int g, int32_t rounds) const { return corr_of(gst[g], rounds); }
    // The correction blocks over a row table (a group's, or the level's st).
    static lv::SegTable corr_of(const lv::SegTable &base, int32_t rounds) {
        lv::SegTable t = base;
        int32_t gb = 0;
        for (int i = 0; i < t.nseg; ++i) {
            const int lg = t.s[i].lanes_log2;
            const int32_t per = lv::corr_items(lv::corr_lanes(lg), rounds);
            t.s[i].block_start = gb;
            gb += (t.s[i].count + per - 1) / per;
        }
        t.chunk_block_start = gb;
        const int32_t per = lv::corr_items(lv::kRowThreads, rounds);
        gb += (t.nchunks + per - 1) / per;
        t.sq_block_start = gb;
        t.total_blocks = gb;
        return t;
    }

    lv::ChunkedTier chunked_tier() {
        lv::ChunkedTier t;
        t.chunks = ins.p;
        t.hvs = hvs.p;
        t.items = items.p;
        t.out_key = out_key.p;
        t.out_val = out_val.p;
        t.item_best = item_best.p;
        t.item_K = item_K.p;
        t.poff = poff.p;
        t.ptot = ptot.p;
        t.gkeys = gkeys.p;
        t.gvals = gvals.p;
        t.best = hbest.p;
        t.kcur = hkcur.p;
        t.merge_limit = merge_limit;
        t.ovf = hovf.p;
        t.done = hdone.p;
        t.hint = hhint.p;
        t.hdist = hdist.p;
        t.chunk_order = tier_xcd ? corder.p : nullptr;
        t.item_order = tier_xcd ? iorder.p : nullptr;
        t.n_chunk_order = n_corder;
        t.n_item_order = n_iorder;
        t.range_blk = tier_xcd ? (V + 7) / 8 : 0;
        int clg = 6;  // table_log2's floor
        while ((1 << clg) < 2 * chunk_edges && clg < lv::kChunkCapLog2) ++clg;
        t.chunk_cap_log2 = clg;
        // input levels: a chunk's communities barely repeat there (on
        // aggregated levels rows that stay this diverse are rare, and merging
        // a chunk's duplicates keeps stage B's work down)
        t.direct = cur_buf < 0 ? 1 : 0;
        return t;
    }

    // Grow the global hash arena; new memory is cleared (the arena invariant).
    lv_status arena(int64_t slots) {
        if (slots <= gkeys.cap && slots <= gvals.cap) return LV_OK;
        gkeys.release();
        gvals.release();
        LV_TRY(gkeys.reserve(slots));
        LV_TRY(gvals.reserve(slots));
        lv::launch_clear_tables(gkeys.p, gvals.p, gkeys.cap, stream);
        LV_LAUNCHED("clear arena");
        return LV_OK;
    }

    // Tier bounds and sizes from opt (zero fields take the defaults).  They
    // change the kernel mapping only, never a result.
    void apply_options() {
        // LDS aggregation tiers the device can hold; the others stay empty.
        for (int t = 0; t < lv::kAggTiers - 1; ++t) {
            const bool fits = lv::merge_lds_bytes(t) <= max_lds;
            agg_bounds.hi[t] = fits ? lv::kAggTierMax[t] : (t > 0 ? agg_bounds.hi[t - 1] : 0);
        }
        if (opt.agg_lds_max > 0)
            for (int t = 0; t < lv::kAggTiers - 1; ++t)
                agg_bounds.hi[t] = std::min(agg_bounds.hi[t], opt.agg_lds_max);
        // Bucket bounds: 0..5 are the reference groups; group 6 splits into
        // four one-workgroup tiers and the chunked tier.  The upper tier's
        // 72 KB table needs that much LDS per workgroup.
        // the largest one-workgroup bound whose table (up to 2 x the bound,
        // 12 bytes a slot) fits a workgroup's LDS
        int32_t cap = kWgMax;
        while (cap > 320 && 24 * cap > max_lds) cap /= 2;
        int wg_max = opt.heavy_wg_max > 0 ? opt.heavy_wg_max : cap;
        wg_max = std::max(320, std::min(wg_max, cap));
        // group 6's tier bounds, ascending and at most wg_max (empty tiers
        // are allowed); group 5's split between the two wave tables
        auto bound = [&](int32_t v, int32_t def, int32_t lo) {
            return std::max(lo, std::min(wg_max, v > 0 ? v : def));
        };
        const int32_t wg_a = bound(opt.wg_bound_a, kWgA, 320);
        const int32_t wg_lo = bound(opt.wg_bound_b, kWgSplit, wg_a);
        const int32_t wg_mid = bound(opt.wg_bound_c, kWgMid, wg_lo);
        const int32_t wsplit =
            opt.wave_split > 0 ? std::max(85, std::min(255, opt.wave_split)) : kWave256Hi;
        const int32_t his[] = {4,   8,    16,    32,     84,     wsplit,
                               319, wg_a, wg_lo, wg_mid, wg_max, 0x7fffffff};
        static_assert(sizeof(his) / sizeof(his[0]) == lv::kNumBuckets, "one bound per bucket");
        for (int b = 0; b < lv::kNumBuckets; ++b) bb.hi[b] = his[b];
        // tables of 1.5 x the bucket's degree bound (load <= 2/3) by default,
        // 1.25 x for 1025..2048 (five workgroups per CU instead of four), in
        // 64-slot steps and at least 512: smaller tables, more rows in flight
        // per CU.  Every table holds more slots than its bound (>= 1.2 x).
        table_pct = opt.table_pct > 0 ? std::max(120, std::min(200, opt.table_pct)) : 0;
        auto slots_for = [&](int32_t b, int def_pct) {
            const int64_t p = table_pct ? table_pct : def_pct;
            return (int)std::max<int64_t>(512, ((int64_t)b * p / 100 + 63) & ~(int64_t)63);
        };
        wg_slots[0] = slots_for(wg_a, 150);
        wg_slots[1] = slots_for(wg_lo, 150);
        wg_slots[2] = slots_for(wg_mid, 125);
        wg_slots[3] = slots_for(wg_max, 150);
        static const int kDefLanes[4] = {128, 256, 256, 1024};
        for (int t = 0; t < 4; ++t) {
            const int l = opt.wg_lanes[t];
            wg_lanes[t] = (l == 128 || l == 256 || l == 512 || l == 1024) ? l : kDefLanes[t];
        }
        // sub-wave buckets 1..3 (degree bounds 8, 16, 32): lanes per vertex.
        // Four lanes (a quad: every exchange is DPP) with 2, 4 or 8 entries
        // each put 16 vertices in flight per wave.
        static const int kDefSub[3] = {4, 4, 4};
        for (int b = 1; b < 4; ++b) {
            const int l = opt.sub_lanes[b - 1];
            const int bound = 4 << b;
            sub_x[b] = (l >= 4 && l <= bound && (l & (l - 1)) == 0 &&
                        lv::move_sub_shape(l, bound / l)) ? l : kDefSub[b - 1];
        }
        hinted_on = opt.hinted >= 0;
        nmv_pull = opt.nmv_pull;
        heavy_try = opt.heavy_lds >= 0 && 12 * kHeavyTrySlots <= max_lds;
        heavy_limit = opt.heavy_lds > 0 ? std::min(opt.heavy_lds, lv::kHeavyTryLimit)
                                        : lv::kHeavyTryLimit;
        chunk_edges = opt.chunk_edges > 0 ? std::max(16, std::min(opt.chunk_edges, lv::kChunkEdges))
                                          : lv::kChunkEdges;
        part_target = opt.part_target > 0 ? std::max(16, opt.part_target) : lv::kPartTarget;
        merge_limit = opt.merge_limit > 0 ? std::min(opt.merge_limit, lv::kMergeLimit) : lv::kMergeLimit;
        incr_den = opt.q_incremental != 0 ? opt.q_incremental : 2;
        compact_on = opt.compact >= 0;
    }

    // ---------------------------------------------------------------- setup
    lv_status create(const int32_t *rowptr_h, const int32_t *col_h, const float *w_h, int32_t nV,
                     int64_t nE, const lv_options *o) {
        if (o) opt = *o;
        device = opt.device;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
            return fail(LV_ERR_NODEVICE, "no HIP device visible");
        if (device < 0 || device >= ndev) return fail(LV_ERR_ARG, "device ordinal out of range");
        LV_HIP(hipSetDevice(device));
        LV_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) ==
                hipSuccess &&
            lds > 0)
            max_lds = lds;
        lv::set_wg_lds_limits(max_lds);
        lv::set_merge_lds_limits(max_lds);
        (void)hipGetLastError();
        apply_options();
        V0 = V = nV;
        Vorig = nV;
        E0 = E = nE;

        // M = (sum of CSR weights)/2 in FP64, CSR order (utils.cu:36-46).
        double s = 0.0;
        for (int64_t e = 0; e < nE; ++e) s += (double)w_h[e];
        M = s / 2.0;
        M32 = (float)M;
        {
            volatile float two_m = 2.0f * M32;
            volatile float d = two_m * M32;
            den = d;
        }

        const int64_t Vp = (int64_t)nV + 1;
        LV_TRY(dev_alloc(&rp0, Vp));
        LV_TRY(dev_alloc(&col0, nE));
        LV_TRY(dev_alloc(&w0, nE));
        for (int i = 0; i < 2; ++i) {
            LV_TRY(dev_alloc(&rpb[i], Vp));
            LV_TRY(dev_alloc(&colb[i], nE));
            LV_TRY(dev_alloc(&wb[i], nE));
        }
        LV_TRY(dev_alloc(&k, Vp));
        LV_TRY(dev_alloc(&sigma, Vp));
        LV_TRY(dev_alloc(&comm, Vp));
        LV_TRY(dev_alloc(&size, Vp));
        LV_TRY(dev_alloc(&newc, Vp));
        LV_TRY(dev_alloc(&orig2comm, Vp));
        LV_TRY(dev_alloc(&crec, Vp));
        LV_TRY(dev_alloc(&dk, Vp));
        LV_TRY(dev_alloc(&margin, Vp));
        LV_TRY(dev_alloc(&kown, Vp));
        LV_TRY(dev_alloc(&hist, Vp + 8));
        LV_TRY(dev_alloc(&active, Vp + 8));
        LV_TRY(dev_alloc(&nmv, Vp));
        LV_TRY(dev_alloc(&fkeys, Vp));
        LV_TRY(dev_alloc(&forder, Vp));
        LV_TRY(dev_alloc(&fd_off, 64));
        LV_TRY(dev_alloc(&mlp, lv::kParams));
        bit_words = (int32_t)((Vp + 31) / 32);
        LV_TRY(dev_alloc(&bits, 3 * (int64_t)bit_words));
        LV_TRY(dev_alloc(&keys8, Vp));
        LV_TRY(dev_alloc(&order, Vp));
        LV_TRY(dev_alloc(&d_off, 64));
        LV_TRY(dev_alloc(&d_bedges, 16));
        scratch_elems = std::max(lv::partition_scratch_elems(Vp, lv::kNumBuckets),
                                 lv::scan_scratch_elems(Vp) + 8) +
                        64;
        LV_TRY(dev_alloc(&scratch, scratch_elems));
        LV_TRY(dev_alloc(&d_q, 8));
        LV_TRY(dev_alloc(&d_moved, lv::kMoveSlots * lv::kMoveSlotStride));
        LV_TRY(dev_alloc(&d_own_slots, lv::kMoveSlots * lv::kMoveSlotStride));
        LV_TRY(dev_alloc(&d_sown, 2 + lv::kSelfLoopBlocks));
        LV_TRY(dev_alloc(&d_ctl, 1));
        LV_HIP(hipHostMalloc(reinterpret_cast<void **>(&h_ctl), sizeof(PhaseCtl) * kCtlRing,
                             hipHostMallocDefault));
        lv::fill_f64(d_own_slots, 0.0, lv::kMoveSlots * lv::kMoveSlotStride, stream);
        LV_TRY(dev_alloc(&a_newid, Vp));
        LV_TRY(dev_alloc(&a_keep, Vp));
        LV_TRY(dev_alloc(&newid_ref, Vp + 1));
        for (int i = 0; i < 2; ++i) {
            LV_TRY(dev_alloc(&ref_of[i], Vp));
            LV_TRY(dev_alloc(&cmp_of[i], Vp));
        }
        LV_TRY(dev_alloc(&a_mstart, Vp));
        LV_TRY(dev_alloc(&a_fill, Vp));
        LV_TRY(dev_alloc(&a_members, Vp));
        LV_TRY(dev_alloc(&a_mscan, Vp + 1));
        LV_TRY(dev_alloc(&a_cnt, Vp));
        LV_TRY(dev_alloc(&a_cstart, Vp));
        LV_TRY(dev_alloc(&a_order, Vp));
        LV_TRY(dev_alloc(&a_off, 16));
        LV_TRY(dev_alloc(&a_scalars, 16));
        LV_TRY(dev_alloc(&a_keys, Vp));
        LV_TRY(dev_alloc(&tmp_col, nE));
        LV_TRY(dev_alloc(&tmp_w, nE));
        LV_TRY(stage(1 << 20));

        EventPair ev;
        LV_HIP(ev.create());
        hipEvent_t e0 = ev.a, e1 = ev.b;
        LV_HIP(hipEventRecord(e0, stream));
        LV_HIP(hipMemcpyAsync(rp0, rowptr_h, sizeof(int32_t) * Vp, hipMemcpyHostToDevice, stream));
        if (nE > 0) {
            LV_HIP(hipMemcpyAsync(col0, col_h, sizeof(int32_t) * nE, hipMemcpyHostToDevice, stream));
            LV_HIP(hipMemcpyAsync(w0, w_h, sizeof(float) * nE, hipMemcpyHostToDevice, stream));
        }
        Vref0 = nV;
        {
            int32_t *flag = a_scalars + 12;
            lv::fill_i32(flag, 0, 1, stream);
            if (nE > 0)
                hipLaunchKernelGGL(k_weight_sign, dim3((unsigned)std::min<int64_t>(blocks_for(nE, 256), 4096)),
                                   dim3(256), 0, stream, w0, (int64_t)nE, flag);
            LV_LAUNCHED("weight signs");
            LV_HIP(hipMemcpyAsync(h_stage, flag, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
            LV_HIP(hipStreamSynchronize(stream));
            int32_t bad = 1;
            std::memcpy(&bad, h_stage, sizeof bad);
            weights_nonneg = bad == 0;
        }
        LV_TRY(compact_level0(rowptr_h));
        LV_HIP(hipEventRecord(e1, stream));
        LV_HIP(hipStreamSynchronize(stream));
        float ms = 0.0f;
        LV_HIP(hipEventElapsedTime(&ms, e0, e1));
        memory_ms = ms;  // upload, and the compaction of level 0
        return reset();
    }

    // Level 0 keeps only its vertices of positive degree when some, but not
    // all, have none: rowptr compacted, columns renumbered (in place, on the
    // device), maps to the caller's ids kept for every returned array.
    lv_status compact_level0(const int32_t *rowptr_h) {
        if (!compact_on) return LV_OK;
        int32_t vc = 0;
        for (int32_t v = 0; v < V0; ++v) vc += rowptr_h[v + 1] > rowptr_h[v] ? 1 : 0;
        if (vc == 0 || vc == V0) return LV_OK;
        const int64_t Vp = (int64_t)V0 + 1;
        LV_TRY(dev_alloc(&ref_of0, Vp));
        LV_TRY(dev_alloc(&cmp_of0, Vp));
        hipStream_t s = stream;
        hipLaunchKernelGGL(k_degree_flags, dim3(blocks_for(V0, 256)), dim3(256), 0, s, rp0, V0,
                           a_newid);
        lv::exclusive_scan(a_newid, a_newid, V0, nullptr, scratch, s);
        hipLaunchKernelGGL(k_level0_maps, dim3(blocks_for(V0, 256)), dim3(256), 0, s, rp0, V0,
                           a_newid, ref_of0, cmp_of0);
        hipLaunchKernelGGL(k_compact_rowptr, dim3(blocks_for((int64_t)vc + 1, 256)), dim3(256), 0, s,
                           rp0, ref_of0, vc, (int32_t)E0, rpb[0]);
        LV_HIP(hipMemcpyAsync(rp0, rpb[0], sizeof(int32_t) * ((size_t)vc + 1),
                              hipMemcpyDeviceToDevice, s));
        if (E0 > 0)
            hipLaunchKernelGGL(k_relabel, dim3((unsigned)std::min<int64_t>(blocks_for(E0, 256), 65536)),
                               dim3(256), 0, s, col0, E0, cmp_of0);
        LV_LAUNCHED("level 0 compaction");
        compact0 = true;
        V0 = vc;
        return LV_OK;
    }

    // Back to the singleton state of level 0.  When level 0 is current and
    // set up, only the community state is rewritten (no re-bucketing).
    lv_status reset() {
        if (o2c0)
            LV_HIP(hipMemcpyAsync(orig2comm, o2c0, sizeof(int32_t) * (size_t)Vorig,
                                  hipMemcpyDeviceToDevice, stream));
        else
            lv::iota_i32(orig2comm, Vorig, stream);
        if (cur_buf < 0 && level_ready) {
            lv::launch_level_state(V, k, comm, size, sigma, crec, stream);
            state_changed(true);
            LV_LAUNCHED("reset");
            return LV_OK;
        }
        cur_buf = -1;
        rows_sorted = false;
        V = V0;
        E = E0;
        map_cur = 0;
        maps_identity = !compact0;
        Vref = Vref0;
        if (compact0) {
            LV_HIP(hipMemcpyAsync(ref_of[0], ref_of0, sizeof(int32_t) * (size_t)V0,
                                  hipMemcpyDeviceToDevice, stream));
            LV_HIP(hipMemcpyAsync(cmp_of[0], cmp_of0, sizeof(int32_t) * (size_t)Vref0,
                                  hipMemcpyDeviceToDevice, stream));
        }
        LV_LAUNCHED("reset");
        level_ready = false;
        col_shadow = false;
        return LV_OK;
    }

    lv_status setup_level() {
        glist_ok = false;
        hipStream_t s = stream;
        const int32_t *rp = rowptr();
        lv::launch_degree_keys(rp, V, bb, keys8, s);
        lv::stable_partition(keys8, V, lv::kNumBuckets, order, d_off, scratch, s);
        // frontier kernels write the entries of listed (positive-degree)
        // vertices only; the others stay out of every frontier
        lv::fill_u8(reinterpret_cast<uint8_t *>(fkeys), 0xFF, V, s);
        lv::fill_u8(active, 0, V, s);
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, d_bedges);
        for (int b = 1; b < lv::kNumBuckets; ++b)
            hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, d_bedges + b);
        if (V > 0)
            hipLaunchKernelGGL(k_bucket_edges, dim3(blocks_for(V, 256)), dim3(256), 0, s, rp, order,
                               d_off, lv::kNumBuckets, d_bedges);
        LV_LAUNCHED("bucketing");
        LV_TRY(stage(4096));
        LV_HIP(hipMemcpyAsync(h_stage, d_off, sizeof(int32_t) * (lv::kNumBuckets + 1),
                              hipMemcpyDeviceToHost, s));
        LV_HIP(hipMemcpyAsync((char *)h_stage + 256, d_bedges,
                              sizeof(unsigned long long) * lv::kNumBuckets, hipMemcpyDeviceToHost,
                              s));
        LV_TRY(sync());
        std::memcpy(h_off, h_stage, sizeof(int32_t) * (lv::kNumBuckets + 1));
        {
            unsigned long long tmp[lv::kNumBuckets];
            std::memcpy(tmp, (char *)h_stage + 256, sizeof tmp);
            for (int b = 0; b < lv::kNumBuckets; ++b) h_bedges[b] = (int64_t)tmp[b];
        }

        // split table of the heavy rows (XCD-local passes); needs only h_off
        LV_TRY(setup_xcd());

        // chunked tier: row chunks (stage A, strength, modularity), per-vertex
        // partitions (stage B) and overflow tables, planned and written on the
        // device (k_tier_plan, k_tier_fill); only their totals come back
        const int HB = lv::kNumBuckets - 1;
        nh = h_off[HB + 1] - h_off[HB];
        nins = nitems = 0;
        tier_xcd = false;
        n_corder = n_iorder = 0;
        if (nh > 0) {
            // with sorted rows: chunks cut at the column-block boundaries and
            // partitions by id range, both in an XCD-aligned dispatch order
            // (measured slower than the hash partitions on RMAT-22: opt-in)
            tier_xcd = xcd_ok && opt.xcd_rows == 2;
            const int32_t *split = tier_xcd ? xsplit.p + 9 * (int64_t)(h_off[HB] - xh_lo) : nullptr;
            LV_TRY(hvs.reserve(nh));
            LV_TRY(hplan.reserve(nh));
            LV_TRY(tier_tot.reserve(1));
            if (split)
                hipLaunchKernelGGL(k_tier_plan<true>, dim3(1), dim3(kPlanThreads), 0, s, rp,
                                   order + h_off[HB], nh, (int32_t)chunk_edges,
                                   (int32_t)part_target, split, hvs.p, hplan.p, tier_tot.p);
            else
                hipLaunchKernelGGL(k_tier_plan<false>, dim3(1), dim3(kPlanThreads), 0, s, rp,
                                   order + h_off[HB], nh, (int32_t)chunk_edges,
                                   (int32_t)part_target, split, hvs.p, hplan.p, tier_tot.p);
            LV_LAUNCHED("chunked tier plan");
            LV_TRY(stage(sizeof(TierTotals)));
            LV_HIP(hipMemcpyAsync(h_stage, tier_tot.p, sizeof(TierTotals), hipMemcpyDeviceToHost, s));
            LV_TRY(sync());
            TierTotals tt;
            std::memcpy(&tt, h_stage, sizeof tt);
            if (tt.nchunks > INT32_MAX || tt.nitems > INT32_MAX || tt.pbase > INT32_MAX)
                return fail(LV_ERR_ARG, "chunked tier tables exceed 2^31 entries (raise chunk_edges)");
            nins = (int32_t)tt.nchunks;
            nitems = (int32_t)tt.nitems;
            const int64_t pbase = tt.pbase, tbase = tt.nitems, slots = tt.slots, outs = tt.outs;
            if (tier_xcd) {
                int64_t Lc = 0, Li = 0;
                for (int x = 0; x < 8; ++x) {
                    Lc = std::max(Lc, tt.xc[x]);
                    Li = std::max(Li, tt.xi[x]);
                }
                n_corder = (int32_t)(8 * Lc);
                n_iorder = (int32_t)(8 * Li);
                LV_TRY(corder.reserve(n_corder));
                LV_TRY(iorder.reserve(n_iorder));
                lv::fill_i32(corder.p, -1, n_corder, s);
                lv::fill_i32(iorder.p, -1, n_iorder, s);
            }
            LV_TRY(arena(slots));
            LV_TRY(ins.reserve(nins));
            LV_TRY(items.reserve(nitems));
            LV_TRY(out_key.reserve(outs));
            LV_TRY(out_val.reserve(outs));
            LV_TRY(item_best.reserve(nitems));
            LV_TRY(item_K.reserve(nitems));
            LV_TRY(poff.reserve(pbase));
            LV_TRY(ptot.reserve(tbase));
            LV_TRY(hbest.reserve(nh));
            LV_TRY(hkcur.reserve(nh));
            LV_TRY(hovf.reserve(nh));
            LV_TRY(hdone.reserve(nh));
            LV_TRY(hhint.reserve(nh));
            LV_TRY(hdist.reserve(nh));
            hipLaunchKernelGGL(k_tier_fill, dim3(nh), dim3(256), 0, s, hvs.p, hplan.p,
                               (int32_t)chunk_edges, split, ins.p, items.p, corder.p, iorder.p);
            LV_LAUNCHED("chunked tier fill");
            lv::fill_i32(ptot.p, 0, tbase, s);
            lv::fill_f64(hkcur.p, 0.0, nh, s);
            lv::fill_i32(hovf.p, 0, nh, s);
            lv::fill_i32(hdone.p, 0, nh, s);
            lv::fill_i32(hdist.p, 0, nh, s);
            hipLaunchKernelGGL(k_zero_u64_n, dim3(blocks_for(nh, 256)), dim3(256), 0, s,
                               reinterpret_cast<unsigned long long *>(hbest.p), (int64_t)nh);
            LV_LAUNCHED("chunked tier tables");
        }

        // hints of the small-table attempt: none yet at this level
        {
            const int32_t n7 = h_off[11] - h_off[7];
            if (n7 > 0) {
                LV_TRY(ghint.reserve(n7));
                LV_TRY(hdone_w.reserve(n7));
                lv::fill_i32(ghint.p, 0x7fffffff, n7, s);
                lv::fill_i32(hdone_w.p, 0, n7, s);
            }
        }

        // row-reduction segments: lanes per vertex by bucket
        static const int lanes_lg[] = {2, 3, 4, 5, 6, 6, 6, 8, 8, 8, 8};
        static_assert(sizeof(lanes_lg) / sizeof(lanes_lg[0]) == lv::kNumBuckets - 1,
                      "one entry per bucket below the chunked tier");
        st = make_segs(h_off, nins);
        // (also the self-loop partials of the strength pass: one per block
        // and kSelfLoopBlocks folds)
        LV_TRY(qpart.reserve(std::max<int64_t>(
            2 * ((int64_t)st.total_blocks + st_light.total_blocks + 8 * (int64_t)xst.total_blocks) +
                2 * 256,
            (int64_t)st.total_blocks + lv::kSelfLoopBlocks)));
        // per-group segments for the modularity correction of the applies
        // (block starts are set per sweep by corr_table)
        for (int g = 0; g < lv::kNumGroups; ++g) {
            lv::SegTable &t = gst[g];
            t = lv::SegTable{};
            int32_t gb = 0;
            for (int b = 0; b < HB; ++b) {
                if (lv::kBucketGroup[b] != g) continue;
                const int32_t cnt = h_off[b + 1] - h_off[b];
                if (cnt <= 0) continue;
                lv::Seg sg;
                sg.list_off = h_off[b];
                sg.count = cnt;
                sg.block_start = gb;
                sg.lanes_log2 = lanes_lg[b];
                gb += cnt;
                t.s[t.nseg++] = sg;
            }
            if (t.nseg == 0) {
                t.s[0] = lv::Seg{0, 0, 0, 2};
                t.nseg = 1;
            }
            t.chunk_block_start = gb;
            t.nchunks = lv::kBucketGroup[HB] == g ? nins : 0;
            gb += t.nchunks;
            t.sq_block_start = gb;
            t.total_blocks = gb;
        }

        // column numbering of the move kernels (and of the modularity row pass)
        col_shadow = false;
        if (col_order_wanted()) {
            LV_TRY(c_pos.reserve(V));
            LV_TRY(c_col.reserve(E));
            LV_TRY(c_comm.reserve(V));
            lv::launch_col_numbering(order, V, col(), E, c_pos.p, c_col.p, s);
            LV_LAUNCHED("column numbering");
            col_shadow = true;
        }

        // strengths and singleton state (utils.cu:84-89; computeEdgesSum)
        // and the level's self-loop weight (the singleton state's weight
        // inside communities): summed with the strengths, or by a binary
        // search per row on levels the aggregation left sorted; level 0's CSR
        // never changes, so it is summed once
        const bool lvl0 = cur_buf < 0;
        const bool loops = G == 0 && opt.exact_sums >= 0 && !(lvl0 && sown0_ok);
        const bool loops_sorted = loops && rows_sorted && !lvl0;
        lv::fill_f64(k, 0.0, V, s);
        if (loops && !loops_sorted)
            lv::launch_strength(rp, w(), order, st, ins.p, k, s, col(), qpart.p,
                                d_sown + (lvl0 ? 0 : 1));
        else
            lv::launch_strength(rp, w(), order, st, ins.p, k, s);
        if (loops_sorted)
            lv::launch_selfloop_weight(rp, col(), w(), V, d_sown + 2, d_sown + 1, s);
        LV_LAUNCHED("strengths and self-loop weight");
        if (G == 0 && opt.exact_sums >= 0) (lvl0 ? sown0_ok : sown_next_ok) = true;
        lv::launch_level_state(V, k, comm, size, sigma, crec, s);
        state_changed(true);
        LV_LAUNCHED("level state");
        if (G > 0) LV_TRY(update_graph_scales());
        level_ready = true;
        return LV_OK;
    }

    // Batches: each graph's V_g / (2 M_g) at the current level (the model's
    // community-total feature uses its own graph's numbers).
    lv_status update_graph_scales() {
        std::vector<int32_t> vp((size_t)G + 1);
        LV_HIP(hipMemcpyAsync(vp.data(), vptr, sizeof(int32_t) * ((size_t)G + 1),
                              hipMemcpyDeviceToHost, stream));
        LV_TRY(sync());
        std::vector<float> sc((size_t)G);
        for (int32_t g = 0; g < G; ++g)
            sc[(size_t)g] = gM[(size_t)g] > 0.0
                                ? (float)((double)(vp[(size_t)g + 1] - vp[(size_t)g]) / (2.0 * gM[(size_t)g]))
                                : 0.0f;
        LV_HIP(hipMemcpyAsync(gvscale, sc.data(), sizeof(float) * (size_t)G, hipMemcpyHostToDevice,
                              stream));
        return sync();
    }

    // Whether the level's move kernels read bucket-ordered columns: large
    // single-graph levels whose vertices all have entries (all of them are
    // in `order`).  Needs h_off.
    bool col_order_wanted() const {
        return G == 0 && opt.col_order >= 0 && V > 0 && h_off[lv::kNumBuckets] == V &&
               (opt.col_order > 0 || V >= kColOrderMinVertices);
    }

    // Split table and segment tables of the XCD-local modularity pass; off
    // when disabled, when the level has no heavy rows or when one is unsorted,
    // and by default on levels with bucket-ordered columns: the plain row
    // pass reads those (their hub neighbours share lines) and was faster
    // than the column-block split at RMAT-22 (861 vs 1068 us a pass).
    lv_status setup_xcd() {
        xcd_ok = false;
        st_light = lv::SegTable{};
        xst = lv::SegTable{};
        constexpr int XB = 4;  // first heavy bucket (degree > 32)
        xh_lo = h_off[XB];
        xh_n = h_off[lv::kNumBuckets] - h_off[XB];
        // worth it only when comm (4V bytes) is well beyond one XCD's 4 MB
        // L2; below that the plain row pass already hits L2 often
        if (opt.xcd_rows < 0 || xh_n <= 0 || V < 8) return LV_OK;
        if (opt.xcd_rows == 0 && (V < kXcdMinVertices || col_order_wanted())) return LV_OK;
        hipStream_t s = stream;
        LV_TRY(xsplit.reserve(9 * (int64_t)xh_n));
        int32_t *flag = a_scalars + 8;
        lv::fill_i32(flag, 0, 1, s);
        lv::launch_xcd_splits(rowptr(), col(), order + xh_lo, xh_n, V, xsplit.p, flag, s);
        LV_LAUNCHED("xcd splits");
        LV_TRY(stage(64));
        LV_HIP(hipMemcpyAsync(h_stage, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        LV_TRY(sync());
        int32_t unsorted = 1;
        std::memcpy(&unsorted, h_stage, sizeof unsorted);
        if (unsorted) return LV_OK;
        static const int lanes_light[XB] = {2, 3, 4, 5};
        int32_t blk = 0;
        for (int b = 0; b < XB; ++b) {
            const int32_t cnt = h_off[b + 1] - h_off[b];
            if (cnt <= 0) continue;
            const int per = lv::kRowThreads >> lanes_light[b];
            st_light.s[st_light.nseg++] = lv::Seg{h_off[b], cnt, blk, lanes_light[b]};
            blk += (cnt + per - 1) / per;
        }
        if (st_light.nseg == 0) {
            st_light.s[0] = lv::Seg{0, 0, 0, 2};
            st_light.nseg = 1;
        }
        st_light.chunk_block_start = blk;
        st_light.nchunks = 0;
        st_light.sq_block_start = blk;
        blk += (int32_t)((V + lv::kSqItems - 1) / lv::kSqItems);
        st_light.total_blocks = blk;
        // heavy rows: a column-block segment of a row has about deg/8 entries
        static const int lanes_heavy[] = {3, 4, 4, 6, 6, 8, 8, 8};
        static_assert(sizeof(lanes_heavy) / sizeof(lanes_heavy[0]) == lv::kNumBuckets - XB,
                      "one entry per heavy bucket");
        int32_t j = 0;
        for (int b = XB; b < lv::kNumBuckets; ++b) {
            const int32_t cnt = h_off[b + 1] - h_off[b];
            if (cnt <= 0) continue;
            const int lg = lanes_heavy[b - XB];
            const int per = lg >= 8 ? 1 : (lv::kRowThreads >> lg);
            xst.s[xst.nseg++] = lv::Seg{h_off[b] - xh_lo, cnt, j, lg};
            j += (cnt + per - 1) / per;
        }
        xst.chunk_block_start = j;
        xst.nchunks = 0;
        xst.sq_block_start = j;
        xst.total_blocks = j;
        xcd_ok = true;
        return LV_OK;
    }

    // Row-reduction segments over a bucket-ordered list with offsets off
    // (buckets below the chunked tier), then the chunk blocks and the
    // sum-of-squares blocks.
    lv::SegTable make_segs(const int32_t *off, int32_t nchunks) const {
        static const int lanes[] = {2, 3, 4, 5, 6, 6, 6, 8, 8, 8, 8};
        static_assert(sizeof(lanes) / sizeof(lanes[0]) == lv::kNumBuckets - 1,
                      "one entry per bucket below the chunked tier");
        lv::SegTable t{};
        int32_t blk = 0;
        for (int b = 0; b < lv::kNumBuckets - 1; ++b) {
            const int32_t cnt = off[b + 1] - off[b];
            if (cnt <= 0) continue;
            const int per = lanes[b] >= 8 ? 1 : (lv::kRowThreads >> lanes[b]);
            t.s[t.nseg++] = lv::Seg{off[b], cnt, blk, lanes[b]};
            blk += (cnt + per - 1) / per;
        }
        if (t.nseg == 0) {  // keep the segment search well-defined
            t.s[0] = lv::Seg{0, 0, 0, 2};
            t.nseg = 1;
        }
        t.chunk_block_start = blk;
        t.nchunks = nchunks;
        blk += nchunks;
        t.sq_block_start = blk;
        blk += (int32_t)((V + lv::kSqItems - 1) / lv::kSqItems);
        t.total_blocks = std::max(blk, 1);
        return t;
    }

    // ---------------------------------------------------------------- phase 1
    // Move kernels of group g; pruned: only the frontier lists (forder by
    // h_foff), heavy rows by their active flags.
    // obs: the evaluations' margins must be exact (a frontier or the caller
    // reads them); otherwise candidates that cannot move the vertex are
    // skipped without reading their records.
    lv_status run_group(int g, bool pruned = false, bool obs = true) {
        hipStream_t s = stream;
        lv::MoveArgs a = move_args();
        a.obs = obs ? 1 : 0;
        if (pruned) a.active = active;
        // (not on large input levels, whose rows see nearly as many
        // communities as entries: no row would qualify)
        const bool hinted_here = hinted_on && !pruned && !(cur_buf < 0 && V >= kBoundsMinVertices);
        if (g == lv::kBucketGroup[7] && hinted_here && h_off[12] > h_off[7]) {
            // plain sweeps: the small-table attempt over the rows of the four
            // one-workgroup tiers (hints and flags indexed from bucket 7) and
            // of the chunked tier (its own hints and flags), before their
            // kernels, which skip the rows it decides
            hipEvent_t th = tbegin();
            const int32_t nh2 = heavy_try ? nh : 0;
            lv::launch_move_hinted(a, order + h_off[7], h_off[11] - h_off[7], ghint.p, hdone_w.p,
                                   nh2, hhint.p, hdone.p, s);
            tend(th, LV_K_MOVE_WG, 0.0, LV_KD_HINTED);
        }
        for (int b = 0; b < lv::kNumBuckets; ++b) {
            if (lv::kBucketGroup[b] != g) continue;
            const int32_t nfull = h_off[b + 1] - h_off[b];
            const int32_t n = pruned ? h_foff[b + 1] - h_foff[b] : nfull;
            if (n <= 0) continue;
            const int32_t *list = pruned ? forder + h_foff[b] : order + h_off[b];
            // pruned: edges estimated from the bucket's mean degree
            const double bytes =
                12.0 * (double)h_bedges[b] * ((double)n / (double)nfull) + 32.0 * (double)n;
            hipEvent_t t0 = nullptr;
            switch (b) {
            case 0: case 1: case 2: case 3:
                t0 = tbegin();
                lv::launch_move_sub(sub_x[b], (4 << b) / sub_x[b], a, list, n, s);
                tend(t0, LV_K_MOVE_SUB, bytes, LV_KD_MOVE_BUCKET0 + b);
                break;
            case 4: case 5: case 6:
                t0 = tbegin();
                // degrees <= 128 (bucket 4, and bucket 5 unless its bound is
                // raised): two rows per wave
                lv::launch_move_wave(b == 4 ? 128 : b == 5 ? (bb.hi[5] > 128 ? 256 : 192) : 448,
                                     a, list, n, s);
                tend(t0, LV_K_MOVE_WAVE, bytes, LV_KD_MOVE_BUCKET0 + b);
                break;
            case 7: case 8: case 9: case 10: {
                static const int kd[4] = {LV_KD_MOVE_BUCKET0 + 7, LV_KD_MOVE_BUCKET8,
                                          LV_KD_MOVE_BUCKET8 + 1, LV_KD_MOVE_BUCKET8 + 2};
                // hints and flags of the small-table attempt (run above)
                const bool hinted = hinted_here;
                int32_t *hp = hinted ? ghint.p + (h_off[b] - h_off[7]) : nullptr;
                int32_t *dp = hinted ? hdone_w.p + (h_off[b] - h_off[7]) : nullptr;
                t0 = tbegin();
                lv::launch_move_wg(b - 7, wg_lanes[b - 7], wg_slots[b - 7],
                                   table_pct ? table_pct : 150, a, list, n, hp, dp, s);
                tend(t0, LV_K_MOVE_WG, bytes, kd[b - 7]);
                break;
            }
            default: {
                // chunked tier: the family and each of its three kernels;
                // stage A streams the heavy rows (col, w, comm[col])
                const lv::ChunkedTier t = chunked_tier();
                t0 = tbegin();
                if (heavy_try) {
                    // rows whose last evaluation saw few distinct communities
                    // are decided in LDS; the stages below skip them.  The
                    // full heavy list (index = heavy vertex), frontier by flags.
                    hipEvent_t tt = tbegin();
                    lv::launch_heavy_try(kHeavyTrySlots, heavy_limit, a, order + h_off[b], nfull,
                                         t, s);
                    tend(tt, -1, 0.0, LV_KD_HEAVY_LDS);
                }
                hipEvent_t ta = tbegin();
                lv::launch_chunk_aggregate(a, t, nins, s);
                tend(ta, -1, 12.0 * (double)h_bedges[b] * ((double)n / (double)nfull),
                     LV_KD_CHUNK_AGGREGATE);
                hipEvent_t tb = tbegin();
                lv::launch_chunk_merge(a, t, nitems, s);
                tend(tb, -1, 12.0 * (double)h_bedges[b] * ((double)n / (double)nfull),
                     LV_KD_CHUNK_MERGE);
                hipEvent_t tc = tbegin();
                lv::launch_heavy_final(a, t, nh, s);
                tend(tc, -1, 0.0, LV_KD_CHUNK_FINAL);
                tend(t0, LV_K_MOVE_GLOBAL, bytes);
                break;
            }
            }
        }
        LV_LAUNCHED("move kernels");
        return LV_OK;
    }

    // Range of group g's buckets in a bucket-ordered list with offsets off.
    void group_range(int g, int32_t &lo, int32_t &hi, const int32_t *off = nullptr) const {
        if (!off) off = h_off;
        lo = -1;
        hi = -1;
        for (int b = 0; b < lv::kNumBuckets; ++b) {
            if (lv::kBucketGroup[b] != g) continue;
            if (lo < 0) lo = off[b];
            hi = off[b + 1];
        }
    }

    lv_status enqueue_modularity() {
        const bool lvl0 = cur_buf < 0;
        if (singleton_state && (lvl0 ? sown0_ok : sown_next_ok)) {
            // singletons: own = the level's self-loop weight; S_c^2 by the
            // same blocks and order as a full pass, so Q is the same bits
            hipEvent_t t0 = tbegin();
            hipLaunchKernelGGL(k_singleton_own, dim3(1), dim3(64), 0, stream, d_q,
                               d_sown + (lvl0 ? 0 : 1), d_own_slots);
            lv::launch_modularity_from_own(sigma, V, M, qpart.p, d_q, d_own_slots, stream);
            tend(t0, LV_K_MODULARITY, 8.0 * (double)V);
            LV_LAUNCHED("modularity");
            own_valid = true;
            return LV_OK;
        }
        return enqueue_modularity_full();
    }

    lv_status enqueue_modularity_full() {
        if (xcd_ok) {
            double e_light = 0.0, e_heavy = 0.0;
            for (int b = 0; b < lv::kNumBuckets; ++b) (b < 4 ? e_light : e_heavy) += (double)h_bedges[b];
            hipEvent_t t0 = tbegin();
            hipEvent_t tr = tbegin();
            lv::launch_modularity_rows(rowptr(), col(), w(), comm, sigma, V, order, st_light, ins.p,
                                       qpart.p, stream, col_shadow ? c_col.p : nullptr,
                                       c_comm.p);
            tend(tr, -1, 12.0 * e_light + 16.0 * (double)V, LV_KD_MODULARITY_ROWS);
            hipEvent_t tx = tbegin();
            lv::launch_modularity_xcd(col(), w(), comm, order + xh_lo, xst, xsplit.p, qpart.p,
                                      st_light.total_blocks, stream);
            tend(tx, -1, 12.0 * e_heavy, LV_KD_MODULARITY_XCD);
            lv::launch_modularity_reduce(st_light.total_blocks + 8 * xst.total_blocks,
                                         st_light.sq_block_start,
                                         st_light.total_blocks - st_light.sq_block_start, M,
                                         qpart.p, d_q, stream);
            tend(t0, LV_K_MODULARITY, 12.0 * (double)E + 16.0 * (double)V);
            LV_LAUNCHED("modularity");
            own_valid = true;
            return LV_OK;
        }
        hipEvent_t t0 = tbegin();
        hipEvent_t tr = tbegin();
        lv::launch_modularity_rows(rowptr(), col(), w(), comm, sigma, V, order, st, ins.p, qpart.p,
                                   stream, col_shadow ? c_col.p : nullptr, c_comm.p);
        tend(tr, -1, 12.0 * (double)E + 16.0 * (double)V, LV_KD_MODULARITY_ROWS);
        lv::launch_modularity_finish(st, M, qpart.p, d_q, stream);
        tend(t0, LV_K_MODULARITY, 12.0 * (double)E + 16.0 * (double)V);
        LV_LAUNCHED("modularity");
        own_valid = true;
        return LV_OK;
    }

    lv_status modularity(double *q) {
        if (!level_ready) LV_TRY(setup_level());
        LV_TRY(enqueue_modularity());
        LV_TRY(stage(64));
        LV_HIP(hipMemcpyAsync(h_stage, d_q, sizeof(double), hipMemcpyDeviceToHost, stream));
        LV_TRY(sync());
        std::memcpy(q, h_stage, sizeof(double));
        return LV_OK;
    }

    // Frontier of a pruned sweep from the state after the last sweep:
    // active flags and bucket keys, compacted into forder by bucket; the
    // vertices left out keep their community (newc = comm).
    lv_status select_frontier(const lv_frontier &fr) {
        hipStream_t s = stream;
        const lv::FeatureArgs fa = feature_args();
        if (fr.kind == kFrontierBatchGraphs) {
            lv::launch_graph_frontier(rowptr(), V, gid, gact, bb, fkeys, active, s);
        } else if (fr.kind == kFrontierBatchModel) {
            if (!mlp_set) return fail(LV_ERR_STATE, "frontier model not set (lv_batch_set_model)");
            // the vertices outside glist (inactive graphs) keep key -1 from
            // the list's build
            if (!glist_ok) return fail(LV_ERR_STATE, "graph list not built");
            lv::launch_frontier(fa, glist, glist_n, mlp, fr.threshold, fr.heavy_threshold,
                                fr.heavy_degree, bb, fkeys, active, nullptr, s, fr.and_neighbor);
            lv::launch_graph_model_frontier(rowptr(), glist, glist_n, gid, gact, gfull, bb, fkeys,
                                            active, s);
        } else if (fr.kind == LV_FRONTIER_MODEL) {
            if (!mlp_set) return fail(LV_ERR_STATE, "frontier model not set (lv_set_model)");
            hipEvent_t t0 = tbegin();
            lv::launch_frontier(fa, order, h_off[lv::kNumBuckets], mlp, fr.threshold,
                                fr.heavy_threshold, fr.heavy_degree, bb, fkeys, active, nullptr,
                                s, fr.and_neighbor);
            // per vertex: rowptr 8, k 8, comm 4, record 8, margin 4, kown 4,
            // hist 1, nmv 4 read; keys 1, active 1 written
            tend(t0, -1, 43.0 * (double)V, LV_KD_FRONTIER);
        } else {
            lv::launch_frontier_rule(fr.kind, fa, order, h_off[lv::kNumBuckets], fr.random_p,
                                     fr.seed + 0x9e37ull * sweep_count, bb, fkeys, active, s);
        }
        lv::stable_partition(fkeys, V, lv::kNumBuckets, forder, fd_off, scratch, s);
        if (V > 0)
            LV_HIP(hipMemcpyAsync(newc, comm, sizeof(int32_t) * V, hipMemcpyDeviceToDevice, s));
        LV_LAUNCHED("frontier");
        LV_TRY(stage(4096));
        LV_HIP(hipMemcpyAsync(h_stage, fd_off, sizeof(int32_t) * (lv::kNumBuckets + 1),
                              hipMemcpyDeviceToHost, s));
        LV_TRY(sync());
        std::memcpy(h_foff, h_stage, sizeof(int32_t) * (lv::kNumBuckets + 1));
        return LV_OK;
    }

    // One sweep: groups 0..6, each = moves then apply; then Q.  After a sweep
    // that moved few vertices the applies carry the weight inside
    // communities forward (correction over the movers' rows) and Q needs
    // only the sum of S_c^2; otherwise a full modularity pass follows.
    // fr: evaluate only the frontier it selects (pruned sweep).  track: count
    // the neighbour moves afterwards (a frontier feature).
    lv_status sweep(double *q_after, int64_t *moved, double *sweep_ms,
                    const lv_frontier *fr = nullptr, bool track = false,
                    int64_t *evaluated = nullptr) {
        if (!level_ready) LV_TRY(setup_level());
        hipStream_t s = stream;
        // the correction walks the movers' rows; a full pass walks all rows
        const bool incremental = incr_den > 0 && own_valid && last_moved_edges >= 0 &&
                                 last_moved_edges * (int64_t)incr_den <= E;
        hipEvent_t e0 = take_event(), e1 = take_event();
        if (e0) (void)hipEventRecord(e0, s);
        const bool pruned = fr != nullptr && fr->kind != LV_FRONTIER_ALL;
        if (pruned) LV_TRY(select_frontier(*fr));
        if (evaluated)
            *evaluated = pruned ? h_foff[lv::kNumBuckets] : h_off[lv::kNumBuckets];
        LV_TRY(enqueue_sweep(pruned ? fr : nullptr, track, incremental, last_moved));
        if (e1) (void)hipEventRecord(e1, s);
        const size_t mv_bytes = sizeof(unsigned long long) * lv::kMoveSlots * lv::kMoveSlotStride;
        LV_TRY(stage(64 + mv_bytes));
        LV_HIP(hipMemcpyAsync(h_stage, d_q, sizeof(double), hipMemcpyDeviceToHost, s));
        LV_HIP(hipMemcpyAsync((char *)h_stage + 64, d_moved, mv_bytes, hipMemcpyDeviceToHost, s));
        if (G > 0)
            LV_HIP(hipMemcpyAsync(h_gsums.data(), gsums, sizeof(double) * 3 * (size_t)G,
                                  hipMemcpyDeviceToHost, s));
        LV_TRY(sync());
        std::memcpy(q_after, h_stage, sizeof(double));
        unsigned long long mv = 0, me = 0;
        {
            const unsigned long long *slots =
                reinterpret_cast<const unsigned long long *>((char *)h_stage + 64);
            for (int i = 0; i < lv::kMoveSlots; ++i) {
                mv += slots[i * lv::kMoveSlotStride];
                me += slots[i * lv::kMoveSlotStride + 1];
            }
        }
        *moved = (int64_t)mv;
        last_moved = (int64_t)mv;
        last_moved_edges = (int64_t)me;
        if (sweep_ms) {
            float ms = 0.0f;
            if (e0 && e1) (void)hipEventElapsedTime(&ms, e0, e1);
            *sweep_ms = ms;
        }
        if (e0) ev_pool.push_back(e0);
        if (e1) ev_pool.push_back(e1);
        return LV_OK;
    }

    // Correction rows per lane group: about one mover per block when the
    // share of movers of an earlier sweep (prev_moved, -1 = unknown) holds.
    int32_t rounds_for(int64_t prev_moved) const {
        const int64_t nact = std::max<int64_t>(h_off[lv::kNumBuckets], 1);
        return (int32_t)std::min<int64_t>(
            1024, std::max<int64_t>(1, nact / std::max<int64_t>(prev_moved, 1)));
    }

    // The kernels of one sweep: groups 0..6 (moves, then the apply), then the
    // modularity, and with track the neighbour moves.  fr: a pruned sweep
    // over the frontier select_frontier built.  incremental: the modularity
    // is carried forward from the last one (only the movers' rows are
    // walked) instead of a full pass; both give the same value bit for bit.
    lv_status enqueue_sweep(const lv_frontier *fr, bool track, bool incremental,
                            int64_t prev_moved) {
        const int32_t rounds = rounds_for(prev_moved);
        LV_TRY(enqueue_groups(fr, track, incremental, rounds));
        return enqueue_tail(fr, track, incremental, rounds);
    }

    // Groups 0..6 of a sweep.  Batches with an incremental sweep keep mover
    // bitmaps and add each group's correction per graph; a single graph's
    // weight inside communities is carried forward once per sweep
    // (enqueue_tail).  Under a resident phase's gate these kernels do nothing
    // once the phase has ended.
    lv_status enqueue_groups(const lv_frontier *fr, bool track, bool incremental, int32_t rounds) {
        hipStream_t s = stream;
        const bool pruned = fr != nullptr;
        const bool with_marks = incremental && G > 0;
        // a single graph's movers in one bitmap for the sweep (the first half
        // of bits), when its modularity may be carried forward
        const bool sweep_marks = G == 0 && (incremental || gate_ptr != nullptr);
        uint32_t *smovers = sweep_marks ? bits : nullptr;
        sweep_count++;
        singleton_state = false;
        if (sweep_marks) lv::fill_i32(reinterpret_cast<int32_t *>(bits), 0, bit_words, s);
        lv::launch_hist_shift(hist, V, s, gate_ptr);
        hipLaunchKernelGGL(k_zero_u64_n, dim3(blocks_for(lv::kMoveSlots * lv::kMoveSlotStride, 256)),
                           dim3(256), 0, s, d_moved, (int64_t)lv::kMoveSlots * lv::kMoveSlotStride);
        const lv::MoveArgs a = move_args();
        for (int g = 0; g < lv::kNumGroups; ++g) {
            int32_t lo, hi;
            group_range(g, lo, hi);
            if (lo < 0 || hi <= lo) continue;
            LV_TRY(run_group(g, pruned, track));
            hipEvent_t t0 = tbegin();
            lv::GroupMarks gm = marks();
            gm.rounds = rounds;
            if (G > 0) {
                gm.gid = gid;
                gm.gin = gin;
            }
            const lv::SegTable ct = corr_table(g, rounds);
            // pruned: only frontier vertices can move
            int32_t alo = lo, ahi = hi;
            if (pruned) group_range(g, alo, ahi, h_foff);
            const int32_t *alist = (pruned ? forder : order) + alo;
            lv::launch_apply(a, alist, ahi - alo, order, &ct, ins.p, comm, size, sigma, crec,
                             with_marks ? &gm : nullptr, hist, d_moved, d_own_slots, s, smovers);
            tend(t0, LV_K_APPLY, 16.0 * (double)(ahi - alo));
            LV_LAUNCHED("apply");
            if (with_marks) step++;
        }
        return LV_OK;
    }

    // The modularity after a sweep's groups (and the neighbour moves).  A
    // single graph's incremental sweep adds the change of the weight inside
    // communities over the movers' rows (newc holds, after a sweep, each
    // vertex's community before it) and recomputes the sum of S_c^2.
    lv_status enqueue_tail(const lv_frontier *fr, bool track, bool incremental, int32_t rounds) {
        hipStream_t s = stream;
        const bool pruned = fr != nullptr;
        if (G > 0) {
            // a recount walks whole graphs: the graph frontier lists them, the
            // model's does not (then every row is walked)
            LV_TRY(enqueue_graph_sums(pruned && fr->kind == kFrontierBatchGraphs, incremental));
        } else if (incremental) {
            hipEvent_t tq = tbegin();
            lv::launch_sweep_delta(rowptr(), col(), w(), comm, newc, dk, bits, order,
                                   corr_of(st, rounds), ins.p, rounds, d_own_slots, s);
            lv::launch_modularity_from_own(sigma, V, M, qpart.p, d_q, d_own_slots, s);
            tend(tq, LV_K_MODULARITY, 8.0 * (double)V);
            LV_LAUNCHED("modularity");
        } else {
            LV_TRY(enqueue_modularity());
        }
        if (track) {
            lv::fill_i32(nmv, 0, V, s);
            // Many movers (judged by the last sweep known on the host): every
            // row counts its moved neighbours; few: the movers add one at
            // each neighbour.  The counts are the same.
            const bool pull =
                nmv_pull > 0 ||
                (nmv_pull == 0 && (last_moved_edges < 0 || last_moved_edges * kPullDen >= E));
            if (pull) {
                lv::launch_neighbor_count(rowptr(), col(), hist, V, lv::kNmvMaxDegree, order, st,
                                          ins.p, bits + 2 * (int64_t)bit_words, nmv, s);
            } else {
                // movers' neighbours lie in their own graph: the active
                // graphs' list covers every count that changes
                const bool gl = G > 0 && glist_ok;
                lv::launch_neighbor_moves(rowptr(), col(), gl ? glist : order,
                                          gl ? glist_n : h_off[lv::kNumBuckets], hist, nmv, s);
            }
            LV_LAUNCHED("neighbour moves");
        }
        return LV_OK;
    }

    // Features and move labels of traced sweeps, copied to the caller.
    struct TraceSink {
        float *feat;      // cap x V x 8
        uint8_t *moved;   // cap x V
        int32_t cap;
        int32_t n;
    };

    // Community array of the level (comm or decisions) in the reference's ids.
    lv_status communities_out(const int32_t *dev, int32_t *host) {
        if (maps_identity) {
            if (V > 0)
                LV_HIP(hipMemcpyAsync(host, dev, sizeof(int32_t) * (size_t)V, hipMemcpyDeviceToHost,
                                      stream));
            return sync();
        }
        std::vector<int32_t> rof, c((size_t)std::max(V, 1));
        LV_TRY(host_maps(&rof, nullptr));
        if (V > 0)
            LV_HIP(hipMemcpyAsync(c.data(), dev, sizeof(int32_t) * (size_t)V, hipMemcpyDeviceToHost,
                                  stream));
        LV_TRY(sync());
        for (int32_t r = 0; r < Vref; ++r) host[r] = r;  // vertices without edges: own community
        for (int32_t i = 0; i < V; ++i) host[rof[(size_t)i]] = rof[(size_t)c[(size_t)i]];
        return LV_OK;
    }

    // Community state from comm (V entries, stored ids): sizes, totals and
    // records rebuilt.
    lv_status set_state(const int32_t *c) {
        if (V > 0) {
            LV_HIP(hipMemcpyAsync(comm, c, sizeof(int32_t) * V, hipMemcpyHostToDevice, stream));
            lv::fill_i32(size, 0, V, stream);
            lv::fill_f64(sigma, 0.0, V, stream);
            hipLaunchKernelGGL(k_state_from_comm, dim3(blocks_for(V, 256)), dim3(256), 0, stream, V,
                               comm, k, size, sigma);
            lv::launch_records_all(V, sigma, size, crec, stream);
            state_changed();
            LV_LAUNCHED("state from communities");
        }
        return sync();
    }

    // ---- arrays in the reference's numbering at a compact level
    // Host copies of the level's maps (empty when the numberings agree).
    lv_status host_maps(std::vector<int32_t> *rof, std::vector<int32_t> *cof) {
        if (maps_identity) {
            if (rof) rof->clear();
            if (cof) cof->clear();
            return LV_OK;
        }
        if (rof) {
            rof->resize((size_t)std::max(V, 1));
            if (V > 0)
                LV_HIP(hipMemcpyAsync(rof->data(), ref_of[map_cur], sizeof(int32_t) * (size_t)V,
                                      hipMemcpyDeviceToHost, stream));
        }
        if (cof) {
            cof->resize((size_t)std::max(Vref, 1));
            if (Vref > 0)
                LV_HIP(hipMemcpyAsync(cof->data(), cmp_of[map_cur], sizeof(int32_t) * (size_t)Vref,
                                      hipMemcpyDeviceToHost, stream));
        }
        return sync();
    }

    // dst[Vref rows of w] from src[V rows] through rof; rows of vertices
    // without edges get fill[0..w).
    template <typename T>
    static void expand_rows(const T *src, T *dst, const std::vector<int32_t> &rof, int32_t Vr,
                            int w, const T *fill) {
        for (int64_t r = 0; r < Vr; ++r)
            for (int j = 0; j < w; ++j) dst[r * w + j] = fill[j];
        for (size_t i = 0; i < rof.size(); ++i)
            for (int j = 0; j < w; ++j) dst[(int64_t)rof[i] * w + j] = src[i * w + j];
    }

    // Per-vertex device array (V rows of w) to the caller in the reference's
    // numbering (Vref rows), rows of vertices without edges set to fill.
    template <typename T>
    lv_status copy_out_rows(const T *dev, T *host, int w, const T *fill) {
        if (maps_identity) {
            if (V > 0)
                LV_HIP(hipMemcpyAsync(host, dev, sizeof(T) * (size_t)V * w, hipMemcpyDeviceToHost,
                                      stream));
            return sync();
        }
        std::vector<int32_t> rof;
        LV_TRY(host_maps(&rof, nullptr));
        rof.resize((size_t)V);
        std::vector<T> tmp((size_t)std::max<int64_t>((int64_t)V * w, 1));
        if (V > 0)
            LV_HIP(hipMemcpyAsync(tmp.data(), dev, sizeof(T) * (size_t)V * w, hipMemcpyDeviceToHost,
                                  stream));
        LV_TRY(sync());
        expand_rows(tmp.data(), host, rof, Vref, w, fill);
        return LV_OK;
    }

    // Features of the level's vertices without edges and the model's
    // probability for them.
    lv_status isolated_defaults(float *feat8, float *prob) {
        LV_TRY(tfeat.reserve(16));
        lv::launch_isolated_features(feature_args(), tfeat.p, stream);
        if (prob && mlp_set) lv::launch_mlp_forward(tfeat.p, 1, mlp, tfeat.p + 8, stream);
        LV_LAUNCHED("isolated features");
        float h[16] = {0};
        LV_HIP(hipMemcpyAsync(h, tfeat.p, sizeof(float) * 16, hipMemcpyDeviceToHost, stream));
        LV_TRY(sync());
        std::memcpy(feat8, h, sizeof(float) * lv::kFeat);
        if (prob) *prob = mlp_set ? h[8] : 0.0f;
        return LV_OK;
    }

    lv_status trace_features(TraceSink *t) {
        if (V <= 0 && Vref <= 0) return LV_OK;
        float *dst = t->feat + (int64_t)t->n * Vref * lv::kFeat;
        float iso[lv::kFeat] = {0};
        if (!maps_identity) LV_TRY(isolated_defaults(iso, nullptr));
        if (V <= 0) {
            expand_rows<float>(nullptr, dst, {}, Vref, lv::kFeat, iso);
            return LV_OK;
        }
        LV_TRY(tfeat.reserve((int64_t)V * lv::kFeat));
        lv::launch_features(feature_args(), tfeat.p, stream);
        LV_LAUNCHED("features");
        return copy_out_rows<float>(tfeat.p, dst, lv::kFeat, iso);
    }

    lv_status trace_labels(TraceSink *t) {
        if (Vref <= 0) return LV_OK;
        uint8_t *dst = t->moved + (int64_t)t->n * Vref;
        const uint8_t zero = 0;
        LV_TRY(copy_out_rows<uint8_t>(hist, dst, 1, &zero));
        for (int32_t v = 0; v < Vref; ++v) dst[v] &= 1u;
        return LV_OK;
    }

    // Phase (modularity_optimisation.cu:284-363): sweeps while the gain
    // reaches min_gain and is positive.
    //
    // With a frontier fr (learned pruning): sweep i is full when i is a
    // multiple of fr->full_every, pruned otherwise; a pruned sweep that would
    // end the phase is followed by a full sweep, and only a full sweep ends
    // it.  With a trace sink every sweep is full and its features (before)
    // and moves (after) are recorded.
    lv_status optimise(float min_gain, int32_t *changed, int32_t *sweeps_out, double *ms_out,
                       int64_t *edges_out, const lv_frontier *fr = nullptr,
                       int64_t *evaluated_out = nullptr, int32_t *full_out = nullptr,
                       TraceSink *trace = nullptr) {
        if (!level_ready) LV_TRY(setup_level());
        if (fr == nullptr && trace == nullptr && G == 0) {
            int32_t sw = 0;
            if (!sweeps_out) sweeps_out = &sw;
            LV_TRY(optimise_resident(min_gain, changed, sweeps_out, ms_out, edges_out));
            if (evaluated_out) *evaluated_out = (int64_t)*sweeps_out * h_off[lv::kNumBuckets];
            if (full_out) *full_out = *sweeps_out;
            return LV_OK;
        }
        double q0 = 0.0;
        LV_TRY(modularity(&q0));
        const int k_full = fr ? (fr->full_every > 0 ? fr->full_every : 4) : 1;
        int32_t ch = 0, sweeps = 0, fulls = 0;
        int64_t evaluated = 0, edges = 0;
        double ms_total = 0.0;
        // Every sweep is judged by the reference's rule against the sweep
        // before it.  A pruned sweep that converged is followed by a full
        // sweep (the convergence check), judged against the state before that
        // pruned sweep.  Without a frontier every sweep is full and this is
        // exactly the reference's loop.
        double q_prev = q0, q_before_pruned = q0;
        bool confirm = false;
        for (int64_t i = 0;; ++i) {
            const bool full = fr == nullptr || confirm || i % k_full == 0 || i >= kMaxPrunedSweeps;
            const bool next_may_prune = fr != nullptr && (i + 1) % k_full != 0;
            const bool track = trace != nullptr || next_may_prune;
            if (trace && trace->n < trace->cap) LV_TRY(trace_features(trace));
            double q1 = 0.0, ms = 0.0;
            int64_t mv = 0, ev = 0;
            LV_TRY(sweep(&q1, &mv, &ms, full ? nullptr : fr, track, &ev));
            if (trace && trace->n < trace->cap) {
                LV_TRY(trace_labels(trace));
                trace->n++;
            }
            sweeps++;
            fulls += full ? 1 : 0;
            evaluated += ev;
            edges += full ? E : 0;
            ms_total += ms;
            if (q1 - q_prev > 0.0) ch = 1;
            // A sweep without positive gain also ends the phase (identical
            // for min_gain > 0; guarantees termination for min_gain <= 0).
            const double gain = q1 - (confirm ? q_before_pruned : q_prev);
            const bool stop = mv == 0 || !(gain > 0.0) || !(gain >= (double)min_gain);
            if (stop && full) break;
            // A pruned sweep that lowered Q ends the phase as a full one
            // would; one that converged (no move, or a gain below min_gain)
            // is checked by a full sweep, since the frontier may have missed
            // movers.
            if (stop && q1 - q_prev < 0.0) break;
            confirm = stop;
            if (stop) q_before_pruned = q_prev;
            q_prev = q1;
        }
        compose_phase();
        LV_LAUNCHED("compose");
        if (changed) *changed = ch;
        if (sweeps_out) *sweeps_out = sweeps;
        if (ms_out) *ms_out = ms_total;
        if (edges_out) *edges_out = edges;
        if (evaluated_out) *evaluated_out = evaluated;
        if (full_out) *full_out = fulls;
        return LV_OK;
    }

    // A plain phase resident on the device: the termination test runs on the
    // device after every sweep (k_phase_step), and the host keeps the next
    // sweep's groups enqueued while the current one runs instead of waiting
    // for each sweep's modularity.  Iteration i enqueues sweep i's groups,
    // then waits for sweep i-1's control block: if sweep i-1 ended the phase,
    // sweep i's groups find the gate set and do nothing; otherwise it
    // enqueues sweep i's modularity, chosen (full pass or incremental) from
    // sweep i-1's movers exactly as the synchronous loop chooses it.  Results
    // are those of the synchronous loop: the same sweeps, judged by the same
    // rule on the same modularity values.
    lv_status optimise_resident(float min_gain, int32_t *changed, int32_t *sweeps_out,
                                double *ms_out, int64_t *edges_out) {
        hipStream_t s = stream;
        LV_TRY(enqueue_modularity());  // Q before the first sweep (and the weight inside communities)
        hipLaunchKernelGGL(k_phase_start, dim3(1), dim3(1), 0, s, d_ctl, d_q);
        LV_LAUNCHED("phase start");
        EventPair span;
        LV_HIP(span.create());
        hipEvent_t ev[kCtlRing];
        for (auto &e : ev) {
            e = take_event();
            if (!e) return fail(LV_ERR_HIP, "hipEventCreate failed");
        }
        struct Release {
            lv_graph *g;
            hipEvent_t *ev;
            ~Release() {
                g->gate_ptr = nullptr;
                for (int i = 0; i < kCtlRing; ++i) g->ev_pool.push_back(ev[i]);
            }
        } release{this, ev};
        LV_HIP(hipEventRecord(span.a, s));
        gate_ptr = &d_ctl->done;
        int64_t pm = -1, pme = -1;  // movers of sweep i-1 and their row entries (-1: none yet)
        int64_t i = 0;
        for (;; ++i) {
            LV_TRY(enqueue_groups(nullptr, false, false, 1));
            if (i >= 1) {
                LV_HIP(hipEventSynchronize(ev[(i - 1) % kCtlRing]));
                const PhaseCtl &c1 = h_ctl[(i - 1) % kCtlRing];
                if (c1.done) break;
                pm = (int64_t)c1.moved;
                pme = (int64_t)c1.moved_edges;
            }
            const bool incremental = incr_den > 0 && pme >= 0 && pme * (int64_t)incr_den <= E;
            LV_TRY(enqueue_tail(nullptr, false, incremental, rounds_for(pm)));
            hipLaunchKernelGGL(k_phase_step, dim3(1), dim3(1), 0, s, d_ctl, d_q, d_moved, min_gain);
            LV_LAUNCHED("phase step");
            LV_HIP(hipMemcpyAsync(&h_ctl[i % kCtlRing], d_ctl, sizeof(PhaseCtl),
                                  hipMemcpyDeviceToHost, s));
            LV_HIP(hipEventRecord(ev[i % kCtlRing], s));
        }
        LV_HIP(hipEventRecord(span.b, s));
        LV_TRY(sync());
        const PhaseCtl c = h_ctl[(i - 1) % kCtlRing];  // after the last sweep run
        float ms = 0.0f;
        LV_HIP(hipEventElapsedTime(&ms, span.a, span.b));
        own_valid = true;
        last_moved = (int64_t)c.moved;
        last_moved_edges = (int64_t)c.moved_edges;
        compose_phase();
        LV_LAUNCHED("compose");
        if (changed) *changed = c.changed;
        if (sweeps_out) *sweeps_out = c.sweeps;
        if (ms_out) *ms_out = ms;
        if (edges_out) *edges_out = (int64_t)c.sweeps * E;
        return LV_OK;
    }

    // ---------------------------------------------------------------- phase 2
    // End of a phase: the original vertices follow their level vertex into its
    // community (updateOriginalToCommunity, modularity_optimisation.cu:251-257).
    void compose_phase() {
        if (maps_identity)
            lv::launch_compose(orig2comm, Vorig, comm, stream);
        else
            lv::launch_compose_ref(orig2comm, Vorig, cmp_of[map_cur], ref_of[map_cur], comm,
                                   stream);
    }

    // Renumbering flags of an aggregation (community_aggregation.cu:267-275):
    // a_newid[c] = 1 for the communities that become vertices of the next
    // level, the non-empty ones; when dropping, only those holding a vertex
    // of positive degree (the others are singletons without edges, kept in
    // the reference's numbering only).  *drop tells whether it drops.
    lv_status renumber_flags(bool allow_drop, bool *drop) {
        *drop = allow_drop && compact_on && G == 0 && E > 0;
        if (*drop) {
            lv::fill_i32(a_newid, 0, V, stream);
            lv::launch_has_degree(rowptr(), comm, V, a_newid, stream);
            lv::launch_present_compact(size, V, a_newid, stream);
        } else {
            lv::launch_present(size, V, a_newid, stream);
        }
        if ((*drop || !maps_identity) && V > 0)
            LV_HIP(hipMemcpyAsync(a_keep, a_newid, sizeof(int32_t) * (size_t)V,
                                  hipMemcpyDeviceToDevice, stream));
        return LV_OK;
    }

    // After the scan of the flags (a_newid, newV vertices next): the
    // reference's renumbering of the same communities, which also counts the
    // singletons without edges, moves the original vertices' map
    // (community_aggregation.cu:222-228) and gives the maps of the next
    // level.  *nVref = the next level's vertex count in the reference's ids.
    lv_status finish_renumbering(bool drop, int32_t newV, int32_t *nVref) {
        hipStream_t s = stream;
        if (!drop && maps_identity) {
            lv::launch_compose(orig2comm, Vorig, a_newid, s);
            *nVref = newV;
            return LV_OK;
        }
        if (maps_identity) lv::iota_i32(ref_of[map_cur], V, s);
        int32_t *nref_d = a_scalars + 3;
        lv::fill_i32(newid_ref, 1, Vref, s);
        lv::launch_ref_present(size, ref_of[map_cur], V, newid_ref, s);
        lv::exclusive_scan(newid_ref, newid_ref, Vref, nref_d, scratch, s);
        lv::launch_compose(orig2comm, Vorig, newid_ref, s);
        const int nm = map_cur ^ 1;
        lv::launch_next_ref_of(a_keep, a_newid, ref_of[map_cur], newid_ref, V, ref_of[nm], s);
        lv::fill_i32(cmp_of[nm], -1, Vref, s);
        lv::launch_cmp_of(ref_of[nm], newV, cmp_of[nm], s);
        LV_LAUNCHED("reference renumbering");
        LV_TRY(stage(64));
        LV_HIP(hipMemcpyAsync(h_stage, nref_d, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        LV_TRY(sync());
        std::memcpy(nVref, h_stage, sizeof(int32_t));
        map_cur = nm;
        maps_identity = *nVref == newV;  // nothing dropped: the maps are the identity
        return LV_OK;
    }

    lv_status aggregate() {
        if (!level_ready) LV_TRY(setup_level());
        if (opt.agg_mode == 0) return aggregate_sorted();
        hipStream_t s = stream;
        hipEvent_t t0 = tbegin();
        const int32_t *rp = rowptr();
        int32_t *newV_d = a_scalars;
        int32_t *etot_d = a_scalars + 1;
        // renumbering (community_aggregation.cu:267-275)
        bool drop = false;
        LV_TRY(renumber_flags(false, &drop));
        lv::exclusive_scan(a_newid, a_newid, V, newV_d, scratch, s);
        if (G > 0) lv::launch_graph_offsets(vptr, G, V, a_newid, newV_d, vptr_next, s);
        lv::exclusive_scan(size, a_mstart, V, nullptr, scratch, s);
        lv::fill_i32(a_fill, 0, V, s);
        lv::launch_members(comm, V, a_mstart, a_fill, a_members, s);
        lv::launch_member_degrees(rp, a_members, V, a_mscan, s);
        lv::exclusive_scan(a_mscan, a_mscan, V, a_mscan + V, scratch, s);
        lv::launch_agg_keys(size, a_mstart, a_mscan, V, agg_bounds, a_keys, s);
        lv::stable_partition(a_keys, V, lv::kAggTiers, a_order, a_off, scratch, s);
        lv::fill_i32(a_cnt, 0, V, s);
        LV_LAUNCHED("aggregation setup");
        LV_TRY(stage(4096));
        LV_HIP(hipMemcpyAsync(h_stage, a_off, sizeof(int32_t) * (lv::kAggTiers + 1),
                              hipMemcpyDeviceToHost, s));
        LV_HIP(hipMemcpyAsync((char *)h_stage + 64, newV_d, sizeof(int32_t),
                              hipMemcpyDeviceToHost, s));
        LV_TRY(sync());
        int32_t toff[lv::kAggTiers + 1];
        std::memcpy(toff, h_stage, sizeof toff);
        int32_t newV = 0;
        std::memcpy(&newV, (char *)h_stage + 64, sizeof newV);

        lv::AggArgs a;
        a.rowptr = rp;
        a.col = col();
        a.w = w();
        a.comm = comm;
        a.size = size;
        a.newid = a_newid;
        a.mstart = a_mstart;
        a.members = a_members;
        a.mscan = a_mscan;
        a.cnt = a_cnt;
        a.tmp_col = tmp_col;
        a.tmp_w = tmp_w;
        for (int t = 0; t < lv::kAggTiers - 1; ++t)
            lv::launch_merge_lds(t, a, a_order + toff[t], toff[t + 1] - toff[t], s);
        LV_LAUNCHED("merge (LDS tiers)");

        const int GT = lv::kAggTiers - 1;
        const int32_t ng = toff[GT + 1] - toff[GT];
        if (ng > 0) {
            const int32_t *glist = a_order + toff[GT];
            LV_TRY(a_plan.reserve(ng));
            LV_TRY(a_tot.reserve(1));
            LV_TRY(a_toff.reserve(ng));
            LV_TRY(a_soff.reserve(ng));
            LV_TRY(a_lg.reserve(ng));
            hipLaunchKernelGGL(k_agg_plan, dim3(1), dim3(kPlanThreads), 0, s, glist, ng, size,
                               a_mstart, a_mscan, newV, kAggChunk, kSlotChunk, a_toff.p, a_soff.p,
                               a_lg.p, a_plan.p, a_tot.p);
            LV_LAUNCHED("aggregation plan");
            LV_TRY(stage(sizeof(AggTotals)));
            LV_HIP(hipMemcpyAsync(h_stage, a_tot.p, sizeof(AggTotals), hipMemcpyDeviceToHost, s));
            LV_TRY(sync());
            AggTotals at;
            std::memcpy(&at, h_stage, sizeof at);
            if (at.nins > INT32_MAX || at.next > INT32_MAX)
                return fail(LV_ERR_ARG, "aggregation tables exceed 2^31 chunks");
            LV_TRY(arena(at.slots));
            LV_TRY(a_ins.reserve(at.nins));
            LV_TRY(a_ext.reserve(at.next));
            LV_TRY(sortA.reserve(at.sorts));
            LV_TRY(sortB.reserve(at.sorts));
            hipLaunchKernelGGL(k_agg_fill, dim3(ng), dim3(256), 0, s, glist, a_plan.p, a_toff.p,
                               a_lg.p, kAggChunk, kSlotChunk, a_ins.p, a_ext.p);
            LV_LAUNCHED("aggregation tables");
            lv::launch_agg_global(a, a_ins.p, (int32_t)at.nins, a_ext.p, (int32_t)at.next,
                                  glist, ng, a_toff.p, a_lg.p, gkeys.p, gvals.p, sortA.p, sortB.p,
                                  a_soff.p, s);
            LV_LAUNCHED("merge (global tier)");
        }

        // CSR offsets and compaction (compressEdges, community_aggregation.cu:196-220)
        const int nb = cur_buf == 0 ? 1 : 0;
        lv::exclusive_scan(a_cnt, a_cstart, V, etot_d, scratch, s);
        lv::launch_new_rowptr(size, a_newid, a_cstart, V, newV_d, etot_d, rpb[nb], s);
        for (int t = 0; t < lv::kAggTiers; ++t)
            lv::launch_copy_rows(a_order + toff[t], toff[t + 1] - toff[t], t >= 2, a_mstart,
                                 a_mscan, a_cnt, a_cstart, tmp_col, tmp_w, colb[nb], wb[nb], s);
        LV_LAUNCHED("aggregation output");
        // updateOriginalToCommunity (community_aggregation.cu:222-228)
        int32_t nVref = 0;
        LV_TRY(finish_renumbering(drop, newV, &nVref));
        tend(t0, LV_K_AGGREGATE, 24.0 * (double)E + 40.0 * (double)V);
        LV_HIP(hipMemcpyAsync(h_stage, etot_d, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        LV_TRY(sync());
        int32_t newE = 0;
        std::memcpy(&newE, h_stage, sizeof newE);
        cur_buf = nb;
        V = newV;
        Vref = nVref;
        E = newE;
        level_ready = false;
        col_shadow = false;
        rows_sorted = false;
        if (G > 0) {
            std::swap(vptr, vptr_next);
            lv::launch_graph_ids(vptr, G, gid, s);
            LV_LAUNCHED("graph ids");
        }
        return LV_OK;
    }

    // ---------------------------------------------------------------- batch
    // Per-graph (own, sq, moved) of the current state into gsums (device).
    // pruned: only the graphs of the current frontier (their rows are the
    // frontier lists; the others' sums are not read).
    // Per-graph sums after a union sweep.  incremental: the weight inside
    // communities was carried forward by the applies' corrections (gin);
    // otherwise it is recomputed from the rows (of the active graphs when
    // pruned) and stored in gin.
    lv_status enqueue_graph_sums(bool pruned = false, bool incremental = false) {
        hipStream_t s = stream;
        hipEvent_t t0 = tbegin();
        const uint8_t *ga = pruned ? gact : nullptr;
        if (incremental) {
            lv::launch_graph_sums(vptr, G, nullptr, sigma, hist, ga, gin, gsums, s);
            tend(t0, LV_K_MODULARITY, 17.0 * (double)V);
        } else {
            lv::fill_f64(own_v, 0.0, V, s);
            if (pruned) {
                const lv::SegTable fst = make_segs(h_foff, nins);
                lv::launch_vertex_own(rowptr(), col(), w(), comm, forder, fst, ins.p, active, own_v,
                                      s);
            } else {
                lv::launch_vertex_own(rowptr(), col(), w(), comm, order, st, ins.p, nullptr, own_v,
                                      s);
            }
            lv::launch_graph_sums(vptr, G, own_v, sigma, hist, ga, gin, gsums, s);
            tend(t0, LV_K_MODULARITY, 12.0 * (double)E + 16.0 * (double)V);
            own_valid = true;
        }
        LV_LAUNCHED("graph sums");
        return LV_OK;
    }

    // glist / gl_off from the uploaded gact: frontier flags and keys of the
    // active graphs (launch_graph_frontier), partitioned by bucket.
    lv_status build_graph_list() {
        hipStream_t s = stream;
        lv::launch_graph_frontier(rowptr(), V, gid, gact, bb, fkeys, active, s);
        lv::stable_partition(fkeys, V, lv::kNumBuckets, glist, gl_off, scratch, s);
        LV_LAUNCHED("graph list");
        int32_t off[lv::kNumBuckets + 1];
        LV_TRY(stage(4096));
        LV_HIP(hipMemcpyAsync(h_stage, gl_off, sizeof(off), hipMemcpyDeviceToHost, s));
        LV_TRY(sync());
        std::memcpy(off, h_stage, sizeof(off));
        glist_n = off[lv::kNumBuckets];
        glist_ok = true;
        return LV_OK;
    }

    double graph_q(int32_t g) const {
        const double Mg = gM[g];
        return h_gsums[3 * g] / (2.0 * Mg) - h_gsums[3 * g + 1] / (4.0 * Mg * Mg);
    }

    lv_status setup_batch(int32_t nG, const int32_t *vp, const int32_t *rowptr_h, const float *w_h) {
        G = nG;
        gmaxV = 0;
        for (int32_t g = 0; g < G; ++g) gmaxV = std::max(gmaxV, vp[g + 1] - vp[g]);
        h_vptr0.assign(vp, vp + G + 1);
        gM.assign(G, 0.0);
        std::vector<float> m32(G), dn(G);
        for (int32_t g = 0; g < G; ++g) {
            double s = 0.0;  // (sum of the graph's CSR weights)/2 in CSR order, as lv_create
            for (int64_t e = rowptr_h[vp[g]]; e < rowptr_h[vp[g + 1]]; ++e) s += (double)w_h[e];
            gM[g] = s / 2.0;
            m32[g] = (float)gM[g];
            volatile float two_m = 2.0f * m32[g];
            volatile float d = two_m * m32[g];
            dn[g] = d;
        }
        LV_TRY(dev_alloc(&vptr, G + 1));
        LV_TRY(dev_alloc(&vptr_next, G + 1));
        LV_TRY(dev_alloc(&vptr0, G + 1));
        LV_TRY(dev_alloc(&gid, (int64_t)V0 + 1));
        LV_TRY(dev_alloc(&gid0, (int64_t)V0 + 1));
        LV_TRY(dev_alloc(&gfinal, (int64_t)V0 + 1));
        LV_TRY(dev_alloc(&gM32, G));
        LV_TRY(dev_alloc(&gden, G));
        LV_TRY(dev_alloc(&gact, G + 8));
        LV_TRY(dev_alloc(&gdone, G + 8));
        LV_TRY(dev_alloc(&gsums, 3 * (int64_t)G));
        LV_TRY(dev_alloc(&own_v, (int64_t)V0 + 1));
        LV_TRY(dev_alloc(&gin, (int64_t)G + 1));
        LV_TRY(dev_alloc(&gfull, (int64_t)G + 8));
        LV_TRY(dev_alloc(&glist, (int64_t)V0 + 1));
        LV_TRY(dev_alloc(&gl_off, 64));
        LV_TRY(dev_alloc(&gvscale, (int64_t)G + 1));
        h_gsums.assign(3 * (size_t)G, 0.0);
        LV_HIP(hipHostMalloc(reinterpret_cast<void **>(&h_up), kUpSlots * up_stride(),
                            hipHostMallocDefault));
        LV_HIP(hipMemcpy(vptr0, vp, sizeof(int32_t) * (G + 1), hipMemcpyHostToDevice));
        LV_HIP(hipMemcpy(gM32, m32.data(), sizeof(float) * G, hipMemcpyHostToDevice));
        LV_HIP(hipMemcpy(gden, dn.data(), sizeof(float) * G, hipMemcpyHostToDevice));
        lv::launch_graph_ids(vptr0, G, gid0, stream);
        LV_LAUNCHED("graph ids");
        return sync();
    }

    // Per graph: plain phases and aggregations exactly as lv_run would run
    // them on that graph alone; all graphs advance together, a graph whose
    // phase ended waits (outside the frontier) for the others.
    // Batched runs: every graph follows lv_run's phase rule (or, with a model
    // frontier mfr, lv_run_pruned's: its sweep i of a phase is full when
    // i % full_every == 0 or to confirm a converged pruned sweep) on its own
    // sums, and leaves the union's frontier when its phase ends.
    lv_status run_batch(float min_gain, lv_batch_result *res, lv_stats *tot,
                        const lv_frontier *mfr = nullptr) {
        hipStream_t s = stream;
        EventPair ev;
        LV_HIP(ev.create());
        hipEvent_t e0 = ev.a, e1 = ev.b;
        LV_HIP(hipEventRecord(e0, s));
        LV_TRY(reset());
        LV_HIP(hipMemcpyAsync(vptr, vptr0, sizeof(int32_t) * (G + 1), hipMemcpyDeviceToDevice, s));
        LV_HIP(hipMemcpyAsync(gid, gid0, sizeof(int32_t) * V0, hipMemcpyDeviceToDevice, s));
        std::vector<uint8_t> done(G, 0), act(G, 0), changed(G, 0);
        std::vector<double> qprev(G, 0.0);
        // pruned runs: per graph the sweep index in its phase, the pending
        // confirmation and Q before the pruned sweep it confirms
        std::vector<int64_t> gi(G, 0);
        std::vector<uint8_t> gconf(G, 0), full(G, 1), full_up(G, 2);
        std::vector<double> qbp(G, 0.0);
        const int k_full = mfr ? (mfr->full_every > 0 ? mfr->full_every : 4) : 1;
        lv_frontier mf{};
        if (mfr) {
            mf = *mfr;
            mf.kind = kFrontierBatchModel;
        }
        std::vector<lv_batch_result> r(G);
        for (auto &x : r) x = lv_batch_result{};
        lv_stats st{};
        double agg_ms = 0.0;
        int32_t remaining = G;
        lv_frontier fr{};
        fr.kind = kFrontierBatchGraphs;
        while (remaining > 0) {
            if (!level_ready) LV_TRY(setup_level());
            // Q at the phase start
            LV_TRY(enqueue_graph_sums());
            LV_HIP(hipMemcpyAsync(h_gsums.data(), gsums, sizeof(double) * 3 * (size_t)G,
                                  hipMemcpyDeviceToHost, s));
            LV_TRY(sync());
            int32_t nact = 0;
            for (int32_t g = 0; g < G; ++g) {
                act[g] = done[g] ? 0 : 1;
                changed[g] = 0;
                nact += act[g];
                if (act[g]) qprev[g] = graph_q(g);
                gi[g] = 0;
                gconf[g] = 0;
            }
            LV_TRY(upload_flags(gact, act, kUpAct));
            if (mfr) LV_TRY(build_graph_list());
            st.levels++;
            while (nact > 0) {
                double q1 = 0.0, ms = 0.0;
                int64_t mv = 0, ev = 0;
                const bool all = nact == G;
                const lv_frontier *sfr = all ? nullptr : &fr;
                if (mfr) {
                    bool any_pruned = false;
                    for (int32_t g = 0; g < G; ++g) {
                        full[g] = (gconf[g] || gi[g] % k_full == 0 || gi[g] >= kMaxPrunedSweeps)
                                      ? 1 : 0;
                        if (act[g] && !full[g]) any_pruned = true;
                    }
                    if (any_pruned) {
                        if (full != full_up) {
                            LV_TRY(upload_flags(gfull, full, kUpFull));
                            full_up = full;
                        }
                        sfr = &mf;
                    }
                }
                // the frontier features need the neighbour moves of a sweep
                // when some graph's next sweep may be pruned (lv_optimise's
                // rule per graph)
                bool track = false;
                if (mfr)
                    for (int32_t g = 0; g < G && !track; ++g)
                        track = act[g] && (gi[g] + 1) % k_full != 0;
                LV_TRY(sweep(&q1, &mv, &ms, sfr, track, &ev));
                st.sweeps++;
                st.sweep_ms += ms;
                st.evaluated += ev;
                bool flip = false;
                for (int32_t g = 0; g < G; ++g) {
                    if (!act[g]) continue;
                    const double q = graph_q(g);
                    const double moved_g = h_gsums[3 * g + 2];
                    r[g].sweeps++;
                    if (q - qprev[g] > 0.0) changed[g] = 1;
                    bool end = false;
                    if (!mfr) {
                        // the phase rule of lv_optimise, per graph
                        const double gain = q - qprev[g];
                        end = moved_g == 0.0 || !(gain > 0.0) || !(gain >= (double)min_gain);
                        qprev[g] = q;
                    } else {
                        // lv_optimise with a frontier, per graph
                        const double gain = q - (gconf[g] ? qbp[g] : qprev[g]);
                        const bool stop =
                            moved_g == 0.0 || !(gain > 0.0) || !(gain >= (double)min_gain);
                        st.full_sweeps += full[g] ? 1 : 0;
                        end = (stop && full[g]) || (stop && q - qprev[g] < 0.0);
                        if (!end) {
                            gconf[g] = stop ? 1 : 0;
                            if (stop) qbp[g] = qprev[g];
                            gi[g]++;
                        }
                        qprev[g] = q;
                    }
                    if (end) {
                        act[g] = 0;
                        nact--;
                        flip = true;
                    }
                }
                if (flip && nact > 0) {
                    LV_TRY(upload_flags(gact, act, kUpAct));
                    if (mfr) LV_TRY(build_graph_list());
                }
            }
            lv::launch_compose(orig2comm, Vorig, comm, s);
            LV_LAUNCHED("compose");
            std::vector<uint8_t> now(G, 0);
            bool any_now = false;
            for (int32_t g = 0; g < G; ++g) {
                if (done[g]) continue;
                r[g].levels++;
                if (!changed[g]) {
                    done[g] = 1;
                    now[g] = 1;
                    any_now = true;
                    remaining--;
                    r[g].modularity = qprev[g];
                }
            }
            if (any_now) {
                std::vector<int32_t> vp(G + 1);
                LV_HIP(hipMemcpyAsync(vp.data(), vptr, sizeof(int32_t) * (G + 1), hipMemcpyDeviceToHost, s));
                LV_TRY(upload_flags(gdone, now, kUpDone));
                lv::launch_graph_final(orig2comm, V0, gid0, gdone, vptr, gfinal, s);
                LV_LAUNCHED("final assignment");
                LV_TRY(sync());
                for (int32_t g = 0; g < G; ++g)
                    if (now[g]) r[g].final_V = vp[g + 1] - vp[g];
            }
            if (remaining == 0) break;
            hipEvent_t a0 = take_event(), a1 = take_event();
            if (a0) (void)hipEventRecord(a0, s);
            LV_TRY(aggregate());
            if (a1) (void)hipEventRecord(a1, s);
            LV_TRY(sync());
            float ams = 0.0f;
            if (a0 && a1) (void)hipEventElapsedTime(&ams, a0, a1);
            agg_ms += ams;
            if (a0) ev_pool.push_back(a0);
            if (a1) ev_pool.push_back(a1);
        }
        LV_HIP(hipEventRecord(e1, s));
        LV_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        LV_HIP(hipEventElapsedTime(&ms, e0, e1));
        double qsum = 0.0;
        for (int32_t g = 0; g < G; ++g) qsum += r[g].modularity;
        st.modularity = G > 0 ? qsum / G : 0.0;
        st.algo_ms = ms;
        st.aggregate_ms = agg_ms;
        st.memory_ms = memory_ms;
        st.final_V = V;
        st.final_E = E;
        if (res) std::memcpy(res, r.data(), sizeof(lv_batch_result) * (size_t)G);
        if (tot) *tot = st;
        return LV_OK;
    }

    // Aggregation by radix sort (community_aggregation.cu:248-347 computes
    // the same CSR by per-community hash merges): keys (new id of the row's
    // community, new id of the column's community) of every CSR entry are
    // sorted, equal keys summed in FP64, and the unique keys are the next
    // level's rows and columns in ascending new id.
    lv_status aggregate_sorted() {
        hipStream_t s = stream;
        hipEvent_t t0 = tbegin();
        int32_t *newV_d = a_scalars;
        int32_t *etot_d = a_scalars + 1;
        int32_t *U_d = a_scalars + 2;
        // renumbering (community_aggregation.cu:267-275)
        bool drop = false;
        LV_TRY(renumber_flags(true, &drop));
        lv::exclusive_scan(a_newid, a_newid, V, newV_d, scratch, s);
        if (G > 0) lv::launch_graph_offsets(vptr, G, V, a_newid, newV_d, vptr_next, s);
        LV_LAUNCHED("renumbering");
        LV_TRY(stage(64));
        LV_HIP(hipMemcpyAsync(h_stage, newV_d, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        LV_TRY(sync());
        int32_t newV = 0;
        std::memcpy(&newV, h_stage, sizeof newV);
        const int nb = cur_buf == 0 ? 1 : 0;
        if (E > 0) {
            LV_TRY(skA.reserve(E));
            LV_TRY(skB.reserve(E));
            LV_TRY(svA.reserve(E));
            LV_TRY(svB.reserve(E));
            LV_TRY(shead.reserve(E));
            LV_TRY(spos.reserve(E));
            LV_TRY(sacc.reserve(E));
            LV_TRY(sscratch.reserve(std::max(lv::radix_scratch_elems(E), lv::scan_scratch_elems(E) + 8)));
            // the level's bucket-ordered columns, when it has them
            if (col_shadow) LV_TRY(c_cn.reserve(V));
            lv::launch_agg_pairs(rowptr(), col(), w(), comm, a_newid, V, order, st, ins.p, a_mstart,
                                 skA.p, svA.p, s, col_shadow ? c_col.p : nullptr, c_pos.p,
                                 c_cn.p);
            // (row id, column id) as one key of 2*bits bits, column id low:
            // 8-bit digits over it (stable LSD)
            int bits = 1;
            while ((1ll << bits) < (int64_t)newV) ++bits;
            int shifts[8], np = 0;
            for (int b = 0; b < 2 * bits; b += 8) shifts[np++] = b;
            const bool in_b = lv::radix_sort_pairs(skA.p, svA.p, skB.p, svB.p, E, shifts, np,
                                                   bits, sscratch.p, s);
            uint64_t *keys = in_b ? skB.p : skA.p;
            float *vals = in_b ? svB.p : svA.p;
            uint64_t *ukey = in_b ? skA.p : skB.p;  // the free key buffer
            lv::launch_run_heads(keys, E, shead.p, s);
            lv::exclusive_scan(shead.p, spos.p, E, U_d, sscratch.p, s);
            lv::fill_f64(sacc.p, 0.0, E, s);
            lv::launch_run_reduce(keys, vals, shead.p, spos.p, E, ukey, sacc.p, s);
            lv::fill_i32(a_cnt, 0, (int64_t)newV + 1, s);
            lv::launch_unique_rows(ukey, sacc.p, U_d, a_cnt, colb[nb], wb[nb], E, s);
        } else {
            lv::fill_i32(a_cnt, 0, (int64_t)newV + 1, s);
            lv::fill_i32(U_d, 0, 1, s);
        }
        lv::exclusive_scan(a_cnt, rpb[nb], (int64_t)newV + 1, etot_d, scratch, s);
        LV_LAUNCHED("sort aggregation");
        // updateOriginalToCommunity (community_aggregation.cu:222-228)
        int32_t nVref = 0;
        LV_TRY(finish_renumbering(drop, newV, &nVref));
        tend(t0, LV_K_AGGREGATE, 24.0 * (double)E + 40.0 * (double)V);
        LV_HIP(hipMemcpyAsync(h_stage, U_d, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        LV_TRY(sync());
        int32_t newE = 0;
        std::memcpy(&newE, h_stage, sizeof newE);
        cur_buf = nb;
        V = newV;
        Vref = nVref;
        E = newE;
        level_ready = false;
        col_shadow = false;
        rows_sorted = true;
        if (G > 0) {
            std::swap(vptr, vptr_next);
            lv::launch_graph_ids(vptr, G, gid, s);
            LV_LAUNCHED("graph ids");
        }
        return LV_OK;
    }

    // The original vertices of an earlier run map into this handle's level 0
    // by o2c (n entries below its vertex count); reset() and lv_assignment
    // use it.  The
    // run's total weight M_run replaces the one summed from this level.
    lv_status set_initial_map(const int32_t *o2c, int32_t n, double M_run) {
        for (int32_t x = 0; x < n; ++x)
            if (o2c[x] < 0 || o2c[x] >= Vref0) return fail(LV_ERR_ARG, "map entry out of range");
        if (!(M_run >= 0.0)) return fail(LV_ERR_ARG, "bad total weight");
        // the run's M (aggregated fp32 weights need not sum to it exactly)
        M = M_run;
        M32 = (float)M;
        {
            volatile float two_m = 2.0f * M32;
            volatile float d = two_m * M32;
            den = d;
        }
        if ((int64_t)n + 1 > (int64_t)Vref0 + 1) {
            dev_free(orig2comm);
            orig2comm = nullptr;
            LV_TRY(dev_alloc(&orig2comm, (int64_t)n + 1));
        }
        dev_free(o2c0);
        o2c0 = nullptr;
        LV_TRY(dev_alloc(&o2c0, (int64_t)n + 1));
        LV_HIP(hipMemcpy(o2c0, o2c, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
        Vorig = n;
        return reset();
    }

    lv_status destroy() {
        if (stream) (void)hipStreamSynchronize(stream);
        dev_free(rp0);
        dev_free(col0);
        dev_free(w0);
        for (int i = 0; i < 2; ++i) {
            dev_free(rpb[i]);
            dev_free(colb[i]);
            dev_free(wb[i]);
        }
        dev_free(k);
        dev_free(sigma);
        dev_free(comm);
        dev_free(size);
        dev_free(newc);
        dev_free(orig2comm);
        dev_free(o2c0);
        dev_free(ref_of0);
        dev_free(cmp_of0);
        dev_free(crec);
        dev_free(dk);
        dev_free(margin);
        dev_free(kown);
        dev_free(hist);
        dev_free(active);
        dev_free(nmv);
        dev_free(fkeys);
        dev_free(forder);
        dev_free(fd_off);
        dev_free(mlp);
        tfeat.release();
        xsplit.release();
        dev_free(vptr);
        dev_free(vptr_next);
        dev_free(vptr0);
        dev_free(gid);
        dev_free(gid0);
        dev_free(gfinal);
        dev_free(gM32);
        dev_free(gden);
        dev_free(gact);
        dev_free(gdone);
        dev_free(gsums);
        dev_free(own_v);
        dev_free(gin);
        dev_free(gfull);
        dev_free(glist);
        dev_free(gl_off);
        dev_free(gvscale);
        if (h_up) (void)hipHostFree(h_up);
        h_up = nullptr;
        dev_free(bits);
        dev_free(keys8);
        dev_free(order);
        dev_free(d_off);
        dev_free(d_bedges);
        dev_free(scratch);
        dev_free(d_q);
        dev_free(d_moved);
        dev_free(d_own_slots);
        dev_free(d_ctl);
        if (h_ctl) (void)hipHostFree(h_ctl);
        h_ctl = nullptr;
        dev_free(a_newid);
        dev_free(a_keep);
        dev_free(newid_ref);
        for (int i = 0; i < 2; ++i) {
            dev_free(ref_of[i]);
            dev_free(cmp_of[i]);
        }
        dev_free(a_mstart);
        dev_free(a_fill);
        dev_free(a_members);
        dev_free(a_mscan);
        dev_free(a_cnt);
        dev_free(a_cstart);
        dev_free(a_order);
        dev_free(a_off);
        dev_free(a_scalars);
        dev_free(a_keys);
        dev_free(tmp_col);
        dev_free(tmp_w);
        ins.release();
        hvs.release();
        items.release();
        out_key.release();
        out_val.release();
        poff.release();
        ptot.release();
        ghint.release();
        hdone_w.release();
        hplan.release();
        tier_tot.release();
        hbest.release();
        hkcur.release();
        hovf.release();
        hdone.release();
        hhint.release();
        hdist.release();
        corder.release();
        iorder.release();
        gkeys.release();
        gvals.release();
        c_pos.release();
        c_col.release();
        c_comm.release();
        c_cn.release();
        item_best.release();
        item_K.release();
        qpart.release();
        a_ins.release();
        a_ext.release();
        a_toff.release();
        a_soff.release();
        a_lg.release();
        a_plan.release();
        a_tot.release();
        sortA.release();
        sortB.release();
        skA.release();
        skB.release();
        svA.release();
        svB.release();
        shead.release();
        spos.release();
        sscratch.release();
        sacc.release();
        if (h_stage) (void)hipHostFree(h_stage);
        h_stage = nullptr;
        for (auto &r : pending) {
            ev_pool.push_back(r.a);
            ev_pool.push_back(r.b);
        }
        pending.clear();
        for (auto e : ev_pool) (void)hipEventDestroy(e);
        ev_pool.clear();
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
        return LV_OK;
    }
};

// ====================================================================== C-ABI

extern "C" {

const char *lv_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"

namespace {

// Shape checks of a CSR handed to lv_create / lv_batch_create.
lv_status check_csr(const int32_t *rowptr, const int32_t *col, const float *w, int32_t V,
                    int64_t E) {
    if (V < 0 || E < 0 || E > 0x7fffffffLL) return fail(LV_ERR_ARG, "V/E out of range");
    if (!rowptr || (E > 0 && (!col || !w))) return fail(LV_ERR_ARG, "null CSR array");
    if (rowptr[0] != 0 || (int64_t)rowptr[V] != E) return fail(LV_ERR_ARG, "rowptr[0]/rowptr[V] mismatch");
    for (int32_t v = 0; v < V; ++v)
        if (rowptr[v + 1] < rowptr[v]) return fail(LV_ERR_ARG, "rowptr not monotone");
    for (int64_t e = 0; e < E; ++e)
        if (col[e] < 0 || col[e] >= V) return fail(LV_ERR_ARG, "column index out of range");
    if (E > 0 && !lv::csr_symmetric(rowptr, col, w, V))
        return fail(LV_ERR_ARG, "CSR is not symmetric: an entry (v, u, w) has no matching (u, v, w)");
    return LV_OK;
}

// Weights of entries [lo, hi) rounded to their own grid into wq (allocated
// on the first change; untouched entries are copied then).
lv_status grid_round(const float *w, int64_t E, int64_t lo, int64_t hi, std::vector<float> &wq) {
    int32_t shift = 0;
    int64_t changed = 0;
    LV_TRY(lv_weight_grid(w + lo, hi - lo, &shift, &changed));
    if (changed == 0) return LV_OK;
    if (wq.empty()) wq.assign(w, w + E);
    LV_TRY(lv_quantize_weights(w + lo, hi - lo, shift, wq.data() + lo));
    return LV_OK;
}

// A handle over a CSR already checked and rounded.
lv_status create_handle(const int32_t *rowptr, const int32_t *col, const float *w, int32_t V,
                        int64_t E, const lv_options *opt, lv_graph **out) {
    lv_graph *g = new (std::nothrow) lv_graph();
    if (!g) return fail(LV_ERR_OOM, "host allocation failed");
    lv_status st = g->create(rowptr, col, w, V, E, opt);
    if (st != LV_OK) {
        std::string msg = g_last_error;
        g->destroy();
        delete g;
        g_last_error = msg;
        return st;
    }
    *out = g;
    return LV_OK;
}

}  // namespace

extern "C" {

lv_status lv_create(const int32_t *rowptr, const int32_t *col, const float *w, int32_t V,
                    int64_t E, const lv_options *opt, lv_graph **out) {
    if (!out) return fail(LV_ERR_ARG, "out is null");
    *out = nullptr;
    LV_TRY(check_csr(rowptr, col, w, V, E));
    std::vector<float> wq;
    if (!opt || opt->exact_sums >= 0) LV_TRY(grid_round(w, E, 0, E, wq));
    return create_handle(rowptr, col, wq.empty() ? w : wq.data(), V, E, opt, out);
}

void lv_destroy(lv_graph *g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    g->destroy();
    delete g;
}

lv_status lv_reset(lv_graph *g) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    LV_HIP(hipSetDevice(g->device));
    return g->reset();
}

lv_status lv_optimise(lv_graph *g, float min_gain, int32_t *changed) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    LV_HIP(hipSetDevice(g->device));
    return g->optimise(min_gain, changed, nullptr, nullptr, nullptr);
}

lv_status lv_aggregate(lv_graph *g) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    LV_HIP(hipSetDevice(g->device));
    return g->aggregate();
}

lv_status lv_modularity(lv_graph *g, double *q) {
    if (!g || !q) return fail(LV_ERR_ARG, "null argument");
    LV_HIP(hipSetDevice(g->device));
    return g->modularity(q);
}

lv_status lv_sweep(lv_graph *g, double *q_after, int64_t *moved) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    LV_HIP(hipSetDevice(g->device));
    double q1 = 0.0;
    int64_t mv = 0;
    LV_TRY(g->sweep(&q1, &mv, nullptr));
    if (q_after) *q_after = q1;
    if (moved) *moved = mv;
    return LV_OK;
}

}  // extern "C"

namespace {

// Level loop (gpulouvain.cu:30-44): phase, then aggregation while the phase
// changed the communities; fr selects pruned phases.
lv_status run_levels(lv_graph *g, float min_gain, const lv_frontier *fr, lv_stats *stats) {
    LV_HIP(hipSetDevice(g->device));
    EventPair ev, aev;
    LV_HIP(ev.create());
    LV_HIP(aev.create());
    hipEvent_t e0 = ev.a, e1 = ev.b;
    LV_HIP(hipEventRecord(e0, g->stream));
    lv_stats st{};
    double agg_ms = 0.0;
    for (;;) {
        int32_t changed = 0, sw = 0, full = 0;
        double ms = 0.0;
        int64_t edges = 0, ev = 0;
        LV_TRY(g->optimise(min_gain, &changed, &sw, &ms, &edges, fr, &ev, &full));
        if (st.levels == 0) {
            st.level0_sweeps = sw;
            st.level0_sweep_ms = ms;
        }
        st.levels++;
        st.sweeps += sw;
        st.full_sweeps += full;
        st.evaluated += ev;
        st.sweep_ms += ms;
        st.edges_swept += edges;
        if (!changed) break;
        hipEvent_t a0 = aev.a, a1 = aev.b;
        LV_HIP(hipEventRecord(a0, g->stream));
        LV_TRY(g->aggregate());
        LV_HIP(hipEventRecord(a1, g->stream));
        LV_HIP(hipEventSynchronize(a1));
        float ams = 0.0f;
        LV_HIP(hipEventElapsedTime(&ams, a0, a1));
        agg_ms += ams;
    }
    double q = 0.0;
    LV_TRY(g->modularity(&q));
    LV_HIP(hipEventRecord(e1, g->stream));
    LV_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    LV_HIP(hipEventElapsedTime(&ms, e0, e1));
    st.modularity = q;
    st.algo_ms = ms;
    st.aggregate_ms = agg_ms;
    st.memory_ms = g->memory_ms;
    st.final_V = g->Vref;
    st.final_E = g->E;
    if (stats) *stats = st;
    return LV_OK;
}

lv_status check_frontier(const lv_frontier *fr) {
    if (!fr) return fail(LV_ERR_ARG, "null frontier");
    if (fr->kind < LV_FRONTIER_MODEL || fr->kind > LV_FRONTIER_ALL)
        return fail(LV_ERR_ARG, "unknown frontier kind");
    if (fr->full_every < 0) return fail(LV_ERR_ARG, "full_every must be >= 0");
    if (fr->heavy_degree < 0) return fail(LV_ERR_ARG, "heavy_degree must be >= 0");
    return LV_OK;
}

}  // namespace

extern "C" {

lv_status lv_run(lv_graph *g, float min_gain, lv_stats *stats) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    return run_levels(g, min_gain, nullptr, stats);
}

lv_status lv_run_pruned(lv_graph *g, float min_gain, const lv_frontier *fr, lv_stats *stats) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    LV_TRY(check_frontier(fr));
    return run_levels(g, min_gain, fr, stats);
}

lv_status lv_optimise_pruned(lv_graph *g, float min_gain, const lv_frontier *fr,
                             int32_t *changed, int32_t *sweeps, int64_t *evaluated) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    LV_TRY(check_frontier(fr));
    LV_HIP(hipSetDevice(g->device));
    return g->optimise(min_gain, changed, sweeps, nullptr, nullptr, fr, evaluated, nullptr);
}

lv_status lv_sweep_pruned(lv_graph *g, const lv_frontier *fr, int32_t track, double *q_after,
                          int64_t *moved, int64_t *evaluated) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    LV_TRY(check_frontier(fr));
    LV_HIP(hipSetDevice(g->device));
    double q1 = 0.0;
    int64_t mv = 0, ev = 0;
    LV_TRY(g->sweep(&q1, &mv, nullptr, fr, track != 0, &ev));
    if (q_after) *q_after = q1;
    if (moved) *moved = mv;
    if (evaluated) *evaluated = ev;
    return LV_OK;
}

lv_status lv_trace_phase(lv_graph *g, float min_gain, float *feat, uint8_t *moved,
                         int32_t max_sweeps, int32_t *traced, int32_t *changed) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    if (max_sweeps < 0 || (max_sweeps > 0 && (!feat || !moved)))
        return fail(LV_ERR_ARG, "trace buffers");
    LV_HIP(hipSetDevice(g->device));
    lv_graph::TraceSink t{feat, moved, max_sweeps, 0};
    int32_t sw = 0;
    LV_TRY(g->optimise(min_gain, changed, &sw, nullptr, nullptr, nullptr, nullptr, nullptr, &t));
    if (traced) *traced = t.n;
    return LV_OK;
}

lv_status lv_features(lv_graph *g, float *feat) {
    if (!g || !feat) return fail(LV_ERR_ARG, "null argument");
    LV_HIP(hipSetDevice(g->device));
    if (!g->level_ready) LV_TRY(g->setup_level());
    lv_graph::TraceSink t{feat, nullptr, 1, 0};
    return g->trace_features(&t);
}

lv_status lv_move_state(lv_graph *g, float *margin, float *kown, uint8_t *hist, int32_t *nmv) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    LV_HIP(hipSetDevice(g->device));
    if (!g->level_ready) LV_TRY(g->setup_level());
    // vertices without edges: never evaluated, never moved
    const float zf = 0.0f;
    const uint8_t zb = 0;
    const int32_t zi = 0;
    if (margin) LV_TRY(g->copy_out_rows<float>(g->margin, margin, 1, &zf));
    if (kown) LV_TRY(g->copy_out_rows<float>(g->kown, kown, 1, &zf));
    if (hist) LV_TRY(g->copy_out_rows<uint8_t>(g->hist, hist, 1, &zb));
    if (nmv) LV_TRY(g->copy_out_rows<int32_t>(g->nmv, nmv, 1, &zi));
    return g->sync();
}

lv_status lv_set_model(lv_graph *g, const float *params) {
    if (!g || !params) return fail(LV_ERR_ARG, "null argument");
    LV_HIP(hipSetDevice(g->device));
    LV_HIP(hipMemcpyAsync(g->mlp, params, sizeof(float) * lv::kParams, hipMemcpyHostToDevice,
                          g->stream));
    LV_TRY(g->sync());
    g->mlp_set = true;
    return LV_OK;
}

lv_status lv_predict(lv_graph *g, float threshold, float *prob, uint8_t *active) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    if (!g->mlp_set) return fail(LV_ERR_STATE, "frontier model not set (lv_set_model)");
    LV_HIP(hipSetDevice(g->device));
    if (!g->level_ready) LV_TRY(g->setup_level());
    // vertices without edges: the model's output for their features, never active
    float iso[lv::kFeat], piso = 0.0f;
    LV_TRY(g->isolated_defaults(iso, &piso));
    const uint8_t zb = 0;
    if (g->V <= 0) {
        std::vector<int32_t> none;
        if (prob) lv_graph::expand_rows<float>(nullptr, prob, none, g->Vref, 1, &piso);
        if (active) lv_graph::expand_rows<uint8_t>(nullptr, active, none, g->Vref, 1, &zb);
        return LV_OK;
    }
    LV_TRY(g->tfeat.reserve(g->V));
    lv::launch_frontier(g->feature_args(), nullptr, 0, g->mlp, threshold, threshold, 0, g->bb,
                        g->fkeys, g->active, g->tfeat.p, g->stream);
    LV_LAUNCHED("frontier model");
    if (prob) LV_TRY(g->copy_out_rows<float>(g->tfeat.p, prob, 1, &piso));
    if (active) LV_TRY(g->copy_out_rows<uint8_t>(g->active, active, 1, &zb));
    return g->sync();
}

lv_status lv_assignment(lv_graph *g, int32_t *orig2comm, int32_t *level_V) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    LV_HIP(hipSetDevice(g->device));
    if (orig2comm && g->Vorig > 0)
        LV_HIP(hipMemcpyAsync(orig2comm, g->orig2comm, sizeof(int32_t) * g->Vorig,
                              hipMemcpyDeviceToHost, g->stream));
    LV_TRY(g->sync());
    if (level_V) *level_V = g->Vref;
    return LV_OK;
}

lv_status lv_level_size(lv_graph *g, int32_t *V, int64_t *E) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    if (V) *V = g->Vref;
    if (E) *E = g->E;
    return LV_OK;
}

lv_status lv_level_graph(lv_graph *g, int32_t *rowptr, int32_t *col, float *w) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    LV_HIP(hipSetDevice(g->device));
    std::vector<int32_t> rof, rp_c;
    LV_TRY(g->host_maps(&rof, nullptr));
    const bool expand = !g->maps_identity;
    int32_t *rp_dst = rowptr;
    if (expand && rowptr) {
        rp_c.resize((size_t)g->V + 1);
        rp_dst = rp_c.data();
    }
    if (rp_dst)
        LV_HIP(hipMemcpyAsync(rp_dst, g->rowptr(), sizeof(int32_t) * ((size_t)g->V + 1),
                              hipMemcpyDeviceToHost, g->stream));
    if (col && g->E > 0)
        LV_HIP(hipMemcpyAsync(col, g->col(), sizeof(int32_t) * g->E, hipMemcpyDeviceToHost,
                              g->stream));
    if (w && g->E > 0)
        LV_HIP(hipMemcpyAsync(w, g->w(), sizeof(float) * g->E, hipMemcpyDeviceToHost, g->stream));
    LV_TRY(g->sync());
    if (!expand) return LV_OK;
    // rows of vertices without edges are empty; columns in reference ids
    if (col)
        for (int64_t e = 0; e < g->E; ++e) col[e] = rof[(size_t)col[e]];
    if (rowptr) {
        int32_t i = 0;
        for (int32_t r = 0; r <= g->Vref; ++r) {
            while (i < g->V && rof[(size_t)i] < r) ++i;
            rowptr[r] = rp_c[(size_t)i];
        }
    }
    return LV_OK;
}

namespace {
const char kMapMagic[8] = {'L', 'V', 'O', '2', 'C', '0', '1', '\0'};
}

lv_status lv_dump_level(lv_graph *g, const char *prefix, int32_t with_mtx) {
    if (!g || !prefix) return fail(LV_ERR_ARG, "null argument");
    LV_HIP(hipSetDevice(g->device));
    int32_t V = 0;
    int64_t E = 0;
    LV_TRY(lv_level_size(g, &V, &E));
    std::vector<int32_t> rp((size_t)V + 1), cl((size_t)std::max<int64_t>(E, 1));
    std::vector<float> w((size_t)std::max<int64_t>(E, 1));
    LV_TRY(lv_level_graph(g, rp.data(), cl.data(), w.data()));
    std::vector<int32_t> o2c((size_t)std::max(g->Vorig, 1));
    int32_t lv_V = 0;
    LV_TRY(lv_assignment(g, o2c.data(), &lv_V));
    const std::string base(prefix);
    LV_TRY(lv_write_csr((base + ".lvcsr").c_str(), rp.data(), cl.data(), w.data(), V, E));
    if (with_mtx) LV_TRY(lv_write_mtx((base + ".mtx").c_str(), rp.data(), cl.data(), w.data(), V));
    const std::string mp = base + ".o2c";
    FILE *f = std::fopen(mp.c_str(), "wb");
    if (!f) return fail(LV_ERR_IO, "cannot create '" + mp + "'");
    const int64_t hdr[2] = {(int64_t)g->Vorig, (int64_t)V};
    const double Mrun = g->M;
    bool ok = std::fwrite(kMapMagic, 1, 8, f) == 8 && std::fwrite(hdr, sizeof hdr, 1, f) == 1 &&
              std::fwrite(&Mrun, sizeof Mrun, 1, f) == 1 &&
              std::fwrite(o2c.data(), sizeof(int32_t), (size_t)g->Vorig, f) == (size_t)g->Vorig;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(LV_ERR_IO, "cannot write '" + mp + "'");
    return LV_OK;
}

lv_status lv_create_from_dump(const char *prefix, const lv_options *opt, lv_graph **out) {
    if (!prefix || !out) return fail(LV_ERR_ARG, "null argument");
    *out = nullptr;
    const std::string base(prefix);
    const std::string mp = base + ".o2c";
    std::vector<int32_t> o2c;
    int64_t hdr[2] = {0, 0};
    double Mrun = 0.0;
    {
        FILE *f = std::fopen(mp.c_str(), "rb");
        if (!f) return fail(LV_ERR_IO, "cannot open '" + mp + "'");
        char magic[8] = {0};
        bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMapMagic, 8) == 0 &&
                  std::fread(hdr, sizeof hdr, 1, f) == 1 && hdr[0] >= 0 && hdr[0] <= 0x7fffffffLL &&
                  std::fread(&Mrun, sizeof Mrun, 1, f) == 1;
        if (ok) {
            o2c.resize((size_t)std::max<int64_t>(hdr[0], 1));
            ok = std::fread(o2c.data(), sizeof(int32_t), (size_t)hdr[0], f) == (size_t)hdr[0];
        }
        std::fclose(f);
        if (!ok) return fail(LV_ERR_ARG, "'" + mp + "' is not a vertex map of a level dump");
    }
    int32_t *rp = nullptr, *cl = nullptr;
    float *w = nullptr;
    int32_t V = 0;
    int64_t E = 0;
    LV_TRY(lv_read_csr((base + ".lvcsr").c_str(), &rp, &cl, &w, &V, &E));
    lv_status st = (int64_t)V == hdr[1] ? LV_OK : fail(LV_ERR_ARG, "dump files disagree on V");
    lv_graph *g = nullptr;
    if (st == LV_OK) st = lv_create(rp, cl, w, V, E, opt, &g);
    lv_free_host(rp);
    lv_free_host(cl);
    lv_free_host(w);
    if (st == LV_OK) {
        st = g->set_initial_map(o2c.data(), (int32_t)hdr[0], Mrun);
        if (st != LV_OK) {
            lv_destroy(g);
            return st;
        }
        *out = g;
    }
    return st;
}

lv_status lv_get_communities(lv_graph *g, int32_t *comm) {
    if (!g || !comm) return fail(LV_ERR_ARG, "null argument");
    LV_HIP(hipSetDevice(g->device));
    if (!g->level_ready) LV_TRY(g->setup_level());
    return g->communities_out(g->comm, comm);
}

lv_status lv_set_communities(lv_graph *g, const int32_t *comm) {
    if (!g || !comm) return fail(LV_ERR_ARG, "null argument");
    for (int32_t v = 0; v < g->Vref; ++v)
        if (comm[v] < 0 || comm[v] >= g->Vref) return fail(LV_ERR_ARG, "community id out of range");
    LV_HIP(hipSetDevice(g->device));
    if (!g->level_ready) LV_TRY(g->setup_level());
    if (g->maps_identity) return g->set_state(comm);
    // a compact level stores only vertices with edges: those without keep
    // their own singleton community and nobody joins it
    std::vector<int32_t> rof, cof;
    LV_TRY(g->host_maps(&rof, &cof));
    for (int32_t r = 0; r < g->Vref; ++r)
        if (cof[(size_t)r] < 0 && comm[r] != r)
            return fail(LV_ERR_ARG, "a vertex without edges must stay in its own community");
    std::vector<int32_t> c((size_t)std::max(g->V, 1));
    for (int32_t i = 0; i < g->V; ++i) {
        const int32_t x = comm[rof[(size_t)i]];
        if (cof[(size_t)x] < 0)
            return fail(LV_ERR_ARG, "no vertex can join the community of a vertex without edges");
        c[(size_t)i] = cof[(size_t)x];
    }
    return g->set_state(c.data());
}

lv_status lv_decide(lv_graph *g, int32_t group, int32_t *new_comm) {
    if (!g || !new_comm) return fail(LV_ERR_ARG, "null argument");
    if (group < 0 || group >= lv::kNumGroups) return fail(LV_ERR_ARG, "group out of range");
    LV_HIP(hipSetDevice(g->device));
    if (!g->level_ready) LV_TRY(g->setup_level());
    if (g->V > 0) {
        LV_HIP(hipMemcpyAsync(g->newc, g->comm, sizeof(int32_t) * g->V, hipMemcpyDeviceToDevice,
                              g->stream));
        LV_TRY(g->run_group(group));
    }
    return g->communities_out(g->newc, new_comm);
}

lv_status lv_kernel_timing(lv_graph *g, int32_t kernel, lv_kernel_time *out) {
    if (!g || !out || kernel < 0 || kernel >= LV_K_COUNT) return fail(LV_ERR_ARG, "bad argument");
    LV_TRY(g->sync());
    *out = g->ktime[kernel];
    return LV_OK;
}

lv_status lv_kernel_timing_reset(lv_graph *g) {
    if (!g) return fail(LV_ERR_ARG, "null graph");
    LV_TRY(g->sync());
    for (auto &t : g->ktime) t = lv_kernel_time{};
    for (auto &t : g->ktd) t = lv_kernel_time{};
    return LV_OK;
}

lv_status lv_kernel_timing_detail(lv_graph *g, int32_t kernel, lv_kernel_time *out) {
    if (!g || !out || kernel < 0 || kernel >= LV_KD_COUNT) return fail(LV_ERR_ARG, "bad argument");
    LV_TRY(g->sync());
    *out = g->ktd[kernel];
    return LV_OK;
}

// ------------------------------------------------------------ frontier model

}  // extern "C"

namespace {

// Device buffers of one standalone model call, freed on return.
struct ModelBuffers {
    std::vector<void *> ptrs;
    template <typename T>
    lv_status get(T **p, int64_t n) {
        LV_TRY(dev_alloc(p, n));
        ptrs.push_back(*p);
        return LV_OK;
    }
    ~ModelBuffers() {
        for (void *p : ptrs) (void)hipFree(p);
    }
};

lv_status model_device(int32_t device) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(LV_ERR_NODEVICE, "no HIP device visible");
    if (device < 0 || device >= ndev) return fail(LV_ERR_ARG, "device ordinal out of range");
    LV_HIP(hipSetDevice(device));
    return LV_OK;
}

}  // namespace

extern "C" {

lv_status lv_mlp_forward(int32_t device, const float *params, const float *feat, int64_t n,
                         float *prob) {
    if (!params || n < 0 || (n > 0 && (!feat || !prob))) return fail(LV_ERR_ARG, "bad argument");
    LV_TRY(model_device(device));
    if (n == 0) return LV_OK;
    ModelBuffers mb;
    float *dP, *dX, *dY;
    LV_TRY(mb.get(&dP, lv::kParams));
    LV_TRY(mb.get(&dX, n * lv::kFeat));
    LV_TRY(mb.get(&dY, n));
    LV_HIP(hipMemcpy(dP, params, sizeof(float) * lv::kParams, hipMemcpyHostToDevice));
    LV_HIP(hipMemcpy(dX, feat, sizeof(float) * n * lv::kFeat, hipMemcpyHostToDevice));
    lv::launch_mlp_forward(dX, n, dP, dY, nullptr);
    LV_LAUNCHED("mlp forward");
    LV_HIP(hipMemcpy(prob, dY, sizeof(float) * n, hipMemcpyDeviceToHost));
    return LV_OK;
}

lv_status lv_mlp_grad(int32_t device, const float *params, const float *feat, const float *labels,
                      int64_t n, float pos_weight, float *grad, float *loss) {
    if (!params || !grad || n <= 0 || !feat || !labels) return fail(LV_ERR_ARG, "bad argument");
    LV_TRY(model_device(device));
    ModelBuffers mb;
    float *dP, *dX, *dL, *dpart, *dG;
    const int64_t nb = lv::grad_blocks(n);
    LV_TRY(mb.get(&dP, lv::kParams));
    LV_TRY(mb.get(&dX, n * lv::kFeat));
    LV_TRY(mb.get(&dL, n));
    LV_TRY(mb.get(&dpart, nb * (lv::kParams + 1)));
    LV_TRY(mb.get(&dG, lv::kParams + 1));
    LV_HIP(hipMemcpy(dP, params, sizeof(float) * lv::kParams, hipMemcpyHostToDevice));
    LV_HIP(hipMemcpy(dX, feat, sizeof(float) * n * lv::kFeat, hipMemcpyHostToDevice));
    LV_HIP(hipMemcpy(dL, labels, sizeof(float) * n, hipMemcpyHostToDevice));
    lv::launch_train_grad(dX, dL, n, dP, pos_weight, dpart, nullptr);
    lv::launch_train_reduce(dpart, nb, n, dG, nullptr);
    LV_LAUNCHED("mlp gradient");
    std::vector<float> h(lv::kParams + 1);
    LV_HIP(hipMemcpy(h.data(), dG, sizeof(float) * (lv::kParams + 1), hipMemcpyDeviceToHost));
    std::memcpy(grad, h.data(), sizeof(float) * lv::kParams);
    if (loss) *loss = h[lv::kParams];
    return LV_OK;
}

lv_status lv_mlp_train(int32_t device, const float *feat, const float *labels, int64_t n,
                       const lv_train_options *opt, float *params, float *loss_per_epoch) {
    if (!feat || !labels || !params || !opt || n <= 0) return fail(LV_ERR_ARG, "bad argument");
    if (opt->epochs < 0 || opt->batch <= 0) return fail(LV_ERR_ARG, "epochs/batch");
    LV_TRY(model_device(device));
    ModelBuffers mb;
    const int64_t batch = std::min<int64_t>(opt->batch, n);
    const int64_t nb = lv::grad_blocks(batch);
    float *dP, *dX, *dL, *dpart, *dG, *dm, *dv;
    double *dloss;
    const int32_t epochs = opt->epochs;
    LV_TRY(mb.get(&dP, lv::kParams));
    LV_TRY(mb.get(&dm, lv::kParams));
    LV_TRY(mb.get(&dv, lv::kParams));
    LV_TRY(mb.get(&dX, n * lv::kFeat));
    LV_TRY(mb.get(&dL, n));
    LV_TRY(mb.get(&dpart, nb * (lv::kParams + 1)));
    LV_TRY(mb.get(&dG, lv::kParams + 1));
    LV_TRY(mb.get(&dloss, std::max(epochs, 1)));
    LV_HIP(hipMemcpy(dP, params, sizeof(float) * lv::kParams, hipMemcpyHostToDevice));
    LV_HIP(hipMemcpy(dX, feat, sizeof(float) * n * lv::kFeat, hipMemcpyHostToDevice));
    LV_HIP(hipMemcpy(dL, labels, sizeof(float) * n, hipMemcpyHostToDevice));
    lv::fill_i32(reinterpret_cast<int32_t *>(dm), 0, lv::kParams, nullptr);
    lv::fill_i32(reinterpret_cast<int32_t *>(dv), 0, lv::kParams, nullptr);
    lv::fill_f64(dloss, 0.0, std::max(epochs, 1), nullptr);
    const float b1 = opt->beta1 > 0.0f ? opt->beta1 : 0.9f;
    const float b2 = opt->beta2 > 0.0f ? opt->beta2 : 0.999f;
    const float eps = opt->eps > 0.0f ? opt->eps : 1e-8f;
    const float pw = opt->pos_weight > 0.0f ? opt->pos_weight : 1.0f;
    int32_t t = 0;
    for (int32_t e = 0; e < epochs; ++e) {
        for (int64_t lo = 0; lo < n; lo += batch) {
            const int64_t bn = std::min(batch, n - lo);
            lv::launch_train_grad(dX + lo * lv::kFeat, dL + lo, bn, dP, pw, dpart, nullptr);
            lv::launch_train_reduce(dpart, lv::grad_blocks(bn), bn, dG, nullptr);
            lv::launch_accumulate_loss(dG, (double)bn / (double)n, dloss + e, nullptr);
            lv::launch_adam(dP, dG, dm, dv, ++t, opt->lr, b1, b2, eps, nullptr);
        }
        LV_LAUNCHED("mlp training step");
    }
    LV_HIP(hipMemcpy(params, dP, sizeof(float) * lv::kParams, hipMemcpyDeviceToHost));
    if (loss_per_epoch && epochs > 0) {
        std::vector<double> h(epochs);
        LV_HIP(hipMemcpy(h.data(), dloss, sizeof(double) * epochs, hipMemcpyDeviceToHost));
        for (int32_t e = 0; e < epochs; ++e) loss_per_epoch[e] = (float)h[e];
    }
    return LV_OK;
}

// ------------------------------------------------------------ batches

struct lv_batch {
    lv_graph *g;
};

lv_status lv_batch_create(int32_t G, const int32_t *vptr, const int32_t *rowptr, const int32_t *col,
                          const float *w, int64_t E, const lv_options *opt, lv_batch **out) {
    if (!out) return fail(LV_ERR_ARG, "out is null");
    *out = nullptr;
    if (G <= 0 || !vptr) return fail(LV_ERR_ARG, "no graphs");
    if (vptr[0] != 0) return fail(LV_ERR_ARG, "vptr[0] must be 0");
    for (int32_t g = 0; g < G; ++g)
        if (vptr[g + 1] < vptr[g]) return fail(LV_ERR_ARG, "vptr not monotone");
    const int32_t V = vptr[G];
    // shape first (rowptr monotone, E in int32, columns in range), then ranges
    LV_TRY(check_csr(rowptr, col, w, V, E));
    for (int32_t g = 0; g < G; ++g)
        for (int32_t v = vptr[g]; v < vptr[g + 1]; ++v)
            for (int32_t e = rowptr[v]; e < rowptr[v + 1]; ++e)
                if (col[e] < vptr[g] || col[e] >= vptr[g + 1])
                    return fail(LV_ERR_ARG, "an edge leaves its graph's vertex range");
    // each graph on its own grid, as lv_create would round it alone
    std::vector<float> wq;
    if (!opt || opt->exact_sums >= 0)
        for (int32_t g = 0; g < G; ++g)
            LV_TRY(grid_round(w, E, rowptr[vptr[g]], rowptr[vptr[g + 1]], wq));
    const float *wu = wq.empty() ? w : wq.data();
    lv_graph *h = nullptr;
    // graph ranges stay in one numbering: batches store every vertex
    lv_options o = opt ? *opt : lv_options{};
    o.compact = -1;
    LV_TRY(create_handle(rowptr, col, wu, V, E, &o, &h));
    lv_status st = h->setup_batch(G, vptr, rowptr, wu);
    if (st != LV_OK) {
        std::string msg = g_last_error;
        lv_destroy(h);
        g_last_error = msg;
        return st;
    }
    lv_batch *b = new (std::nothrow) lv_batch{h};
    if (!b) {
        lv_destroy(h);
        return fail(LV_ERR_OOM, "host allocation failed");
    }
    *out = b;
    return LV_OK;
}

void lv_batch_destroy(lv_batch *b) {
    if (!b) return;
    lv_destroy(b->g);
    delete b;
}

lv_status lv_batch_run(lv_batch *b, float min_gain, lv_batch_result *per_graph, lv_stats *total) {
    if (!b) return fail(LV_ERR_ARG, "null batch");
    LV_HIP(hipSetDevice(b->g->device));
    return b->g->run_batch(min_gain, per_graph, total);
}

lv_status lv_batch_set_model(lv_batch *b, const float *params) {
    if (!b) return fail(LV_ERR_ARG, "null batch");
    return lv_set_model(b->g, params);
}

lv_status lv_batch_run_pruned(lv_batch *b, float min_gain, const lv_frontier *fr,
                              lv_batch_result *per_graph, lv_stats *total) {
    if (!b) return fail(LV_ERR_ARG, "null batch");
    LV_TRY(check_frontier(fr));
    if (fr->kind != LV_FRONTIER_MODEL)
        return fail(LV_ERR_ARG, "batched pruned runs use the model frontier");
    if (!b->g->mlp_set) return fail(LV_ERR_STATE, "frontier model not set (lv_batch_set_model)");
    LV_HIP(hipSetDevice(b->g->device));
    return b->g->run_batch(min_gain, per_graph, total, fr);
}

lv_status lv_batch_assignment(lv_batch *b, int32_t *comm) {
    if (!b || !comm) return fail(LV_ERR_ARG, "null argument");
    lv_graph *g = b->g;
    LV_HIP(hipSetDevice(g->device));
    if (g->V0 > 0)
        LV_HIP(hipMemcpyAsync(comm, g->gfinal, sizeof(int32_t) * g->V0, hipMemcpyDeviceToHost,
                              g->stream));
    return g->sync();
}

lv_status lv_batch_set_options(lv_batch *b, const lv_options *opt) {
    if (!b || !opt) return fail(LV_ERR_ARG, "null argument");
    lv_graph *g = b->g;
    if (opt->device != g->device) return fail(LV_ERR_ARG, "the device of a batch cannot change");
    LV_HIP(hipSetDevice(g->device));
    LV_TRY(g->sync());
    const lv_options was = g->opt;
    g->opt = *opt;
    g->opt.timing = was.timing;
    g->opt.compact = was.compact;  // batches store every vertex
    g->opt.exact_sums = was.exact_sums;
    g->apply_options();
    // every run starts from level 0 (reset), which is bucketed again
    g->level_ready = false;
    return LV_OK;
}

lv_status lv_batch_kernel_timing(lv_batch *b, int32_t kernel, lv_kernel_time *out) {
    if (!b) return fail(LV_ERR_ARG, "null batch");
    return lv_kernel_timing(b->g, kernel, out);
}

lv_status lv_set_options(lv_graph *g, const lv_options *opt) {
    if (!g || !opt) return fail(LV_ERR_ARG, "null argument");
    if (opt->device != g->device) return fail(LV_ERR_ARG, "the device of a handle cannot change");
    LV_HIP(hipSetDevice(g->device));
    LV_TRY(g->sync());
    const lv_options was = g->opt;
    g->opt = *opt;
    // fixed at creation: the timing flag, the level storage and the weights
    g->opt.timing = was.timing;
    g->opt.compact = was.compact;
    g->opt.exact_sums = was.exact_sums;
    g->apply_options();
    // the current level is bucketed again before its next use; its
    // community state is kept
    if (g->level_ready) {
        std::vector<int32_t> keep((size_t)std::max(g->V, 1));
        if (g->V > 0)
            LV_HIP(hipMemcpy(keep.data(), g->comm, sizeof(int32_t) * g->V, hipMemcpyDeviceToHost));
        g->level_ready = false;
        LV_TRY(g->setup_level());
        if (g->V > 0) LV_TRY(g->set_state(keep.data()));
    }
    return LV_OK;
}

}  // extern "C"


