// Shared device helpers and launch declarations of the Louvain engine.
// gfx950 (CDNA4): 64-lane wavefronts, LDS FP64 atomics (ds_add_f64),
// global FP64 atomics (global_atomic_add_f64, coarse-grained memory).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lv {

// ---------------------------------------------------------------------
// Phase-1 execution buckets.  Buckets 0..5 coincide with the reference's
// degree groups 0..5 (modularity_optimisation.cuh:7-19); group 6 (deg > 319)
// is split by degree into two one-workgroup LDS tiers (tables sized for
// their degree bound, so the lower one runs more workgroups per CU) and a
// chunked tier (rows split over many workgroups, merged by partition in
// LDS).  Splitting a group changes only the kernel mapping: decisions inside
// a group are simultaneous.
// ---------------------------------------------------------------------
constexpr int kNumBuckets = 12;
constexpr int kNumGroups = 7;
constexpr int kBucketGroup[] = {0, 1, 2, 3, 4, 5, 5, 6, 6, 6, 6, 6};
static_assert(sizeof(kBucketGroup) / sizeof(kBucketGroup[0]) == kNumBuckets,
              "one group per bucket");
constexpr int kGroupBound[kNumGroups + 1] = {0, 4, 8, 16, 32, 84, 319, 0x7fffffff};

struct BucketBounds {
    int32_t hi[kNumBuckets];  // bucket b holds hi[b-1] < deg <= hi[b]
};

__host__ __device__ inline int bucket_of(int32_t deg, const BucketBounds &bb) {
    if (deg <= 0) return -1;
#pragma unroll
    for (int b = 0; b < kNumBuckets; ++b)
        if (deg <= bb.hi[b]) return b;
    return kNumBuckets - 1;
}

// Move-kernel operands.
struct MoveArgs {
    const int32_t *rowptr;
    const int32_t *col;
    const float *w;
    const int32_t *comm;
    const int32_t *size;
    const double *k;
    const double *sigma;
    const uint64_t *crec;  // packed community record, see pack_rec
    int32_t *newc;         // decision
    double *dk;            // mover: K[new] - K[cur] (weight change with neighbours fixed)
    float *margin;         // evaluated vertex: bestGain - K[cur]/M (move iff > 0)
    float *kown;           // evaluated vertex: K[cur] (fp32)
    const uint8_t *active; // frontier flags (chunked tier; null = every vertex)
    float M32;  // (float)M
    float den;  // (2.0f * M32) * M32, fp32 as the reference evaluates it
    // Batches of independent graphs: gid[v] = graph of vertex v and the
    // graph's own M32 / den (null for a single graph).
    const int32_t *gid;
    const float *gM32;
    const float *gden;
    // bounds: candidates are evaluated best-first and skipped when a bound
    // shows they cannot change the decision (set only when every weight is
    // >= 0).  obs: the margin observation must be exact too (sweeps whose
    // observations a frontier reads, lv_decide); otherwise candidates that
    // cannot pass the move test are skipped as well.
    int32_t bounds;
    int32_t obs;
    // Sweeps enqueued ahead of the termination test (a device-resident
    // phase): null, or the phase's end flag; kernels return at once when it
    // is set, so a sweep enqueued after the last one does nothing.
    const int32_t *gate;
    // Row gathers of the move kernels.  Entry e of the rows names column
    // ccol[e] and its community is ccomm[ccol[e]]; vertex v's own number
    // among the columns is cself[v].  Without renumbering: ccol = col, ccomm
    // = comm, cself = null (identity).  With it, columns are numbered in
    // bucket order (the rows of high degree, which most entries name, lie
    // together) and ccomm is comm in that order, kept by the applies.
    const int32_t *ccol;
    const int32_t *ccomm;
    const int32_t *cself;
};

__device__ __forceinline__ bool gated_off(const int32_t *gate) { return gate && *gate != 0; }

// Vertex v's number among the move kernels' columns (MoveArgs::cself).
__device__ __forceinline__ int32_t col_self(const MoveArgs &a, int32_t v) {
    return a.cself ? a.cself[v] : v;
}

// M32 and den of vertex v's graph (.x, .y).
__device__ __forceinline__ float2 vertex_mden(const MoveArgs &a, int32_t v) {
    if (a.gid) {
        const int32_t g = a.gid[v];
        return make_float2(a.gM32[g], a.gden[g]);
    }
    return make_float2(a.M32, a.den);
}

// Community record read by the move kernels: the fp32 rounding of the FP64
// community total (the value the gain uses) and the community size, in one
// 8-byte word so that a candidate costs one gather.
__host__ __device__ __forceinline__ uint64_t pack_rec(double sigma, int32_t size) {
    const float s = (float)sigma;
    uint32_t bits;
    __builtin_memcpy(&bits, &s, 4);
    return ((uint64_t)(uint32_t)size << 32) | bits;
}
__device__ __forceinline__ float rec_sigma(uint64_t r) { return __uint_as_float((uint32_t)r); }
__device__ __forceinline__ int32_t rec_size(uint64_t r) { return (int32_t)(r >> 32); }

// computeGain (modularity_optimisation.cu:96-102), fp32 in the same order:
//   K/M + k * ((S_cur - k) - S_c) / (2*M*M).  No contraction into FMAs.
__device__ __forceinline__ float gain_of(float Kc, float M32, float den, float k32, float scur,
                                         float sc) {
#pragma clang fp contract(off)
    float cs = scur - k32;
    float num = k32 * (cs - sc);
    float a = Kc / M32;
    float b = num / den;
    return a + b;
}

// Candidate ordering of computeMove (modularity_optimisation.cu:157,169):
// larger gain wins, equal gains -> lower community id.  Packed into one
// 64-bit key so that an unsigned max is the argmax: high word = gain mapped
// to an order-preserving unsigned, low word = ~community.
__device__ __forceinline__ uint64_t pack_best(float g, int32_t c) {
    uint32_t u = __float_as_uint(g + 0.0f);  // -0 -> +0 (IEEE-equal values)
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return ((uint64_t)u << 32) | (uint64_t)(~(uint32_t)c);
}
__device__ __forceinline__ float best_gain(uint64_t key) {
    uint32_t u = (uint32_t)(key >> 32);
    u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    return __uint_as_float(u);
}
__device__ __forceinline__ int32_t best_comm(uint64_t key) {
    return (int32_t)(~(uint32_t)(key & 0xffffffffu));
}

// Final test of computeMove (modularity_optimisation.cu:192-196).
__device__ __forceinline__ int32_t decide_from(uint64_t best, double kcur, float M32,
                                               int32_t cur) {
#pragma clang fp contract(off)
    float bg = best_gain(best);
    float t = (float)kcur / M32;
    return (bg - t > 0.0f) ? best_comm(best) : cur;
}

// Candidate c with weight Kc from v (strength k32) in cur: singleton rule
// (modularity_optimisation.cu:154-155), then the gain folded into the
// packed argmax.  rc = record of c.
__device__ __forceinline__ void consider(uint64_t &best, int32_t c, double Kc, uint64_t rc,
                                         int32_t cur, int32_t size_cur, float k32, float scur,
                                         float M32, float den) {
    if (!(c < cur || size_cur > 1 || rec_size(rc) > 1)) return;
    const float g = gain_of((float)Kc, M32, den, k32, scur, rec_sigma(rc));
    const uint64_t key = pack_best(g, c);
    best = key > best ? key : best;
}

// The float just below x (toward -inf; finite x).
__device__ __forceinline__ float float_below(float x) {
    const uint32_t b = __float_as_uint(x);
    if (x > 0.0f) return __uint_as_float(b - 1u);
    if (x == 0.0f) return -__uint_as_float(1u);
    return __uint_as_float(b + 1u);
}

// Upper bound of the fp32 gain of a candidate with weight Kc whose total S_c
// has not been read.  With weights >= 0, S_c >= K_c (each member's strength
// counts its edges to v); their FP64 sums may round apart by far less than an
// fp32 ulp, so fp32(S_c) >= float_below(fp32(K_c)).  gain_of falls as S_c
// grows (each fp32 step is monotone and k32 >= 0), so the gain at that lower
// bound bounds the candidate's gain from above.
__device__ __forceinline__ float gain_bound(double Kc, float M32, float den, float k32,
                                            float scur) {
    const float kf = (float)Kc;
    return gain_of(kf, M32, den, k32, scur, float_below(kf));
}

// Whether a candidate whose gain is at most U can change the decision: its
// key must beat best0 (a candidate already evaluated, or the start key), and
// when the margin need not be exact it must also pass decide_from's move test
// U - kcur/M > 0 (t = (float)kcur / M32, as decide_from computes it).
__device__ __forceinline__ bool may_win(float U, int32_t c, uint64_t best0, bool obs, float t) {
#pragma clang fp contract(off)
    if (pack_best(U, c) <= best0) return false;
    if (!obs && !(U - t > 0.0f)) return false;
    return true;
}

__device__ __forceinline__ uint32_t hash_u32(int32_t c) {
    return (uint32_t)c * 2654435761u;
}

// Partition of a community among P merge workgroups; mixes differently from
// hash_u32 so that slots inside a partition stay uniformly used.
__device__ __forceinline__ uint32_t part_of(int32_t c, uint32_t P) {
    uint32_t x = (uint32_t)c;
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return (uint32_t)(((uint64_t)x * P) >> 32);
}

// Open-addressing insert with linear probing; keys -1 = empty.  The table
// must hold more slots than distinct keys.  Values are FP64 sums (exact for
// weights on a dyadic grid, hence independent of insertion order).
__device__ __forceinline__ int lds_insert(int32_t *keys, double *vals, uint32_t shift,
                                          uint32_t mask, int32_t c, double w) {
    uint32_t h = hash_u32(c) >> shift;
    for (;;) {
        int32_t k = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (k == c) {
            atomicAdd(&vals[h], w);
            return (int)h;
        }
        if (k == -1) {
            int32_t old = atomicCAS(&keys[h], -1, c);
            if (old == -1 || old == c) {
                atomicAdd(&vals[h], w);
                return (int)h;
            }
        }
        h = (h + 1) & mask;
    }
}

// Tables of any size n: c's first slot is the high word of hash * n (for a
// power of two n the same slot as hash >> (32 - log2 n)); linear probing.
__device__ __forceinline__ uint32_t slot_of(int32_t c, uint32_t n) {
    return (uint32_t)(((uint64_t)hash_u32(c) * n) >> 32);
}
__device__ __forceinline__ void lds_insert_n(int32_t *keys, double *vals, uint32_t n, int32_t c,
                                             double w) {
    uint32_t h = slot_of(c, n);
    for (;;) {
        int32_t k = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (k == c) {
            atomicAdd(&vals[h], w);
            return;
        }
        if (k == -1) {
            int32_t old = atomicCAS(&keys[h], -1, c);
            if (old == -1 || old == c) {
                atomicAdd(&vals[h], w);
                return;
            }
        }
        h = h + 1 == n ? 0u : h + 1;
    }
}
// lds_insert_n that reports whether it turned an empty slot into c (the
// table's count of distinct keys grows by one).
__device__ __forceinline__ bool lds_insert_claim(int32_t *keys, double *vals, uint32_t n,
                                                 int32_t c, double w) {
    uint32_t h = slot_of(c, n);
    for (;;) {
        int32_t k = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (k == c) {
            atomicAdd(&vals[h], w);
            return false;
        }
        if (k == -1) {
            int32_t old = atomicCAS(&keys[h], -1, c);
            if (old == -1 || old == c) {
                atomicAdd(&vals[h], w);
                return old == -1;
            }
        }
        h = h + 1 == n ? 0u : h + 1;
    }
}
__device__ __forceinline__ double lds_value_n(const int32_t *keys, const double *vals, uint32_t n,
                                              int32_t c) {
    uint32_t h = slot_of(c, n);
    for (;;) {
        const int32_t k = keys[h];
        if (k == c) return vals[h];
        if (k == -1) return 0.0;
        h = h + 1 == n ? 0u : h + 1;
    }
}

// Value under key c in an LDS table after its inserts (0 when absent); the
// table keeps at least one empty slot.
__device__ __forceinline__ double lds_value(const int32_t *keys, const double *vals,
                                            uint32_t shift, uint32_t mask, int32_t c) {
    uint32_t h = hash_u32(c) >> shift;
    for (;;) {
        const int32_t k = keys[h];
        if (k == c) return vals[h];
        if (k == -1) return 0.0;
        h = (h + 1) & mask;
    }
}

__device__ __forceinline__ void global_insert(int32_t *keys, double *vals, uint32_t shift,
                                              uint32_t mask, int32_t c, double w) {
    uint32_t h = hash_u32(c) >> shift;
    for (;;) {
        int32_t k = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == c) {
            atomicAdd(&vals[h], w);
            return;
        }
        if (k == -1) {
            int32_t old = atomicCAS(&keys[h], -1, c);
            if (old == -1 || old == c) {
                atomicAdd(&vals[h], w);
                return;
            }
        }
        h = (h + 1) & mask;
    }
}

// Wave-level reductions over a 64-lane wavefront.
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        uint64_t y = __shfl_xor(x, off, 64);
        x = y > x ? y : x;
    }
    return x;
}
__device__ __forceinline__ double wave_sum_f64(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}
__device__ __forceinline__ int32_t wave_sum_i32(int32_t x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}
__device__ __forceinline__ int64_t wave_sum_i64(int64_t x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

constexpr int kWave = 64;

}  // namespace lv
