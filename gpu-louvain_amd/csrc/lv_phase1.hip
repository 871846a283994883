// Phase 1 (modularity optimisation) kernels for gfx950.
//
// The move decision restates computeMove (modularity_optimisation.cu:111-198):
// for vertex v in community cur, K[c] = sum of w over row entries (u,w) with
// u != v and comm[u] = c; the best candidate c != cur (singleton rule
// :154-155) maximises the fp32 gain (:96-102), ties to the lowest id; v
// moves iff bestGain - K[cur]/M > 0 (:192-196).  Unlike the reference, the
// gain is evaluated once per distinct community after accumulation (the
// reference evaluates it per edge on partial sums, which selects the same
// maximum for non-negative weights), and K is accumulated in FP64.
//
// Kernel shapes (64-lane wavefronts):
//   k_move_sub<X,E> X = 4..32 lanes per vertex, E row entries per lane;
//                   per-community sums by an in-register all-pairs pass over
//                   the segment (row order, so sums are bitwise the
//                   sequential ones), argmax of packed keys by DPP
//                   butterflies.
//   k_move_wave<C,G> one wavefront per vertex, C-slot LDS hash per wave
//                   (G names the bucket's instance).
//   k_move_wg<T,i>  one T-lane workgroup per vertex, LDS hash sized per launch
//                   (i = 0..3: the workgroup buckets in degree order).
//   k_chunk_*       rows beyond one workgroup: stage A aggregates 1024-edge
//                   chunks in LDS and buckets the distinct pairs by hash
//                   partition; stage B merges each (vertex, partition) in
//                   LDS, evaluates gains and combines partitions through a
//                   packed 64-bit atomic max.
#include "lv_phase1.h"
#include "lv_prims.h"

#include <algorithm>

namespace lv {

namespace {

constexpr int ilog2_c(int x) { return x <= 1 ? 0 : 1 + ilog2_c(x >> 1); }

__device__ __forceinline__ double block_sum_f64(double x, double *lds) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    x = wave_sum_f64(x);
    __syncthreads();
    if (lane == 0) lds[wid] = x;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += lds[i];
    return t;
}

// Candidate evaluation that keeps the weight of the best candidate with it.
__device__ __forceinline__ void consider_k(uint64_t &best, double &bestK, int32_t c, double Kc,
                                           uint64_t rc, int32_t cur, int32_t size_cur,
                                           float k32, float scur, float M32, float den) {
    const uint64_t before = best;
    consider(best, c, Kc, rc, cur, size_cur, k32, scur, M32, den);
    if (best != before) bestK = Kc;
}

// ------------------------------------------------------------ lane exchange
// Partner value for a butterfly all-reduction over aligned groups of lanes.
// Steps 1 and 2 swap lanes inside quads (DPP quad_perm); step 4 mirrors the
// two quads of each 8-lane half-row and step 8 the two halves of each 16-lane
// row (DPP row_half_mirror / row_mirror).  The mirrors do not pair lane i
// with lane i ^ step, but after the earlier steps every lane of a block holds
// that block's value, so any lane of the other block serves.  Steps 16 and
// 32 are shuffles.  DPP is VALU work: no LDS traffic.  Every lane of a group
// must be active.
template <int STEP>
__device__ __forceinline__ int32_t xchg_i32(int32_t x) {
    if constexpr (STEP == 1) return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);
    else if constexpr (STEP == 2) return __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);
    else if constexpr (STEP == 4) return __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false);
    else if constexpr (STEP == 8) return __builtin_amdgcn_update_dpp(0, x, 0x140, 0xF, 0xF, false);
    else return __shfl_xor(x, STEP, 64);
}
template <int STEP>
__device__ __forceinline__ uint64_t xchg_u64(uint64_t x) {
    const int32_t lo = xchg_i32<STEP>((int32_t)(uint32_t)x);
    const int32_t hi = xchg_i32<STEP>((int32_t)(uint32_t)(x >> 32));
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
template <int STEP>
__device__ __forceinline__ double xchg_f64(double x) {
    return __longlong_as_double((long long)xchg_u64<STEP>((uint64_t)__double_as_longlong(x)));
}

// Sum over WIDTH lanes, in every lane (FP64 addition is commutative, so the
// lanes of a group end with the same value).
template <int WIDTH>
__device__ __forceinline__ double lanes_sum(double x) {
    if constexpr (WIDTH >= 2) x += xchg_f64<1>(x);
    if constexpr (WIDTH >= 4) x += xchg_f64<2>(x);
    if constexpr (WIDTH >= 8) x += xchg_f64<4>(x);
    if constexpr (WIDTH >= 16) x += xchg_f64<8>(x);
    if constexpr (WIDTH >= 32) x += xchg_f64<16>(x);
    if constexpr (WIDTH >= 64) x += xchg_f64<32>(x);
    return x;
}

template <int STEP>
__device__ __forceinline__ void max_pair_step(uint64_t &best, double &bestK) {
    const uint64_t ob = xchg_u64<STEP>(best);
    const double oK = xchg_f64<STEP>(bestK);
    if (ob > best) {
        best = ob;
        bestK = oK;
    }
}

// Max of packed keys over WIDTH lanes, carrying the weight, in every lane.
// Equal keys name the same community, hence carry the same weight.
template <int WIDTH>
__device__ __forceinline__ void lanes_max_pair(uint64_t &best, double &bestK) {
    if constexpr (WIDTH >= 2) max_pair_step<1>(best, bestK);
    if constexpr (WIDTH >= 4) max_pair_step<2>(best, bestK);
    if constexpr (WIDTH >= 8) max_pair_step<4>(best, bestK);
    if constexpr (WIDTH >= 16) max_pair_step<8>(best, bestK);
    if constexpr (WIDTH >= 32) max_pair_step<16>(best, bestK);
    if constexpr (WIDTH >= 64) max_pair_step<32>(best, bestK);
}

template <int STEP>
__device__ __forceinline__ void argmax_step(double &k, int32_t &src) {
    const double ok = xchg_f64<STEP>(k);
    const int32_t os = xchg_i32<STEP>(src);
    if (ok > k || (ok == k && os < src)) {
        k = ok;
        src = os;
    }
}

// Largest k over WIDTH lanes and the index that goes with it (ties to the
// lowest index), in every lane.
template <int WIDTH>
__device__ __forceinline__ void lanes_argmax_idx(double &k, int32_t &idx) {
    if constexpr (WIDTH >= 2) argmax_step<1>(k, idx);
    if constexpr (WIDTH >= 4) argmax_step<2>(k, idx);
    if constexpr (WIDTH >= 8) argmax_step<4>(k, idx);
    if constexpr (WIDTH >= 16) argmax_step<8>(k, idx);
    if constexpr (WIDTH >= 32) argmax_step<16>(k, idx);
    if constexpr (WIDTH >= 64) argmax_step<32>(k, idx);
}

// Lane (of WIDTH) holding the largest k; ties to the lowest lane.
template <int WIDTH>
__device__ __forceinline__ int lanes_argmax(double k, int lane) {
    int32_t src = lane;
    lanes_argmax_idx<WIDTH>(k, src);
    return src;
}

// Lane l of each quad (DPP quad_perm broadcast).
__device__ __forceinline__ int32_t quad_bcast(int32_t x, int l) {
    switch (l) {
    case 0: return __builtin_amdgcn_update_dpp(0, x, 0x00, 0xF, 0xF, false);
    case 1: return __builtin_amdgcn_update_dpp(0, x, 0x55, 0xF, 0xF, false);
    case 2: return __builtin_amdgcn_update_dpp(0, x, 0xAA, 0xF, 0xF, false);
    default: return __builtin_amdgcn_update_dpp(0, x, 0xFF, 0xF, 0xF, false);
    }
}

// Lane l of each aligned X-lane segment (lane = this lane's index in its
// segment): DPP for X = 4 and 8 (the other quad of an 8-lane segment through
// the half-row mirror), a shuffle above.
template <int X>
__device__ __forceinline__ int32_t seg_bcast(int32_t x, int l, int lane) {
    if constexpr (X == 4) {
        return quad_bcast(x, l);
    } else if constexpr (X == 8) {
        const int32_t y = quad_bcast(x, l & 3);
        const int32_t z = __builtin_amdgcn_update_dpp(0, y, 0x141, 0xF, 0xF, false);
        return ((lane ^ l) & 4) ? z : y;
    } else {
        return __shfl(x, l, X);
    }
}

// Decision plus dk = K[new] - K[cur] for a mover (0 otherwise), and the
// observations the frontier model reads (move margin, own-community weight).
__device__ __forceinline__ void write_decision(const MoveArgs &a, int32_t v, int32_t cur,
                                               uint64_t best, double bestK, double kc) {
#pragma clang fp contract(off)
    const float vM32 = vertex_mden(a, v).x;
    const int32_t nc = decide_from(best, kc, vM32, cur);
    a.newc[v] = nc;
    a.dk[v] = nc != cur ? bestK - kc : 0.0;
    // margin as computeMove sees it: bestGain starts at 0 (no candidate or
    // only negative gains leave it there)
    const float bg = best ? fmaxf(best_gain(best), 0.0f) : 0.0f;
    const float t = (float)kc / vM32;
    a.margin[v] = bg - t;
    a.kown[v] = (float)kc;
}

// ------------------------------------------------------------------ moves

// One X-lane segment per vertex; lane l holds row entries l, l + X, ...,
// l + (E-1) X (deg <= X * E by bucket).  Narrow segments with several
// entries per lane put more vertices, hence more independent gathers, in
// flight per wave; X <= 8 exchanges through DPP only.
// Decision of vertex v by its X-lane segment (lane = index in the segment;
// every lane of the segment calls it together).
template <int X, int E>
__device__ __forceinline__ void move_seg(const MoveArgs &a, int32_t v, int lane) {
    const int32_t rs = a.rowptr[v];
    const int32_t deg = a.rowptr[v + 1] - rs;  // <= X * E by bucket
    const int32_t cur = a.comm[v];
    const int32_t vs = col_self(a, v);
    int32_t u[E];
    float wf[E];
#pragma unroll
    for (int t = 0; t < E; ++t) {
        const int32_t j = t * X + lane;
        u[t] = j < deg ? a.ccol[rs + j] : vs;
        wf[t] = j < deg ? a.w[rs + j] : 0.0f;
    }
    int32_t c[E];
#pragma unroll
    for (int t = 0; t < E; ++t) c[t] = u[t] != vs ? a.ccomm[u[t]] : -1;
    // All-pairs pass in row order: K = sum of the weights of the entries
    // sharing c; the entry of lowest index owns the community.
    double K[E];
    bool own[E];
#pragma unroll
    for (int t = 0; t < E; ++t) {
        K[t] = 0.0;
        own[t] = c[t] >= 0;
    }
#pragma unroll
    for (int s = 0; s < E; ++s) {
#pragma unroll
        for (int l = 0; l < X; ++l) {
            const int32_t j = s * X + l;
            if (j < deg) {  // the same in the whole segment
                const int32_t cj = seg_bcast<X>(c[s], l, lane);
                const float wj = __int_as_float(seg_bcast<X>(__float_as_int(wf[s]), l, lane));
#pragma unroll
                for (int t = 0; t < E; ++t) {
                    if (cj == c[t]) {
                        K[t] += (double)wj;
                        if (j < t * X + lane) own[t] = false;
                    }
                }
            }
        }
    }
    double kc = 0.0;
#pragma unroll
    for (int t = 0; t < E; ++t)
        if (own[t] && c[t] == cur) kc = K[t];
    kc = lanes_sum<X>(kc);  // one lane holds it

    const float2 md = vertex_mden(a, v);
    const float vM32 = md.x, vden = md.y;
    const uint64_t rcur = a.crec[cur];
    const float k32 = (float)a.k[v], scur = rec_sigma(rcur);
    const int32_t size_cur = rec_size(rcur);
    uint64_t best = pack_best(0.0f, cur);
    double bestK = 0.0;
    bool cand[E];
#pragma unroll
    for (int t = 0; t < E; ++t) cand[t] = own[t] && c[t] != cur;
    uint64_t rc[E];
    if (a.bounds) {
        // the heaviest candidate first; the others are read only if their
        // bound can still change the decision
        double hk = -1.0;
        int32_t hj = 0x7fffffff;
#pragma unroll
        for (int t = 0; t < E; ++t)
            if (cand[t] && K[t] > hk) {
                hk = K[t];
                hj = t * X + lane;
            }
        lanes_argmax_idx<X>(hk, hj);
#pragma unroll
        for (int t = 0; t < E; ++t)
            if (t * X + lane == hj)
                consider_k(best, bestK, c[t], K[t], a.crec[c[t]], cur, size_cur, k32, scur, vM32,
                           vden);
        lanes_max_pair<X>(best, bestK);
        const float tc = (float)kc / vM32;
#pragma unroll
        for (int t = 0; t < E; ++t)
            cand[t] = cand[t] && t * X + lane != hj &&
                      may_win(gain_bound(K[t], vM32, vden, k32, scur), c[t], best, a.obs != 0, tc);
    }
#pragma unroll
    for (int t = 0; t < E; ++t) rc[t] = cand[t] ? a.crec[c[t]] : 0ull;
#pragma unroll
    for (int t = 0; t < E; ++t)
        if (cand[t])
            consider_k(best, bestK, c[t], K[t], rc[t], cur, size_cur, k32, scur, vM32, vden);
    lanes_max_pair<X>(best, bestK);
    if (lane == 0) write_decision(a, v, cur, best, bestK, kc);
}

template <int X, int E>
__global__ __launch_bounds__(256) void k_move_sub(MoveArgs a, const int32_t *__restrict__ list,
                                                  int32_t n) {
    if (gated_off(a.gate)) return;
    const int lane = threadIdx.x & (X - 1);
    const int32_t slot = (int32_t)((blockIdx.x * 256u + threadIdx.x) / X);
    if (slot >= n) return;  // whole X-lane segment leaves together
    move_seg<X, E>(a, list[slot], lane);
}

// One wave's LDS hash table of C slots (keys K, FP64 sums W; C a multiple
// of 64): a power of two hashes by shift, other sizes by multiply-shift.
// L: lanes per row (64, or 32 for two rows per wave); lane < L.
template <int C, int L = 64>
__device__ __forceinline__ void wave_tab_clear(int32_t *K, double *W, int lane) {
#pragma unroll
    for (int j = 0; j < C / L; ++j) {
        K[lane + L * j] = -1;
        W[lane + L * j] = 0.0;
    }
}

// Inserts vertex v's row (MAXE entries per lane, deg <= L * MAXE) into the
// row's table; returns v's community.
template <int C, int MAXE, int L = 64>
__device__ __forceinline__ int32_t wave_tab_fill(const MoveArgs &a, int32_t v, int32_t *K,
                                                 double *W, int lane) {
    constexpr bool POW2 = (C & (C - 1)) == 0;
    constexpr int LOG2C = ilog2_c(C);
    const int32_t rs = a.rowptr[v], re = a.rowptr[v + 1];
    const int32_t cur = a.comm[v];
    const int32_t vs = col_self(a, v);
    int32_t u[MAXE];
    float wv[MAXE];
#pragma unroll
    for (int j = 0; j < MAXE; ++j) {
        const int32_t e = rs + lane + L * j;
        u[j] = e < re ? a.ccol[e] : -1;
        wv[j] = e < re ? a.w[e] : 0.0f;
    }
    int32_t cc[MAXE];
#pragma unroll
    for (int j = 0; j < MAXE; ++j) cc[j] = (u[j] >= 0 && u[j] != vs) ? a.ccomm[u[j]] : -1;
#pragma unroll
    for (int j = 0; j < MAXE; ++j) {
        if (cc[j] < 0) continue;
        if constexpr (POW2)
            lds_insert(K, W, 32 - LOG2C, C - 1, cc[j], (double)wv[j]);
        else
            lds_insert_n(K, W, (uint32_t)C, cc[j], (double)wv[j]);
    }
    return cur;
}

// Decision of v (in community cur) from its filled table; the L lanes of
// the row call it together.
template <int C, int L = 64>
__device__ __forceinline__ void wave_tab_decide(const MoveArgs &a, int32_t v, int32_t cur,
                                                const int32_t *K, const double *W, int lane) {
    constexpr int S = C / L;  // slots per lane
    const float k32 = (float)a.k[v];
    const float2 md = vertex_mden(a, v);
    const float vM32 = md.x, vden = md.y;
    const uint64_t rcur = a.crec[cur];
    const float scur = rec_sigma(rcur);
    const int32_t size_cur = rec_size(rcur);
    int32_t kk[S];
    double vv[S];
    uint64_t rr[S];
#pragma unroll
    for (int j = 0; j < S; ++j) {
        kk[j] = K[lane + L * j];
        vv[j] = W[lane + L * j];
    }
    double kc = 0.0;
#pragma unroll
    for (int j = 0; j < S; ++j)
        if (kk[j] == cur) kc = vv[j];
    kc = lanes_sum<L>(kc);
    uint64_t best = pack_best(0.0f, cur);
    double bestK = 0.0;
    if (a.bounds) {
        // the heaviest candidate first; the others are read only if their
        // bound can still change the decision
        double km = -1.0;
        int js = 0;
#pragma unroll
        for (int j = 0; j < S; ++j)
            if (kk[j] >= 0 && kk[j] != cur && vv[j] > km) {
                km = vv[j];
                js = j;
            }
        const int src = lanes_argmax<L>(km, lane);
        if (lane == src && km >= 0.0)
            consider_k(best, bestK, kk[js], vv[js], a.crec[kk[js]], cur, size_cur, k32, scur, vM32,
                       vden);
        lanes_max_pair<L>(best, bestK);  // src holds the best, the others the start key
        const float t = (float)kc / vM32;
        bool ev[S];
#pragma unroll
        for (int j = 0; j < S; ++j)
            ev[j] = kk[j] >= 0 && kk[j] != cur && !(lane == src && j == js) &&
                    may_win(gain_bound(vv[j], vM32, vden, k32, scur), kk[j], best, a.obs != 0, t);
#pragma unroll
        for (int j = 0; j < S; ++j) rr[j] = ev[j] ? a.crec[kk[j]] : 0ull;
#pragma unroll
        for (int j = 0; j < S; ++j)
            if (ev[j]) consider_k(best, bestK, kk[j], vv[j], rr[j], cur, size_cur, k32, scur, vM32, vden);
    } else {
#pragma unroll
        for (int j = 0; j < S; ++j) rr[j] = (kk[j] >= 0 && kk[j] != cur) ? a.crec[kk[j]] : 0ull;
#pragma unroll
        for (int j = 0; j < S; ++j) {
            if (kk[j] < 0 || kk[j] == cur) continue;
            consider_k(best, bestK, kk[j], vv[j], rr[j], cur, size_cur, k32, scur, vM32, vden);
        }
    }
    lanes_max_pair<L>(best, bestK);
    if (lane == 0) write_decision(a, v, cur, best, bestK, kc);
}

// C: LDS slots per row; G names the instance.  448 slots for degrees <= 319:
// load <= 0.71 and 7 blocks per CU.  Each lane holds MAXE row entries:
// 256-slot tables (instance G = 2, 4 per lane) take degrees up to 255, for a
// group-5 split above 128.
// L = 32: two rows per wave, one per half, so that a CU holds twice the
// rows in flight (instance G = 3: degrees <= 84 in 128-slot tables, G = 4:
// degrees <= 128 in 192-slot tables).
template <int C, int G, int L = 64>
__global__ __launch_bounds__(256) void k_move_wave(MoveArgs a, const int32_t *__restrict__ list,
                                                   int32_t n) {
    if (gated_off(a.gate)) return;
    static_assert(C % L == 0, "whole slots per lane");
    constexpr int R = 64 / L;  // rows per wave
    constexpr int MAXE = L == 32 ? (G == 3 ? 3 : 4) : (C == 256 ? 4 : 5);  // edges per lane
    static_assert(L * MAXE >= (L == 32 ? (G == 3 ? 84 : 128) : C == 256 ? 255 : 319),
                  "a row fits the lanes");
    __shared__ int32_t keys_s[4 * R * C];
    __shared__ double vals_s[4 * R * C];
    const int lane = threadIdx.x & (L - 1);
    const int seg = threadIdx.x / L;  // row of the block
    const int32_t slot = blockIdx.x * (4 * R) + seg;
    const bool active = slot < n;
    int32_t *K = keys_s + seg * C;
    double *W = vals_s + seg * C;
    wave_tab_clear<C, L>(K, W, lane);
    __syncthreads();
    int32_t v = 0, cur = 0;
    if (active) {
        v = list[slot];
        cur = wave_tab_fill<C, MAXE, L>(a, v, K, W, lane);
    }
    __syncthreads();
    if (!active) return;  // only communication inside the row's lanes below
    wave_tab_decide<C, L>(a, v, cur, K, W, lane);
}

// Small-table attempt for the rows of the one-workgroup tiers (degrees above
// the wave tiers, up to the chunked tier): one wave per row and a table of
// kHintSlots slots per wave, so that a CU holds many more rows in flight than
// degree-sized tables allow.  On aggregated levels a row's distinct
// communities are soon a small fraction of its degree.  Row i is tried when
// its last evaluation saw at most kHintLimit distinct communities (hint[i];
// INT32_MAX before its first); the lanes read the table's key count before
// every insert and the wave gives up once it reaches kHintLimit, so at most
// 63 keys land beyond it and the table never fills.  done[i] = 1 for a row
// decided here (k_move_wg skips it), 0 otherwise; hint[i] = its distinct
// count.  Decisions are the tier kernels' (the same exact sums, argmax and
// candidate bounds).
constexpr int kHintSlots = 512;
constexpr int kHintLimit = 256;
constexpr int kHintRound = 8;  // row entries per lane in flight together
constexpr int32_t kHintMaxHeavy = 16384;  // longest chunked-tier row tried
static_assert(kHintSlots > kHintLimit + 64, "the attempt's table never fills");
static_assert(kHintSlots % 256 == 0, "four slots per lane per step");

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The list may continue past n with n2 rows of the chunked tier (its heavy
// vertices in order), whose hints and flags are hint2/done2 (the tier's
// own): those rows are left alone unless decided here (done2 = 1 then, which
// k_heavy_final clears); rows longer than kHintMaxHeavy are not tried (one
// wave would hold the launch up).
__global__ __launch_bounds__(256, 6) void k_move_hinted(MoveArgs a, const int32_t *__restrict__ list,
                                                        int32_t n, int32_t *hint, int32_t *done,
                                                        int32_t n2, int32_t *hint2,
                                                        int32_t *done2) {
    if (gated_off(a.gate)) return;
    constexpr int C = kHintSlots;
    __shared__ int32_t keys_s[4 * C];
    __shared__ double vals_s[4 * C];
    __shared__ int32_t fill_s[4];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int32_t *K = keys_s + wid * C;
    double *W = vals_s + wid * C;
    volatile int32_t *fill = &fill_s[wid];
    for (int32_t i = blockIdx.x * 4 + wid; i < n + n2; i += gridDim.x * 4) {
        const bool heavy = i >= n;  // a row of the chunked tier
        int32_t *const hp = heavy ? hint2 + (i - n) : hint + i;
        int32_t *const dp = heavy ? done2 + (i - n) : done + i;
        if (*hp > kHintLimit) {  // the tier kernel takes it
            if (lane == 0 && !heavy) *dp = 0;
            continue;
        }
        const int32_t v = list[i];
        const int32_t rs = a.rowptr[v], re = a.rowptr[v + 1];
        if (heavy && re - rs > kHintMaxHeavy) continue;  // too long for one wave
        wave_tab_clear<C>(K, W, lane);
        if (lane == 0) *fill = 0;
        wave_sync();
        const int32_t cur = a.comm[v];
        const int32_t vs = col_self(a, v);
        double own = 0.0;  // the own community's weight, in registers
        bool dropped = false;
        for (int32_t e0 = rs + lane; e0 < re; e0 += kHintRound * 64) {
            int32_t u[kHintRound];
            float wv[kHintRound];
#pragma unroll
            for (int j = 0; j < kHintRound; ++j) {
                const int32_t e = e0 + j * 64;
                u[j] = e < re ? a.ccol[e] : -1;
                wv[j] = e < re ? a.w[e] : 0.0f;
            }
            int32_t cc[kHintRound];
#pragma unroll
            for (int j = 0; j < kHintRound; ++j)
                cc[j] = (u[j] >= 0 && u[j] != vs) ? a.ccomm[u[j]] : -1;
#pragma unroll
            for (int j = 0; j < kHintRound; ++j) {
                // every lane reads the count before any lane of this step
                // inserts: the step adds at most 64 keys
                const bool full = *fill >= kHintLimit;
                dropped |= full;
                if (full || cc[j] < 0) continue;
                if (cc[j] == cur) {
                    own += (double)wv[j];
                    continue;
                }
                if (lds_insert_claim(K, W, (uint32_t)C, cc[j], (double)wv[j]))
                    atomicAdd(&fill_s[wid], 1);
            }
            if (__any(dropped)) {
                dropped = true;
                break;
            }
        }
        if (dropped) {  // the whole wave: the tier kernel takes the row
            if (lane == 0 && !heavy) *dp = 0;
            wave_sync();
            continue;
        }
        const double kc = lanes_sum<64>(own);
        wave_sync();
        // decision as in k_move_wg, four slots per lane at a time
        const float k32 = (float)a.k[v];
        const float2 md = vertex_mden(a, v);
        const float vM32 = md.x, vden = md.y;
        const uint64_t rcur = a.crec[cur];
        const float scur = rec_sigma(rcur);
        const int32_t size_cur = rec_size(rcur);
        uint64_t best = pack_best(0.0f, cur);
        double bestK = 0.0;
        int32_t slot0 = -1;
        if (a.bounds) {
            // the heaviest candidate first; the others are read only if
            // their bound can still change the decision
            double km = -1.0;
            int32_t ks = -1;
            for (int s2 = lane; s2 < C; s2 += 64)
                if (K[s2] >= 0 && W[s2] > km) {
                    km = W[s2];
                    ks = s2;
                }
            const int src = lanes_argmax<64>(km, lane);
            if (lane == src && ks >= 0)
                consider_k(best, bestK, K[ks], W[ks], a.crec[K[ks]], cur, size_cur, k32, scur,
                           vM32, vden);
            best = __shfl(best, src, 64);
            bestK = __shfl(bestK, src, 64);
            slot0 = __shfl(ks, src, 64);
        }
        const float tc = (float)kc / vM32;
        for (int i0 = lane; i0 < C; i0 += 4 * 64) {
            int32_t kk[4];
            double vv[4];
            bool ev[4];
            uint64_t rr[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int s2 = i0 + j * 64;
                kk[j] = K[s2];
                vv[j] = W[s2];
                ev[j] = kk[j] >= 0 && s2 != slot0 &&
                        (!a.bounds || may_win(gain_bound(vv[j], vM32, vden, k32, scur), kk[j],
                                              best, a.obs != 0, tc));
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) rr[j] = ev[j] ? a.crec[kk[j]] : 0ull;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (ev[j])
                    consider_k(best, bestK, kk[j], vv[j], rr[j], cur, size_cur, k32, scur, vM32,
                               vden);
        }
        lanes_max_pair<64>(best, bestK);
        if (lane == 0) {
            write_decision(a, v, cur, best, bestK, kc);
            *dp = 1;
            *hp = *fill;
        }
        wave_sync();  // every lane is done with the table before the next row clears it
    }
}

// TIER only names the instance (one per workgroup bucket) for profiles.
// Every instance keeps to 64 VGPRs (8 waves per SIMD): with tables of 1.5 x
// the degree, LDS admits that many (two 1024-lane workgroups per CU).
//
// TIER 4 tries the rows of the chunked tier (list = its heavy vertices, in
// their order) in one table of the whole allocation.  Row vi is tried only
// when its last evaluation saw at most kHeavyTryLimit distinct communities
// (hint[vi]).  Each lane reads the table's key count before every insert and
// gives up at `limit` (<= kHeavyTryLimit), so at most T keys land beyond it:
// with cap > kHeavyTryLimit + T the table never fills and every probe ends.
// A row whose inserts all landed is decided here (done[vi] = 1, hint[vi] =
// its count); the chunked tier takes the others.  Rows outside a pruned
// sweep's frontier are skipped, as the chunked tier skips them.
// HINTS (tiers 0-3): rows decided by the small-table attempt (done) are
// skipped and every row evaluated records its distinct communities (hint);
// the instance without them is the plain tier.
template <int T, int TIER, bool HINTS = false>
__global__ __launch_bounds__(T, 8) void k_move_wg(MoveArgs a, const int32_t *__restrict__ list,
                                                  int32_t n, int cap, int pct, int32_t *hint,
                                                  int32_t *done, int32_t limit) {
    if (gated_off(a.gate)) return;
    extern __shared__ double dyn_lds[];
    const int C = cap;  // allocation (slots): W[C] then K[C]
    double *W = dyn_lds;
    int32_t *K = reinterpret_cast<int32_t *>(dyn_lds + C);
    __shared__ uint64_t red_best[T / 64];
    __shared__ double red_K[T / 64];
    __shared__ double red_kc[T / 64];
    __shared__ int32_t red_n[T / 64];   // distinct keys per wave (hints)
    __shared__ int32_t fill_s, stop_s;  // TIER 4: keys in the table, an insert was dropped
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int32_t vi = blockIdx.x; vi < n; vi += gridDim.x) {
        if constexpr (TIER == 4) {
            // the chunked tier takes it, or the small-table attempt decided it
            if (hint[vi] > kHeavyTryLimit || done[vi]) continue;
        } else if constexpr (HINTS) {
            // decided by the small-table attempt (k_move_hinted) this sweep
            if (done[vi]) continue;
        }
        const int32_t v = list[vi];
        if constexpr (TIER == 4) {
            if (a.active && !a.active[v]) continue;  // not in this sweep's frontier
        }
        const int32_t rs = a.rowptr[v], re = a.rowptr[v + 1];
        // table of the vertex: pct % of its degree (150: load <= 2/3), in
        // 64-slot steps, at least 512 slots and at most the allocation (which
        // holds more slots than the tier's degree bound)
        const int32_t deg = re - rs;
        const int Cv =
            TIER == 4 ? C : min(C, max(512, (int)(((int64_t)deg * pct / 100 + 63) & ~(int64_t)63)));
        for (int i = threadIdx.x; i < Cv; i += T) {
            K[i] = -1;
            W[i] = 0.0;
        }
        if constexpr (TIER == 4) {
            if (threadIdx.x == 0) {
                fill_s = 0;
                stop_s = 0;
            }
        }
        __syncthreads();
        const int32_t cur = a.comm[v];
        const int32_t vs = col_self(a, v);
        // TIER 4 (rows of settled levels): entries of v's own community are
        // summed in registers (the key most entries of such a row name: in
        // the table its slot would serialise the lanes' atomics)
        double own = 0.0;
        if constexpr (TIER == 4) {
            volatile int32_t *fill = &fill_s, *stop = &stop_s;
            bool dropped = false;
            for (int32_t e0 = rs + threadIdx.x; e0 < re && !dropped; e0 += 4 * T) {
                int32_t u[4];
                float wv[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int32_t e = e0 + j * T;
                    u[j] = e < re ? a.ccol[e] : -1;
                    wv[j] = e < re ? a.w[e] : 0.0f;
                }
                int32_t cc[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) cc[j] = (u[j] >= 0 && u[j] != vs) ? a.ccomm[u[j]] : -1;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (cc[j] < 0 || dropped) continue;
                    if (cc[j] == cur) {
                        own += (double)wv[j];
                        continue;
                    }
                    if (*fill >= limit) {
                        dropped = true;
                        *stop = 1;
                    } else if (lds_insert_claim(K, W, (uint32_t)Cv, cc[j], (double)wv[j])) {
                        atomicAdd(&fill_s, 1);
                    }
                }
                if (*stop) dropped = true;  // another lane gave up: so does this one
            }
            own = lanes_sum<64>(own);
            if (lane == 0) red_kc[wid] = own;
            __syncthreads();
            const bool gave_up = stop_s != 0;
            if (gave_up) {
                __syncthreads();  // every lane has read stop_s before the next row resets it
                continue;
            }
        } else {
            for (int32_t e0 = rs + threadIdx.x; e0 < re; e0 += 4 * T) {
                int32_t u[4];
                float wv[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int32_t e = e0 + j * T;
                    u[j] = e < re ? a.ccol[e] : -1;
                    wv[j] = e < re ? a.w[e] : 0.0f;
                }
                int32_t cc[4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    cc[j] = (u[j] >= 0 && u[j] != vs) ? a.ccomm[u[j]] : -1;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (cc[j] >= 0) lds_insert_n(K, W, (uint32_t)Cv, cc[j], (double)wv[j]);
            }
            __syncthreads();
        }
        // weight to the own community: TIER 4 summed it over the waves (in a
        // fixed order); the others read it from the table below
        double kc = 0.0;
        if constexpr (TIER == 4) {
#pragma unroll
            for (int w2 = 0; w2 < T / 64; ++w2) kc += red_kc[w2];
        }
        const float k32 = (float)a.k[v];
        const float2 md = vertex_mden(a, v);
        const float vM32 = md.x, vden = md.y;
        const uint64_t rcur = a.crec[cur];
        const float scur = rec_sigma(rcur);
        const int32_t size_cur = rec_size(rcur);
        if (a.bounds) {
            // Each wave evaluates the heaviest candidate among its own slots
            // first; its other slots are read only if their bound can still
            // change the decision.
            if constexpr (TIER != 4) kc = lds_value_n(K, W, (uint32_t)Cv, cur);
            double km = -1.0;
            int32_t ks = -1, nd = 0;
            for (int i = threadIdx.x; i < Cv; i += T) {
                const int32_t kx = K[i];
                if constexpr (HINTS) nd += kx >= 0 ? 1 : 0;
                if (kx >= 0 && kx != cur && W[i] > km) {
                    km = W[i];
                    ks = i;
                }
            }
            const int src = lanes_argmax<64>(km, lane);
            uint64_t best = pack_best(0.0f, cur);
            double bestK = 0.0;
            if (lane == src && ks >= 0)
                consider_k(best, bestK, K[ks], W[ks], a.crec[K[ks]], cur, size_cur, k32, scur, vM32,
                           vden);
            best = __shfl(best, src, 64);
            bestK = __shfl(bestK, src, 64);
            const int32_t slot0 = __shfl(ks, src, 64);
            const float t = (float)kc / vM32;
            for (int i0 = threadIdx.x; i0 < Cv; i0 += 4 * T) {
                int32_t kk[4];
                double vv[4];
                bool ev[4];
                uint64_t rr[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int i = i0 + j * T;
                    kk[j] = i < Cv ? K[i] : -1;
                    vv[j] = i < Cv ? W[i] : 0.0;
                    ev[j] = kk[j] >= 0 && kk[j] != cur && i != slot0 &&
                            may_win(gain_bound(vv[j], vM32, vden, k32, scur), kk[j], best,
                                    a.obs != 0, t);
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) rr[j] = ev[j] ? a.crec[kk[j]] : 0ull;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (ev[j])
                        consider_k(best, bestK, kk[j], vv[j], rr[j], cur, size_cur, k32, scur,
                                   vM32, vden);
            }
            lanes_max_pair<64>(best, bestK);
            if constexpr (HINTS) nd = wave_sum_i32(nd);
            if (lane == 0) {
                red_best[wid] = best;
                red_K[wid] = bestK;
                if constexpr (HINTS) red_n[wid] = nd;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                uint64_t b = red_best[0];
                double bK = red_K[0];
                for (int w2 = 1; w2 < T / 64; ++w2)
                    if (red_best[w2] > b) {
                        b = red_best[w2];
                        bK = red_K[w2];
                    }
                write_decision(a, v, cur, b, bK, kc);
                if constexpr (TIER == 4) {
                    done[vi] = 1;
                    hint[vi] = fill_s;
                } else if constexpr (HINTS) {
                    int32_t c = 0;
                    for (int w2 = 0; w2 < T / 64; ++w2) c += red_n[w2];
                    hint[vi] = c;
                }
            }
            __syncthreads();
            continue;
        }
        uint64_t best = pack_best(0.0f, cur);
        double bestK = 0.0, kcs = 0.0;  // kcs: the own community's slot (TIER < 4)
        int32_t nd = 0;
        for (int i0 = threadIdx.x; i0 < Cv; i0 += 4 * T) {
            int32_t kk[4];
            double vv[4];
            uint64_t rr[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int i = i0 + j * T;
                kk[j] = i < Cv ? K[i] : -1;
                vv[j] = i < Cv ? W[i] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) rr[j] = (kk[j] >= 0 && kk[j] != cur) ? a.crec[kk[j]] : 0ull;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (kk[j] < 0) continue;
                if constexpr (HINTS) ++nd;
                if (kk[j] == cur) {  // TIER < 4 only: TIER 4 keeps it out of the table
                    kcs = vv[j];
                    continue;
                }
                consider_k(best, bestK, kk[j], vv[j], rr[j], cur, size_cur, k32, scur, vM32,
                           vden);
            }
        }
        lanes_max_pair<64>(best, bestK);
        if constexpr (HINTS) nd = wave_sum_i32(nd);
        if constexpr (TIER != 4) kcs = lanes_sum<64>(kcs);
        if (lane == 0) {
            red_best[wid] = best;
            red_K[wid] = bestK;
            if constexpr (HINTS) red_n[wid] = nd;
            if constexpr (TIER != 4) red_kc[wid] = kcs;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t b = red_best[0];
            double bK = red_K[0];
            for (int w2 = 1; w2 < T / 64; ++w2) {
                if (red_best[w2] > b) {
                    b = red_best[w2];
                    bK = red_K[w2];
                }
            }
            if constexpr (TIER != 4) {
                kc = 0.0;
                for (int w2 = 0; w2 < T / 64; ++w2) kc += red_kc[w2];
            }
            write_decision(a, v, cur, b, bK, kc);
            if constexpr (TIER == 4) {
                done[vi] = 1;
                hint[vi] = fill_s;
            } else if constexpr (HINTS) {
                int32_t c = 0;
                for (int w2 = 0; w2 < T / 64; ++w2) c += red_n[w2];
                hint[vi] = c;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------- chunked tier

// Stage A: aggregate one chunk of a heavy row in LDS; write its distinct
// (community, weight) pairs grouped by partition.
// Partition of community c among the P merge items of a heavy vertex: a hash
// (part_of), or with range_blk > 0 id ranges, P/8 per column block of
// range_blk ids, so that the merge items of block x can run on XCD x.
__device__ __forceinline__ uint32_t tier_part(const ChunkedTier &t, int32_t c, uint32_t P) {
    if (t.range_blk > 0) {
        const uint32_t blk = (uint32_t)t.range_blk;
        const uint32_t m = P >> 3;
        const uint32_t x = (uint32_t)c / blk;
        const uint32_t r = (uint32_t)c - x * blk;
        return x * m + (uint32_t)(((uint64_t)r * m) / blk);
    }
    return part_of(c, P);
}

// Stage A merges a chunk's duplicate communities in LDS unless the row's
// last evaluation saw at least kDirectNum / kDirectDen distinct communities
// per entry.
constexpr int64_t kDirectNum = 6, kDirectDen = 10;
static_assert(kChunkEdges <= 4 * 256, "a direct chunk holds four entries per lane");

// smallest power of two >= 2n (load <= 1/2), between 2^6 and 2^cap_lg
__device__ __forceinline__ int table_log2(int64_t n, int cap_lg) {
    const int lg = n > 1 ? 64 - __clzll((unsigned long long)(2 * n - 1)) : 1;
    return min(cap_lg, max(6, lg));
}

// Adds a chunk's own-community sum to its vertex's K_cur (one FP64 atomic
// per chunk); every lane calls it.
__device__ __forceinline__ void chunk_own_sum(const ChunkedTier &t, const Chunk &ch, double own) {
    __shared__ double own_s[4];
    own = lanes_sum<64>(own);
    if ((threadIdx.x & 63) == 0) own_s[threadIdx.x >> 6] = own;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double x = own_s[0] + own_s[1] + own_s[2] + own_s[3];
        if (x != 0.0) atomicAdd(&t.kcur[ch.hv], x);
    }
}

// From the chunk's pair counts per partition (pcnt): its partition offsets
// (poff), the vertex's partition totals (ptot), and pcnt turned into each
// partition's first output slot.  Every lane calls it.
__device__ __forceinline__ void chunk_part_offsets(const ChunkedTier &t, const Chunk &ch, int P,
                                                   int32_t *pcnt) {
    __syncthreads();
    if (threadIdx.x < 64) {  // wave 0: exclusive scan of the partition counts
        const int lane = threadIdx.x;
        int32_t carry = 0;
        for (int b = 0; b < P; b += 64) {
            const int p = b + lane;
            const int32_t x = p < P ? pcnt[p] : 0;
            int32_t inc = x;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int32_t y = __shfl_up(inc, off, 64);
                if (lane >= off) inc += y;
            }
            if (p < P) {
                t.poff[ch.pbase + p] = carry + inc - x;
                if (x > 0) atomicAdd(&t.ptot[ch.ptot_base + p], x);
                pcnt[p] = carry + inc - x;  // becomes the scatter cursor
            }
            carry += __shfl(inc, 63, 64);
        }
        if (lane == 0) t.poff[ch.pbase + P] = carry;
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_chunk_aggregate(MoveArgs a, ChunkedTier t) {
    if (gated_off(a.gate)) return;
    // dynamic LDS sized by the chunk size: vals[C], keys[C], pcnt[kMaxParts]
    extern __shared__ double agg_lds[];
    const int C = 1 << t.chunk_cap_log2;
    double *vals = agg_lds;
    int32_t *keys = reinterpret_cast<int32_t *>(agg_lds + C);
    int32_t *pcnt = keys + C;
    const int32_t ci = t.chunk_order ? t.chunk_order[blockIdx.x] : (int32_t)blockIdx.x;
    if (ci < 0) return;  // padding of the XCD dispatch order
    const Chunk ch = t.chunks[ci];
    if (a.active && !a.active[ch.v]) return;  // not in this sweep's frontier
    if (t.done[ch.hv]) return;                // decided in LDS
    const int P = ch.P;
    const int32_t vs = col_self(a, ch.v);
    const int32_t cur = a.comm[ch.v];
    // entries of the vertex's own community are summed in registers and
    // added to kcur (one FP64 atomic per chunk), the others go to the table
    double own = 0.0;
    for (int p = threadIdx.x; p < P; p += 256) pcnt[p] = 0;
    // On input levels (t.direct), a row whose last evaluation saw distinct
    // communities for most of its entries (every row before its first)
    // gains little from merging a chunk's duplicates here: its entries go to
    // their partitions as they are, and stage B merges them all the same.
    // A chunk holds at most kChunkEdges = 4 x 256 entries: four per lane.
    const int64_t rdeg = a.rowptr[ch.v + 1] - a.rowptr[ch.v];
    if (t.direct && (int64_t)t.hint[ch.hv] * kDirectDen >= rdeg * kDirectNum) {  // block-uniform
        int32_t cc[4], pp[4], pos[4];
        float wf[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t e = ch.lo + threadIdx.x + j * 256;
            const int32_t u = e < ch.hi ? __builtin_nontemporal_load(&a.ccol[e]) : -1;
            wf[j] = e < ch.hi ? __builtin_nontemporal_load(&a.w[e]) : 0.0f;
            cc[j] = (u >= 0 && u != vs) ? a.ccomm[u] : -1;
        }
        __syncthreads();  // pcnt cleared
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            pp[j] = -1;
            pos[j] = 0;
            if (cc[j] == cur) {
                own += (double)wf[j];
            } else if (cc[j] >= 0) {
                pp[j] = (int32_t)tier_part(t, cc[j], (uint32_t)P);
                pos[j] = atomicAdd(&pcnt[pp[j]], 1);
            }
        }
        chunk_own_sum(t, ch, own);  // (its barrier also completes the counts)
        chunk_part_offsets(t, ch, P, pcnt);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (pp[j] < 0) continue;
            const int64_t o = ch.out + pcnt[pp[j]] + pos[j];
            t.out_key[o] = cc[j];
            t.out_val[o] = (double)wf[j];
        }
        return;
    }
    const int cl = table_log2(ch.hi - ch.lo, t.chunk_cap_log2);
    const int Cv = 1 << cl;
    for (int i = threadIdx.x; i < Cv; i += 256) {
        keys[i] = -1;
        vals[i] = 0.0;
    }
    __syncthreads();
    // four entries per lane at a time: their loads, then their gathers, are
    // in flight together ahead of the LDS inserts
    for (int64_t e0 = ch.lo + threadIdx.x; e0 < ch.hi; e0 += 4 * 256) {
        int32_t u[4];
        float wf[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t e = e0 + j * 256;
            u[j] = e < ch.hi ? __builtin_nontemporal_load(&a.ccol[e]) : -1;
            wf[j] = e < ch.hi ? __builtin_nontemporal_load(&a.w[e]) : 0.0f;
        }
        int32_t cc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) cc[j] = (u[j] >= 0 && u[j] != vs) ? a.ccomm[u[j]] : -1;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (cc[j] == cur)
                own += (double)wf[j];
            else if (cc[j] >= 0)
                lds_insert(keys, vals, 32 - cl, (uint32_t)Cv - 1, cc[j], (double)wf[j]);
        }
    }
    chunk_own_sum(t, ch, own);
    for (int s = threadIdx.x; s < Cv; s += 256) {
        const int32_t k = keys[s];
        if (k >= 0) atomicAdd(&pcnt[tier_part(t, k, (uint32_t)P)], 1);
    }
    chunk_part_offsets(t, ch, P, pcnt);
    for (int s = threadIdx.x; s < Cv; s += 256) {
        const int32_t k = keys[s];
        if (k < 0) continue;
        const int32_t pos = atomicAdd(&pcnt[tier_part(t, k, (uint32_t)P)], 1);
        t.out_key[ch.out + pos] = k;
        t.out_val[ch.out + pos] = vals[s];
    }
}

// Stage B: merge partition p of a heavy vertex and evaluate its candidates.
// The partition's segments (one per chunk) are staged in LDS with an
// exclusive scan, so lanes walk one flat range whatever the chunk count.
// A partition holding more pairs than the LDS table takes falls back to the
// vertex's global table (always large enough), so every insert terminates.
constexpr int kMergeThreads = 256;

// Visits the pairs of partition p of vertex h: f(key, value).
template <typename F>
__device__ __forceinline__ void for_partition_pairs(const ChunkedTier &t, const HeavyVertex &h,
                                                    int p, int64_t *seg_base, int32_t *seg_off,
                                                    F &&f) {
    __shared__ int32_t wsum[kMergeThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int32_t ct0 = 0; ct0 < h.nchunks; ct0 += kMergeTile) {
        const int32_t nct = h.nchunks - ct0 < kMergeTile ? h.nchunks - ct0 : kMergeTile;
        // stage segment bases and lengths; block exclusive scan of lengths
        int32_t carry = 0;
        for (int32_t b0 = 0; b0 < nct; b0 += kMergeThreads) {
            const int32_t i = b0 + threadIdx.x;
            int32_t len = 0;
            if (i < nct) {
                const Chunk ch = t.chunks[h.chunk0 + ct0 + i];
                const int32_t lo = t.poff[ch.pbase + p];
                len = t.poff[ch.pbase + p + 1] - lo;
                seg_base[i] = ch.out + lo;
            }
            int32_t inc = len;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int32_t y = __shfl_up(inc, off, 64);
                if (lane >= off) inc += y;
            }
            if (lane == 63) wsum[wid] = inc;
            __syncthreads();
            int32_t woff = 0, tot = 0;
#pragma unroll
            for (int w2 = 0; w2 < kMergeThreads / 64; ++w2) {
                woff += w2 < wid ? wsum[w2] : 0;
                tot += wsum[w2];
            }
            if (i < nct) seg_off[i] = carry + woff + inc - len;
            carry += tot;
            __syncthreads();
        }
        if (threadIdx.x == 0) seg_off[nct] = carry;
        __syncthreads();
        const int32_t total = carry;
        // four pairs per lane at a time: their loads are in flight together
        for (int32_t f0 = threadIdx.x; f0 < total; f0 += 4 * kMergeThreads) {
            int32_t kk[4];
            double vv[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int32_t fj = f0 + j * kMergeThreads;
                kk[j] = -1;
                vv[j] = 0.0;
                if (fj < total) {
                    int32_t a = 0, b = nct - 1;  // last segment with seg_off <= fj
                    while (a < b) {
                        const int32_t mid = (a + b + 1) >> 1;
                        if (seg_off[mid] <= fj) a = mid;
                        else b = mid - 1;
                    }
                    const int64_t idx = seg_base[a] + (fj - seg_off[a]);
                    kk[j] = t.out_key[idx];
                    vv[j] = t.out_val[idx];
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (kk[j] >= 0) f(kk[j], vv[j]);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kMergeThreads) void k_chunk_merge(MoveArgs a, ChunkedTier t) {
    if (gated_off(a.gate)) return;
    constexpr int C = 1 << kMergeCapLog2;
    __shared__ int32_t keys[C];
    __shared__ double vals[C];
    __shared__ int64_t seg_base[kMergeTile];
    __shared__ int32_t seg_off[kMergeTile + 1];
    __shared__ uint64_t red_best[kMergeThreads / 64];
    __shared__ double red_K[kMergeThreads / 64];
    __shared__ int32_t red_n[kMergeThreads / 64];
    const int32_t item = t.item_order ? t.item_order[blockIdx.x] : (int32_t)blockIdx.x;
    if (item < 0) return;  // padding of the XCD dispatch order
    const MergeItem it = t.items[item];
    const HeavyVertex h = t.hvs[it.hv];
    const int32_t v = h.v;
    if (a.active && !a.active[v]) return;  // not in this sweep's frontier
    if (t.done[it.hv]) return;             // decided in LDS
    const int32_t cur = a.comm[v];
    const float k32 = (float)a.k[v];
    const float2 md = vertex_mden(a, v);
    const float vM32 = md.x, vden = md.y;
    const uint64_t rcur = a.crec[cur];
    const float scur = rec_sigma(rcur);
    const int32_t size_cur = rec_size(rcur);
    const int32_t n = t.ptot[h.ptot_base + it.p];
    uint64_t best = 0;
    double bestK = 0.0;
    int32_t nd = 0;  // distinct communities of the partition
    if (n <= t.merge_limit) {
        const int cl = table_log2(n, kMergeCapLog2);
        const int Cv = 1 << cl;
        for (int i = threadIdx.x; i < Cv; i += kMergeThreads) {
            keys[i] = -1;
            vals[i] = 0.0;
        }
        __syncthreads();
        for_partition_pairs(t, h, it.p, seg_base, seg_off, [&](int32_t c, double x) {
            lds_insert(keys, vals, 32 - cl, (uint32_t)Cv - 1, c, x);
        });
        for (int i0 = threadIdx.x; i0 < Cv; i0 += 4 * kMergeThreads) {
            int32_t kk[4];
            double vv[4];
            uint64_t rr[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int i = i0 + j * kMergeThreads;
                kk[j] = i < Cv ? keys[i] : -1;
                vv[j] = i < Cv ? vals[i] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) rr[j] = kk[j] >= 0 ? a.crec[kk[j]] : 0ull;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (kk[j] < 0) continue;
                ++nd;
                consider_k(best, bestK, kk[j], vv[j], rr[j], cur, size_cur, k32, scur, vM32,
                           vden);
            }
        }
    } else {
        int32_t *K = t.gkeys + h.tab_off;
        double *W = t.gvals + h.tab_off;
        if (threadIdx.x == 0) t.ovf[it.hv] = 1;
        const uint32_t cap = 1u << h.tab_lg;
        for_partition_pairs(t, h, it.p, seg_base, seg_off, [&](int32_t c, double x) {
            global_insert(K, W, 32 - h.tab_lg, cap - 1, c, x);
        });
        __threadfence();
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < cap; i += kMergeThreads) {
            const int32_t c = __hip_atomic_load(&K[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (c < 0 || tier_part(t, c, (uint32_t)h.P) != (uint32_t)it.p) continue;
            ++nd;
            const double Kc = __hip_atomic_load(&W[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // the table stays as it is: other partitions of the vertex may
            // still be inserting; k_clear_overflow empties it afterwards
            consider_k(best, bestK, c, Kc, a.crec[c], cur, size_cur, k32, scur, vM32, vden);
        }
    }
    lanes_max_pair<64>(best, bestK);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) nd += __shfl_xor(nd, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red_best[wid] = best;
        red_K[wid] = bestK;
        red_n[wid] = nd;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t b = red_best[0];
        double bK = red_K[0];
        int32_t nn = red_n[0];
        for (int w2 = 1; w2 < kMergeThreads / 64; ++w2) {
            if (red_best[w2] > b) {
                b = red_best[w2];
                bK = red_K[w2];
            }
            nn += red_n[w2];
        }
        atomicAdd(&t.hdist[it.hv], nn);
        t.item_best[item] = b;
        t.item_K[item] = bK;
        if (b != 0)
            atomicMax(reinterpret_cast<unsigned long long *>(&t.best[it.hv]),
                      (unsigned long long)b);
        t.ptot[h.ptot_base + it.p] = 0;  // ready for the next sweep
    }
}

// One wave per heavy vertex: its lanes look for the merge item that holds
// the vertex's best candidate (a community lies in one partition, so one
// item matches), lane 0 writes the decision.
__global__ __launch_bounds__(256) void k_heavy_final(MoveArgs a, int32_t nh, ChunkedTier t) {
    if (gated_off(a.gate)) return;
    const int lane = threadIdx.x & 63;
    const int32_t hv = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (hv >= nh) return;  // the whole wave
    if (t.done[hv]) {      // decided in LDS
        if (lane == 0) t.done[hv] = 0;
        return;
    }
    const HeavyVertex h = t.hvs[hv];
    const int32_t v = h.v;
    const int32_t cur = a.comm[v];
    if (a.active && !a.active[v]) {  // not evaluated: stays
        if (lane == 0) {
            a.newc[v] = cur;
            a.dk[v] = 0.0;
        }
        return;
    }
    uint64_t best = t.best[hv];
    double bestK = 0.0;
    if (best != 0) {
        for (int32_t p0 = 0; p0 < h.P; p0 += 64) {
            const int32_t p = p0 + lane;
            const uint64_t m = __ballot(p < h.P && t.item_best[h.ptot_base + p] == best);
            if (m) {
                bestK = t.item_K[h.ptot_base + p0 + (__ffsll((unsigned long long)m) - 1)];
                break;
            }
        }
    }
    if (lane != 0) return;
    const uint64_t init = pack_best(0.0f, cur);
    if (init > best) best = init;
    write_decision(a, v, cur, best, bestK, t.kcur[hv]);
    t.best[hv] = 0;
    t.kcur[hv] = 0.0;
    // distinct communities of this evaluation: the next sweep tries the row
    // in LDS when they fit
    t.hint[hv] = t.hdist[hv];
    t.hdist[hv] = 0;
}

// Empties the overflow tables the merge items used (one workgroup per heavy
// vertex), restoring the all-empty arena for the next sweep.
__global__ void k_clear_overflow(ChunkedTier t, int32_t nh) {
    const int32_t hv = blockIdx.x;
    if (hv >= nh || !t.ovf[hv]) return;
    const HeavyVertex h = t.hvs[hv];
    const int64_t cap = (int64_t)1 << h.tab_lg;
    for (int64_t i = threadIdx.x; i < cap; i += blockDim.x) {
        t.gkeys[h.tab_off + i] = -1;
        t.gvals[h.tab_off + i] = 0.0;
    }
    if (threadIdx.x == 0) t.ovf[hv] = 0;
}

__global__ void k_clear_tables(int32_t *keys, double *vals, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        keys[i] = -1;
        vals[i] = 0.0;
    }
}

// ------------------------------------------------------------------ apply
// updateVertexCommunity (modularity_optimisation.cu:237-249) with the
// community totals updated by exact FP64 increments (the reference rebuilds
// them, :328-330).  Many movers of a wave often join the same (hub)
// community: lanes sharing the wave leader's key are combined into one
// atomic per field, a few times, before the rest go one by one (joins and
// leaves alike).  A second launch refreshes the records of the touched
// communities and, when
// the sweep tracks modularity incrementally, adds the group's change of the
// weight inside communities:
//   2*dk[v] per mover v (its own view, neighbours fixed), plus for each
//   entry (v,u) with u also a mover of the group
//   w * ([c'u=c'v] - [cu=c'v] - [cv=c'u] + [cu=cv])   (seen from both ends).

__device__ __forceinline__ void wave_peel_update(bool todo, int32_t key, double x, int32_t dn,
                                                 int32_t *size, double *sigma, int peels) {
    const int lane = threadIdx.x & 63;
    for (int it = 0; it < peels; ++it) {
        const uint64_t pend = __ballot(todo);
        if (pend == 0) return;
        const int leader = __ffsll((long long)pend) - 1;
        const int32_t lk = __shfl(key, leader, 64);
        const bool mine = todo && key == lk;
        const uint64_t m = __ballot(mine);
        const double sx = wave_sum_f64(mine ? x : 0.0);
        if (lane == leader) {
            atomicAdd(&sigma[lk], sx);
            atomicAdd(&size[lk], dn * __popcll(m));
        }
        todo = todo && !mine;
    }
    if (todo) {
        atomicAdd(&sigma[key], x);
        atomicAdd(&size[key], dn);
    }
}

// Sets bit `idx`; lanes of the wave sharing the leader's word are combined
// into one atomic OR.
__device__ __forceinline__ void wave_set_bit(bool todo, int32_t idx, uint32_t *bits) {
    const int lane = threadIdx.x & 63;
    const uint64_t pend = __ballot(todo);
    if (pend == 0) return;
    const int leader = __ffsll((long long)pend) - 1;
    const int32_t lw = __shfl(idx >> 5, leader, 64);
    const bool mine = todo && (idx >> 5) == lw;
    uint32_t m = mine ? (1u << (idx & 31)) : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m |= __shfl_xor(m, off, 64);
    if (lane == leader) atomicOr(&bits[lw], m);
    if (todo && !mine) atomicOr(&bits[idx >> 5], 1u << (idx & 31));
}

__device__ __forceinline__ bool test_bit(const uint32_t *bits, int32_t idx) {
    return (bits[idx >> 5] >> (idx & 31)) & 1u;
}

// Movers join their new community and leave the old one: sizes and totals
// by atomics (the lanes of a wave that share the leading lane's community
// are combined first), the move history, the mover bitmap and counts.
__global__ __launch_bounds__(256) void k_apply_join(const int32_t *__restrict__ list, int32_t n,
                                                    const int32_t *__restrict__ rowptr,
                                                    int32_t *newc, int32_t *comm, int32_t *size,
                                                    double *sigma, const double *__restrict__ k,
                                                    uint32_t *movers, uint8_t *hist,
                                                    unsigned long long *moved,
                                                    const int32_t *__restrict__ gate,
                                                    int32_t *ccomm,
                                                    const int32_t *__restrict__ cself) {
    if (gated_off(gate)) return;
    __shared__ unsigned long long blk_moved[2];
    if (threadIdx.x < 2) blk_moved[threadIdx.x] = 0ull;
    __syncthreads();
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    bool mv = false;
    int32_t v = 0, nc = 0, oc = 0;
    double kv = 0.0;
    if (i < n) {
        v = list[i];
        nc = newc[v];
        oc = comm[v];
        if (nc != oc) {
            mv = true;
            kv = k[v];
            comm[v] = nc;
            if (cself) ccomm[cself[v]] = nc;  // the move kernels' copy of comm
            newc[v] = oc;  // from here on newc holds the community left
            hist[v] |= 1;  // move history (bit 0 = this sweep)
        }
    }
    wave_peel_update(mv, nc, kv, 1, size, sigma, 3);
    wave_peel_update(mv, oc, -kv, -1, size, sigma, 1);
    if (movers) wave_set_bit(mv, v, movers);
    const uint64_t m = __ballot(mv);
    // moved[0] += movers, moved[1] += their row entries: summed over the
    // block in LDS, then one atomic pair per block, spread over kMoveSlots
    // cache lines (counters shared by many blocks serialise their atomics)
    const double deg = mv ? (double)(rowptr[v + 1] - rowptr[v]) : 0.0;
    const double wdeg = wave_sum_f64(deg);
    if ((threadIdx.x & 63) == 0 && m) {
        atomicAdd(&blk_moved[0], (unsigned long long)__popcll(m));
        atomicAdd(&blk_moved[1], (unsigned long long)wdeg);
    }
    __syncthreads();
    if (threadIdx.x == 0 && blk_moved[0] != 0) {
        unsigned long long *slot = moved + (int64_t)(blockIdx.x % kMoveSlots) * kMoveSlotStride;
        atomicAdd(slot, blk_moved[0]);
        atomicAdd(slot + 1, blk_moved[1]);
    }
}

constexpr int kClearWords = 4096;

// Blocks [0, fix_start): modularity correction over the group's movers'
// rows (incremental sweeps only); [fix_start, clear_start): records of the
// touched communities; [clear_start, ...): clear the other step's bitmap.
__global__ __launch_bounds__(kRowThreads) void k_apply_finish(
    MoveArgs a, const int32_t *__restrict__ order, const int32_t *__restrict__ list, int32_t n,
    SegTable st,
    const Chunk *__restrict__ chunks, const int32_t *__restrict__ comm,
    const int32_t *__restrict__ size, const double *__restrict__ sigma, uint64_t *crec,
    const uint32_t *__restrict__ movers, uint32_t *clear_bits, int32_t words, double *own,
    int32_t fix_start, int32_t clear_start, int32_t rounds, const int32_t *__restrict__ gid,
    double *gin) {
    if (gated_off(a.gate)) return;
    __shared__ double lds[kRowThreads / 64];
    const int b = blockIdx.x;
    if (b >= clear_start) {
        const int64_t w0 = (int64_t)(b - clear_start) * kClearWords;
        for (int64_t x = w0 + threadIdx.x; x < w0 + kClearWords && x < words; x += kRowThreads)
            clear_bits[x] = 0u;
        return;
    }
    if (b >= fix_start) {
        const int32_t i = (b - fix_start) * kRowThreads + threadIdx.x;
        bool mv = false;
        int32_t oc = 0, nc = 0;
        if (i < n) {
            const int32_t v = list[i];
            oc = a.newc[v];
            nc = comm[v];
            mv = oc != nc;
        }
        // one write per distinct joined community of the wave leader's key
        const uint64_t pend = __ballot(mv);
        if (pend) {
            const int leader = __ffsll((long long)pend) - 1;
            const int32_t lk = __shfl(nc, leader, 64);
            if (mv && (nc != lk || (int)(threadIdx.x & 63) == leader))
                crec[nc] = pack_rec(sigma[nc], size[nc]);
        }
        if (mv) crec[oc] = pack_rec(sigma[oc], size[oc]);
        return;
    }
    double acc = 0.0;
    auto row = [&](int32_t v, int64_t elo, int64_t ehi, int stride, bool first) {
        const int32_t cv_old = a.newc[v], cv_new = comm[v];
        if (first) acc += 2.0 * a.dk[v];
        // four entries a step, so that their gathers are in flight together
        for (int64_t e = elo; e < ehi; e += 4 * (int64_t)stride) {
            int32_t u[4];
            float x[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int64_t ek = e + k * (int64_t)stride;
                u[k] = ek < ehi ? a.col[ek] : v;
                x[k] = ek < ehi ? a.w[ek] : 0.0f;
            }
            bool mv[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) mv[k] = u[k] != v && test_bit(movers, u[k]);
            int32_t cn[4], co[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                cn[k] = mv[k] ? comm[u[k]] : 0;
                co[k] = mv[k] ? a.newc[u[k]] : 0;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (!mv[k]) continue;
                const int d = (int)(cn[k] == cv_new) - (int)(co[k] == cv_new) -
                              (int)(cv_old == cn[k]) + (int)(co[k] == cv_old);
                if (d != 0) acc += (double)d * (double)x[k];
            }
        }
    };
    // The block checks its items (rounds x 256/X vertices of one segment, or
    // rounds heavy-row chunks; at most kCorrItems) at once, keeps the
    // movers' items in LDS and walks only their rows, X lanes per row.
    __shared__ int32_t items[kCorrItems];
    __shared__ int32_t nitems;
    const bool chunked = b >= st.chunk_block_start;
    int X = kRowThreads;
    int32_t first = 0, cnt = 0, list_off = 0;
    if (chunked) {
        const int32_t per = corr_items(kRowThreads, rounds);
        first = (b - st.chunk_block_start) * per;
        cnt = min(per, st.nchunks - first);
    } else {
        int si = 0;
        while (si + 1 < st.nseg && b >= st.s[si + 1].block_start) ++si;
        const Seg sg = st.s[si];
        X = corr_lanes(sg.lanes_log2);
        const int32_t per = corr_items(X, rounds);
        first = (b - sg.block_start) * per;
        cnt = min(per, sg.count - first);
        list_off = sg.list_off;
    }
    if (threadIdx.x == 0) nitems = 0;
    __syncthreads();
    for (int32_t t = (int32_t)threadIdx.x; t < cnt; t += kRowThreads) {
        int32_t item = -1;
        if (chunked) {
            const int32_t v = chunks[first + t].v;
            if (a.newc[v] != comm[v]) item = first + t;
        } else {
            const int32_t v = order[list_off + first + t];
            if (a.newc[v] != comm[v]) item = v;
        }
        if (item >= 0) items[atomicAdd(&nitems, 1)] = item;
    }
    // batches: the block's items lie in ascending vertex order, so one graph
    // holds them all unless the first and the last differ
    __shared__ int32_t graph_of_block;
    if (gin && threadIdx.x == 0 && cnt > 0) {
        const int32_t v0 = chunked ? chunks[first].v : order[list_off + first];
        const int32_t v1 = chunked ? chunks[first + cnt - 1].v : order[list_off + first + cnt - 1];
        graph_of_block = gid[v0] == gid[v1] ? gid[v0] : -1;
    }
    __syncthreads();
    const int32_t nm = nitems;
    if (nm == 0) return;  // block-uniform
    const int32_t gblock = gin ? graph_of_block : -1;
    const int lane = (int)threadIdx.x & (X - 1);
    for (int32_t m = (int32_t)threadIdx.x / X; m < nm; m += kRowThreads / X) {
        int32_t v;
        if (chunked) {
            const Chunk ch = chunks[items[m]];
            v = ch.v;
            row(v, ch.lo + lane, ch.hi, X, lane == 0 && ch.first);
        } else {
            v = items[m];
            row(v, a.rowptr[v] + lane, a.rowptr[v + 1], X, lane == 0);
        }
        if (gin && gblock < 0) {
            // the row's sum over its lane group (over each wave when X = 256;
            // the X lanes of a group are active together)
            const int W = X < 64 ? X : 64;
            double r = acc;
            for (int o = W >> 1; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
            if ((threadIdx.x & (W - 1)) == 0 && r != 0.0) atomicAdd(&gin[gid[v]], r);
            acc = 0.0;
        }
    }
    if (gin && gblock < 0) return;  // block-uniform
    acc = block_sum_f64(acc, lds);
    if (gin) {
        if (threadIdx.x == 0 && acc != 0.0) atomicAdd(&gin[gblock], acc);
        return;
    }
    if (threadIdx.x == 0 && acc != 0.0)
        atomicAdd(&own[(int64_t)(blockIdx.x % kMoveSlots) * kMoveSlotStride], acc);
}

// Degree group of a vertex (modularity_optimisation.cuh:7-19): the order in
// which a sweep decides and applies the groups.
__device__ __forceinline__ int group_of_degree(int32_t deg) {
    int g = 0;
#pragma unroll
    for (int i = 1; i < kNumGroups; ++i) g += deg > kGroupBound[i] ? 1 : 0;
    return g;
}

// Change over one sweep of the weight inside communities.  Per entry
// (v, u, w) of a mover v's row it is e = w ([c1(u) = c1(v)] - [c0(u) = c0(v)])
// (c0, c1: communities before and after the sweep); the entry's mirror in
// u's row has the same w and e.  A mover's 2 dk[v] (dk = K[new] - K[old]
// with its neighbours in the communities its decision saw) covers both
// entries of every neighbour u that did not move exactly.  For a neighbour
// that moved, dk counted 2 w ([cu = c1(v)] - [cu = c0(v)]) with cu = u's
// community at v's decision: c1(u) when u's degree group came earlier, c0(u)
// when it came later or is v's own (its decisions are simultaneous); the
// entry (v, u) replaces that by e, and (u, v) is seen from u's row.  Each
// vertex moves at most once per sweep.  Blocks as k_apply_finish's
// correction: corr_items rows per X-lane group, the movers compacted in LDS,
// only their rows walked, the others' bits tested in the sweep's mover
// bitmap; the sum goes to the kMoveSlots slots.
__global__ __launch_bounds__(kRowThreads) void k_sweep_delta(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const float *__restrict__ w, const int32_t *__restrict__ comm,
    const int32_t *__restrict__ comm0, const double *__restrict__ dk,
    const uint32_t *__restrict__ movers, const int32_t *__restrict__ order, SegTable st,
    const Chunk *__restrict__ chunks, int32_t rounds, double *own) {
    __shared__ double lds[kRowThreads / 64];
    __shared__ int32_t items[kCorrItems];
    __shared__ int32_t nitems;
    const int b = blockIdx.x;
    const bool chunked = b >= st.chunk_block_start;
    int X = kRowThreads;
    int32_t first = 0, cnt = 0, list_off = 0;
    if (chunked) {
        const int32_t per = corr_items(kRowThreads, rounds);
        first = (b - st.chunk_block_start) * per;
        cnt = min(per, st.nchunks - first);
    } else {
        int si = 0;
        while (si + 1 < st.nseg && b >= st.s[si + 1].block_start) ++si;
        const Seg sg = st.s[si];
        X = corr_lanes(sg.lanes_log2);
        const int32_t per = corr_items(X, rounds);
        first = (b - sg.block_start) * per;
        cnt = min(per, sg.count - first);
        list_off = sg.list_off;
    }
    if (threadIdx.x == 0) nitems = 0;
    __syncthreads();
    for (int32_t t = (int32_t)threadIdx.x; t < cnt; t += kRowThreads) {
        const int32_t v = chunked ? chunks[first + t].v : order[list_off + first + t];
        if (test_bit(movers, v)) items[atomicAdd(&nitems, 1)] = chunked ? first + t : v;
    }
    __syncthreads();
    const int32_t nm = nitems;
    if (nm == 0) return;  // block-uniform
    const int lane = (int)threadIdx.x & (X - 1);
    double acc = 0.0;
    for (int32_t m = (int32_t)threadIdx.x / X; m < nm; m += kRowThreads / X) {
        int32_t v;
        int64_t lo, hi;
        bool first_piece;
        if (chunked) {
            const Chunk ch = chunks[items[m]];
            v = ch.v;
            lo = ch.lo;
            hi = ch.hi;
            first_piece = ch.first != 0;
        } else {
            v = items[m];
            lo = rowptr[v];
            hi = rowptr[v + 1];
            first_piece = true;
        }
        if (lane == 0 && first_piece) acc += 2.0 * dk[v];
        const int32_t c1v = comm[v], c0v = comm0[v];
        const int gv = group_of_degree(rowptr[v + 1] - rowptr[v]);
        // four entries a step, so that their gathers are in flight together
        for (int64_t e = lo + lane; e < hi; e += 4 * (int64_t)X) {
            int32_t u[4];
            float x[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int64_t ek = e + k * (int64_t)X;
                u[k] = ek < hi ? col[ek] : v;
                x[k] = ek < hi ? w[ek] : 0.0f;
            }
            bool mv[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) mv[k] = u[k] != v && test_bit(movers, u[k]);
            int32_t c1u[4], c0u[4], du[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                c1u[k] = mv[k] ? comm[u[k]] : 0;
                c0u[k] = mv[k] ? comm0[u[k]] : 0;
                du[k] = mv[k] ? rowptr[u[k] + 1] - rowptr[u[k]] : 0;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (!mv[k]) continue;
                const int32_t cut = group_of_degree(du[k]) < gv ? c1u[k] : c0u[k];
                const int d = (int)(c1u[k] == c1v) - (int)(c0u[k] == c0v) -
                              2 * ((int)(cut == c1v) - (int)(cut == c0v));
                if (d != 0) acc += (double)d * (double)x[k];
            }
        }
    }
    acc = block_sum_f64(acc, lds);
    if (threadIdx.x == 0 && acc != 0.0)
        atomicAdd(&own[(int64_t)(blockIdx.x % kMoveSlots) * kMoveSlotStride], acc);
}

// Self-loop weight: one wave per row (grid-stride), lanes over its entries.
// Weight of the self-loop entries (col[e] == row) of rows sorted by column:
// one thread per row finds its diagonal entry by binary search (a few reads
// per row instead of the row); per-block partials.
__global__ __launch_bounds__(256) void k_selfloop_sorted(const int32_t *__restrict__ rowptr,
                                                         const int32_t *__restrict__ col,
                                                         const float *__restrict__ w, int32_t V,
                                                         double *partials) {
    __shared__ double lds[4];
    double x = 0.0;
    for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * 256) {
        int32_t lo = rowptr[v], hi = rowptr[v + 1];  // first e in [lo, hi) with col[e] >= v
        while (lo < hi) {
            const int32_t m = lo + ((hi - lo) >> 1);
            if (col[m] < (int32_t)v) lo = m + 1; else hi = m;
        }
        if (lo < rowptr[v + 1] && col[lo] == (int32_t)v) x += (double)w[lo];
    }
    x = block_sum_f64(x, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = x;
}

// partials[0, n) folded into kSelfLoopBlocks sums (block b: b, b + grid, ...)
__global__ __launch_bounds__(256) void k_selfloop_fold(const double *__restrict__ partials,
                                                       int32_t n, double *folded) {
    __shared__ double lds[4];
    double x = 0.0;
    for (int32_t i = blockIdx.x + gridDim.x * threadIdx.x; i < n; i += gridDim.x * 256)
        x += partials[i];
    x = block_sum_f64(x, lds);
    if (threadIdx.x == 0) folded[blockIdx.x] = x;
}

__global__ __launch_bounds__(256) void k_selfloop_final(const double *__restrict__ partials,
                                                        int32_t n, double *out) {
    __shared__ double lds[4];
    double x = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) x += partials[i];
    x = block_sum_f64(x, lds);
    if (threadIdx.x == 0) *out = x;
}

__global__ void k_records_all(int32_t V, const double *__restrict__ sigma,
                              const int32_t *__restrict__ size, uint64_t *crec) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < V) crec[c] = pack_rec(sigma[c], size[c]);
}

// ------------------------------------------------------- row reductions

// Weight from a row slice [lo, hi) (stride `st`) to community cv; four
// entries per step so that their gathers are in flight together.
__device__ __forceinline__ double own_weight(const int32_t *__restrict__ col,
                                             const float *__restrict__ w,
                                             const int32_t *__restrict__ comm, int32_t cv,
                                             int64_t lo, int64_t hi, int st) {
    double own = 0.0;
    int64_t e = lo;
    for (; e + 3 * st < hi; e += 4 * st) {
        const int32_t u0 = col[e], u1 = col[e + st], u2 = col[e + 2 * st], u3 = col[e + 3 * st];
        const float w0 = w[e], w1 = w[e + st], w2 = w[e + 2 * st], w3 = w[e + 3 * st];
        const int32_t c0 = comm[u0], c1 = comm[u1], c2 = comm[u2], c3 = comm[u3];
        own += c0 == cv ? (double)w0 : 0.0;
        own += c1 == cv ? (double)w1 : 0.0;
        own += c2 == cv ? (double)w2 : 0.0;
        own += c3 == cv ? (double)w3 : 0.0;
    }
    for (; e < hi; e += st) {
        const int32_t u = col[e];
        const float wf = w[e];
        if (comm[u] == cv) own += (double)wf;
    }
    return own;
}

__device__ __forceinline__ int find_seg(const SegTable &st, int b) {
    int si = 0;
    while (si + 1 < st.nseg && b >= st.s[si + 1].block_start) ++si;
    return si;
}

__global__ void k_degree_keys(const int32_t *__restrict__ rowptr, int32_t V, BucketBounds bb,
                              int8_t *keys) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) keys[v] = (int8_t)bucket_of(rowptr[v + 1] - rowptr[v], bb);
}

// k_v = sum of row weights (computeEdgesSum, modularity_optimisation.cu:23-40), FP64.
// Strengths k_v = sum of v's row weights.  With loops (per-block partials),
// also the weight of the self-loop entries (col[e] == v) per block.
__global__ __launch_bounds__(kRowThreads) void k_strength(const int32_t *__restrict__ rowptr,
                                                          const float *__restrict__ w,
                                                          const int32_t *__restrict__ order,
                                                          SegTable st,
                                                          const Chunk *__restrict__ chunks,
                                                          double *k,
                                                          const int32_t *__restrict__ col,
                                                          double *loops) {
    __shared__ double lds[kRowThreads / 64];
    const int b = blockIdx.x;
    double sl = 0.0;  // self-loop weight seen by this thread
    if (b >= st.chunk_block_start) {
        const Chunk ch = chunks[b - st.chunk_block_start];
        double s = 0.0;
        for (int64_t e = ch.lo + threadIdx.x; e < ch.hi; e += kRowThreads) {
            s += (double)w[e];
            if (loops && col[e] == ch.v) sl += (double)w[e];
        }
        s = block_sum_f64(s, lds);
        if (threadIdx.x == 0) atomicAdd(&k[ch.v], s);
    } else {
        const Seg sg = st.s[find_seg(st, b)];
        const int lg = sg.lanes_log2;
        if (lg <= 6) {
            const int X = 1 << lg;
            const int32_t idx = (b - sg.block_start) * (kRowThreads >> lg) + (threadIdx.x >> lg);
            const int lane = threadIdx.x & (X - 1);
            double s = 0.0;
            int32_t v = -1;
            if (idx < sg.count) {
                v = order[sg.list_off + idx];
                for (int32_t e = rowptr[v] + lane; e < rowptr[v + 1]; e += X) {
                    s += (double)w[e];
                    if (loops && col[e] == v) sl += (double)w[e];
                }
            }
            for (int off = X / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
            if (lane == 0 && v >= 0) k[v] = s;
        } else {
            const int32_t v = order[sg.list_off + (b - sg.block_start)];
            double s = 0.0;
            for (int32_t e = rowptr[v] + threadIdx.x; e < rowptr[v + 1]; e += kRowThreads) {
                s += (double)w[e];
                if (loops && col[e] == v) sl += (double)w[e];
            }
            s = block_sum_f64(s, lds);
            if (threadIdx.x == 0) k[v] = s;
        }
    }
    if (loops) {
        __syncthreads();  // lds is reused
        sl = block_sum_f64(sl, lds);
        if (threadIdx.x == 0) loops[b] = sl;
    }
}

// Column numbering of the move kernels (MoveArgs::ccol): cpos[order[i]] = i,
// order listing every vertex once, bucket by bucket; ccol[e] = cpos[col[e]].
__global__ void k_col_pos(const int32_t *__restrict__ order, int32_t V, int32_t *cpos) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) cpos[order[i]] = i;
}

__global__ void k_col_renumber(const int32_t *__restrict__ col, int64_t E,
                               const int32_t *__restrict__ cpos, int32_t *ccol) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x)
        ccol[e] = cpos[__builtin_nontemporal_load(&col[e])];
}

// ccomm[i] = comm[order[i]]: the communities in that numbering.
__global__ void k_col_comm(const int32_t *__restrict__ order, int32_t V,
                           const int32_t *__restrict__ comm, int32_t *ccomm) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) ccomm[i] = comm[order[i]];
}

__global__ void k_level_state(int32_t V, const double *__restrict__ k, int32_t *comm,
                              int32_t *size, double *sigma, uint64_t *crec) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) {
        comm[v] = v;
        size[v] = 1;
        sigma[v] = k[v];
        crec[v] = pack_rec(k[v], 1);
    }
}

// Fused modularity (calculateToOwnCommunity + the two reductions,
// modularity_optimisation.cu:365-401): one launch produces per-block
// partials of sum_v (weight to own community) and sum_c S_c^2; a one-block
// pass adds them in a fixed order (run-to-run deterministic).
// The row entries name their columns through gcol and gcomm (the move
// kernels' renumbered columns, or col and comm themselves); comm[v] is the
// row's own community.
__global__ __launch_bounds__(kRowThreads) void k_modularity(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ gcol,
    const float *__restrict__ w, const int32_t *__restrict__ comm,
    const int32_t *__restrict__ gcomm, const double *__restrict__ sigma, int32_t V,
    const int32_t *__restrict__ order, SegTable st, const Chunk *__restrict__ chunks,
    double *partials) {
    __shared__ double lds[kRowThreads / 64];
    const int b = blockIdx.x;
    double own = 0.0, sq = 0.0;
    if (b >= st.sq_block_start) {
        const int64_t base = (int64_t)(b - st.sq_block_start) * kSqItems;
        for (int j = threadIdx.x; j < kSqItems; j += kRowThreads) {
            const int64_t c = base + j;
            if (c < V) {
                const double x = sigma[c];
                sq += x * x;
            }
        }
    } else if (b >= st.chunk_block_start) {
        const Chunk ch = chunks[b - st.chunk_block_start];
        own = own_weight(gcol, w, gcomm, comm[ch.v], ch.lo + threadIdx.x, ch.hi, kRowThreads);
    } else {
        const Seg sg = st.s[find_seg(st, b)];
        const int lg = sg.lanes_log2;
        if (lg <= 6) {
            const int X = 1 << lg;
            const int32_t idx = (b - sg.block_start) * (kRowThreads >> lg) + (threadIdx.x >> lg);
            const int lane = threadIdx.x & (X - 1);
            if (idx < sg.count) {
                const int32_t v = order[sg.list_off + idx];
                own = own_weight(gcol, w, gcomm, comm[v], rowptr[v] + lane, rowptr[v + 1], X);
            }
        } else {
            const int32_t v = order[sg.list_off + (b - sg.block_start)];
            own = own_weight(gcol, w, gcomm, comm[v], rowptr[v] + threadIdx.x, rowptr[v + 1],
                             kRowThreads);
        }
    }
    own = block_sum_f64(own, lds);
    sq = block_sum_f64(sq, lds);
    if (threadIdx.x == 0) {
        partials[2 * b] = own;
        partials[2 * b + 1] = sq;
    }
}

// Neighbour moves counted from the rows' side: marks = the vertices that
// moved in the last sweep (hist bit 0) with degree 1..max_degree, one bit
// each; then every row counts its entries whose column is marked.  The same
// counts as the movers adding one per entry of their rows (the CSR is
// symmetric), without atomics on the neighbours: a sweep in which many
// vertices moved pushes to the same hubs from many rows at once.
__global__ void k_mark_movers(const int32_t *__restrict__ rowptr,
                              const uint8_t *__restrict__ hist, int32_t V, int32_t max_degree,
                              int32_t nwords, uint32_t *marks) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    bool m = false;
    if (v < V && (hist[v] & 1u)) {
        const int32_t d = rowptr[v + 1] - rowptr[v];
        m = d > 0 && d <= max_degree;
    }
    const uint64_t b = __ballot(m);
    const int lane = (int)threadIdx.x & 63;
    const int32_t w0 = (v - lane) >> 5;  // the wave's first word (v - lane: multiple of 64)
    if (lane == 0 && w0 < nwords) marks[w0] = (uint32_t)b;
    if (lane == 32 && w0 + 1 < nwords) marks[w0 + 1] = (uint32_t)(b >> 32);
}

// Marked columns in a row slice [lo, hi) (stride st), four entries a step.
__device__ __forceinline__ int32_t marked_in(const int32_t *__restrict__ col,
                                             const uint32_t *__restrict__ marks, int64_t lo,
                                             int64_t hi, int st) {
    int32_t c = 0;
    int64_t e = lo;
    for (; e + 3 * st < hi; e += 4 * st) {
        const int32_t u0 = col[e], u1 = col[e + st], u2 = col[e + 2 * st], u3 = col[e + 3 * st];
        const uint32_t m0 = marks[u0 >> 5], m1 = marks[u1 >> 5], m2 = marks[u2 >> 5],
                       m3 = marks[u3 >> 5];
        c += (int32_t)((m0 >> (u0 & 31)) & 1u) + (int32_t)((m1 >> (u1 & 31)) & 1u) +
             (int32_t)((m2 >> (u2 & 31)) & 1u) + (int32_t)((m3 >> (u3 & 31)) & 1u);
    }
    for (; e < hi; e += st) c += test_bit(marks, col[e]) ? 1 : 0;
    return c;
}

__device__ __forceinline__ int32_t block_sum_i32(int32_t x, int32_t *lds) {
    x = wave_sum_i32(x);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = x;
    __syncthreads();
    int32_t t = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += lds[i];
    return t;
}

// nmv[v] = marked entries of v's row, rows laid out as in the modularity
// pass (st); the chunks of a heavy row add their counts (nmv zero before).
__global__ __launch_bounds__(kRowThreads) void k_neighbor_count(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const int32_t *__restrict__ order, SegTable st, const Chunk *__restrict__ chunks,
    const uint32_t *__restrict__ marks, int32_t *nmv) {
    __shared__ int32_t red[kRowThreads / 64];
    const int b = blockIdx.x;
    if (b >= st.chunk_block_start) {
        const Chunk ch = chunks[b - st.chunk_block_start];
        const int32_t c = block_sum_i32(
            marked_in(col, marks, ch.lo + threadIdx.x, ch.hi, kRowThreads), red);
        if (threadIdx.x == 0 && c != 0) atomicAdd(&nmv[ch.v], c);
        return;
    }
    const Seg sg = st.s[find_seg(st, b)];
    const int lg = sg.lanes_log2;
    if (lg <= 6) {
        const int X = 1 << lg;
        const int32_t idx = (b - sg.block_start) * (kRowThreads >> lg) + (threadIdx.x >> lg);
        const int lane = threadIdx.x & (X - 1);
        int32_t v = -1, c = 0;
        if (idx < sg.count) {
            v = order[sg.list_off + idx];
            c = marked_in(col, marks, rowptr[v] + lane, rowptr[v + 1], X);
        }
        for (int off = X >> 1; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
        if (lane == 0 && v >= 0) nmv[v] = c;
    } else {
        const int32_t v = order[sg.list_off + (b - sg.block_start)];
        const int32_t c = block_sum_i32(
            marked_in(col, marks, rowptr[v] + threadIdx.x, rowptr[v + 1], kRowThreads), red);
        if (threadIdx.x == 0) nmv[v] = c;
    }
}

// ---------------------------------------------------- XCD-local row passes
//
// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs on
// XCD b % 8), and each XCD has its own L2.  The heavy rows (degree > 32,
// sorted by column) are cut at the column-block boundaries x*ceil(V/8):
// workgroup b handles column block b % 8 of its rows only, so every random
// comm[u] gather of an XCD falls in one eighth of comm (2 MB at V = 4M),
// which stays in that XCD's L2.  Row data is streamed with non-temporal
// loads so that it does not evict the slice.

// split[i*9 + x] = first edge of row list[i] with column >= x*blk (x = 0..8;
// entry 0 = row start, 8 = row end).  Meaningful for sorted rows only
// (k_rows_sorted checks).
__global__ void k_xcd_splits(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                             const int32_t *__restrict__ list, int32_t n, int32_t blk,
                             int32_t *split) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * 9) return;
    const int32_t i = (int32_t)(t / 9), x = (int32_t)(t % 9);
    const int32_t v = list[i];
    const int32_t lo = rowptr[v], hi = rowptr[v + 1];
    int32_t r;
    if (x == 0) {
        r = lo;
    } else if (x == 8) {
        r = hi;
    } else {
        const int64_t bound = (int64_t)x * blk;
        int32_t a = lo, z = hi;  // first e in [lo, hi) with col[e] >= bound
        while (a < z) {
            const int32_t m = a + ((z - a) >> 1);
            if ((int64_t)col[m] < bound) a = m + 1; else z = m;
        }
        r = a;
    }
    split[t] = r;
}

// Sortedness of the listed rows: one thread per edge pair.
__global__ void k_rows_sorted(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                              const int32_t *__restrict__ list, int32_t n, int32_t *unsorted) {
    const int32_t i = blockIdx.x;
    if (i >= n) return;
    const int32_t v = list[i];
    const int32_t lo = rowptr[v], hi = rowptr[v + 1];
    bool bad = false;
    for (int32_t e = lo + 1 + threadIdx.x; e < hi; e += blockDim.x) bad |= col[e - 1] > col[e];
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(unsorted, 1);
}

__device__ __forceinline__ double own_weight_nt(const int32_t *__restrict__ col,
                                                const float *__restrict__ w,
                                                const int32_t *__restrict__ comm, int32_t cv,
                                                int64_t lo, int64_t hi, int st) {
    double own = 0.0;
    int64_t e = lo;
    for (; e + 3 * st < hi; e += 4 * st) {
        const int32_t u0 = __builtin_nontemporal_load(&col[e]);
        const int32_t u1 = __builtin_nontemporal_load(&col[e + st]);
        const int32_t u2 = __builtin_nontemporal_load(&col[e + 2 * st]);
        const int32_t u3 = __builtin_nontemporal_load(&col[e + 3 * st]);
        const float w0 = __builtin_nontemporal_load(&w[e]);
        const float w1 = __builtin_nontemporal_load(&w[e + st]);
        const float w2 = __builtin_nontemporal_load(&w[e + 2 * st]);
        const float w3 = __builtin_nontemporal_load(&w[e + 3 * st]);
        const int32_t c0 = comm[u0], c1 = comm[u1], c2 = comm[u2], c3 = comm[u3];
        own += c0 == cv ? (double)w0 : 0.0;
        own += c1 == cv ? (double)w1 : 0.0;
        own += c2 == cv ? (double)w2 : 0.0;
        own += c3 == cv ? (double)w3 : 0.0;
    }
    for (; e < hi; e += st) {
        const int32_t u = __builtin_nontemporal_load(&col[e]);
        const float wf = __builtin_nontemporal_load(&w[e]);
        if (comm[u] == cv) own += (double)wf;
    }
    return own;
}

// Weight inside communities over the heavy rows, by column block: grid =
// 8 * xst.total_blocks; partials of block b at partials[2*(base + b)].
__global__ __launch_bounds__(kRowThreads) void k_modularity_xcd(
    const int32_t *__restrict__ col, const float *__restrict__ w,
    const int32_t *__restrict__ comm, const int32_t *__restrict__ list, SegTable xst,
    const int32_t *__restrict__ split, double *partials, int32_t base) {
    __shared__ double lds[kRowThreads / 64];
    const int b = blockIdx.x;
    const int x = b & 7, j = b >> 3;
    const Seg sg = xst.s[find_seg(xst, j)];
    const int lg = sg.lanes_log2;
    double own = 0.0;
    if (lg <= 6) {
        const int X = 1 << lg;
        const int32_t idx = (j - sg.block_start) * (kRowThreads >> lg) + (threadIdx.x >> lg);
        if (idx < sg.count) {
            const int32_t i = sg.list_off + idx;
            const int32_t v = list[i];
            own = own_weight_nt(col, w, comm, comm[v], split[i * 9 + x] + (threadIdx.x & (X - 1)),
                                split[i * 9 + x + 1], X);
        }
    } else {
        const int32_t i = sg.list_off + (j - sg.block_start);
        const int32_t v = list[i];
        own = own_weight_nt(col, w, comm, comm[v], split[i * 9 + x] + threadIdx.x,
                            split[i * 9 + x + 1], kRowThreads);
    }
    own = block_sum_f64(own, lds);
    if (threadIdx.x == 0) {
        partials[2 * (base + b)] = own;
        partials[2 * (base + b) + 1] = 0.0;
    }
}

// Sum of S_c^2 only (incremental sweeps keep the weight inside communities
// current on their own).
__global__ __launch_bounds__(kRowThreads) void k_sum_squares(const double *__restrict__ sigma,
                                                             int32_t V, double *partials) {
    __shared__ double lds[kRowThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kSqItems;
    double sq = 0.0;
    for (int j = threadIdx.x; j < kSqItems; j += kRowThreads) {
        const int64_t c = base + j;
        if (c < V) {
            const double x = sigma[c];
            sq += x * x;
        }
    }
    sq = block_sum_f64(sq, lds);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = 0.0;
        partials[2 * blockIdx.x + 1] = sq;
    }
}

// Two-level fixed-order reduction of the per-block partials: 256 blocks
// each add a contiguous range, then one block adds the 256 results.  The
// weight inside communities is summed over all nb pairs (its sums are exact,
// so their order does not matter); the sum of S_c^2 only over the nb_sq
// pairs of the S_c^2 blocks from sq_lo, split by nb_sq alone, so that a full
// pass and an incremental sweep (k_sum_squares) add S_c^2 in the same order
// and give the same modularity bit for bit.
constexpr int kQReduceBlocks = 256;

__global__ __launch_bounds__(256) void k_modularity_reduce(const double *__restrict__ partials,
                                                           int32_t nb, int32_t sq_lo, int32_t nb_sq,
                                                           double *partials2) {
    __shared__ double lds[4];
    const int64_t lo = (int64_t)nb * blockIdx.x / kQReduceBlocks;
    const int64_t hi = (int64_t)nb * (blockIdx.x + 1) / kQReduceBlocks;
    const int64_t slo = sq_lo + (int64_t)nb_sq * blockIdx.x / kQReduceBlocks;
    const int64_t shi = sq_lo + (int64_t)nb_sq * (blockIdx.x + 1) / kQReduceBlocks;
    double own = 0.0, sq = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) own += partials[2 * i];
    for (int64_t i = slo + threadIdx.x; i < shi; i += 256) sq += partials[2 * i + 1];
    own = block_sum_f64(own, lds);
    sq = block_sum_f64(sq, lds);
    if (threadIdx.x == 0) {
        partials2[2 * blockIdx.x] = own;
        partials2[2 * blockIdx.x + 1] = sq;
    }
}

// own_slots (incremental): the weight inside communities is q[1] plus the
// corrections the applies left in the slots, which are cleared here.
__global__ __launch_bounds__(256) void k_modularity_final(const double *__restrict__ partials2,
                                                          double M, double *q, double *own_slots) {
    __shared__ double lds[4];
    double own = partials2[2 * threadIdx.x];
    double sq = partials2[2 * threadIdx.x + 1];
    double d = 0.0;
    if (own_slots && threadIdx.x < kMoveSlots) {
        d = own_slots[threadIdx.x * kMoveSlotStride];
        own_slots[threadIdx.x * kMoveSlotStride] = 0.0;
    }
    own = block_sum_f64(own, lds);
    sq = block_sum_f64(sq, lds);
    // fixed-order sum of the slots (the additions inside a slot are exact
    // for weights on a dyadic grid)
    d = block_sum_f64(d, lds);
    if (threadIdx.x == 0) {
        if (own_slots) own = q[1] + d;
        q[0] = own / (2.0 * M) - sq / (4.0 * M * M);
        q[1] = own;
        q[2] = sq;
    }
}

__global__ void k_compose(int32_t *orig2comm, int32_t V0, const int32_t *__restrict__ map) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < V0) orig2comm[x] = map[orig2comm[x]];
}

inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

void launch_degree_keys(const int32_t *rowptr, int32_t V, BucketBounds bb, int8_t *keys,
                        hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_degree_keys, dim3(blocks_for(V, 256)), dim3(256), 0, s, rowptr, V, bb,
                       keys);
}

void launch_strength(const int32_t *rowptr, const float *w, const int32_t *order,
                     const SegTable &st, const Chunk *chunks, double *k, hipStream_t s,
                     const int32_t *col, double *loop_partials, double *loops_out) {
    const int32_t nb = st.chunk_block_start + st.nchunks;
    if (nb <= 0) {
        if (loops_out) fill_f64(loops_out, 0.0, 1, s);
        return;
    }
    hipLaunchKernelGGL(k_strength, dim3(nb), dim3(kRowThreads), 0, s, rowptr, w, order, st, chunks,
                       k, loops_out ? col : nullptr, loops_out ? loop_partials : nullptr);
    if (loops_out) {
        // partials of the nb blocks, then kSelfLoopBlocks folds, in fixed orders
        double *folded = loop_partials + nb;
        hipLaunchKernelGGL(k_selfloop_fold, dim3(kSelfLoopBlocks), dim3(256), 0, s, loop_partials,
                           nb, folded);
        hipLaunchKernelGGL(k_selfloop_final, dim3(1), dim3(256), 0, s, folded,
                           (int32_t)kSelfLoopBlocks, loops_out);
    }
}

void launch_level_state(int32_t V, const double *k, int32_t *comm, int32_t *size, double *sigma,
                        uint64_t *crec, hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_level_state, dim3(blocks_for(V, 256)), dim3(256), 0, s, V, k, comm, size,
                       sigma, crec);
}

void launch_records_all(int32_t V, const double *sigma, const int32_t *size, uint64_t *crec,
                        hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_records_all, dim3(blocks_for(V, 256)), dim3(256), 0, s, V, sigma, size,
                       crec);
}

bool move_sub_shape(int X, int E) {
    switch (X * 64 + E) {
    case 4 * 64 + 1: case 8 * 64 + 1: case 16 * 64 + 1: case 32 * 64 + 1:
    case 4 * 64 + 2: case 4 * 64 + 4: case 4 * 64 + 8:
    case 8 * 64 + 2: case 8 * 64 + 4: case 16 * 64 + 2:
        return true;
    default:
        return false;
    }
}

void launch_move_sub(int X, int E, const MoveArgs &a, const int32_t *list, int32_t n,
                     hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = blocks_for((int64_t)n * X, 256);
#define LV_SUB(x, e)                                                                   \
    case x * 64 + e:                                                                   \
        hipLaunchKernelGGL((k_move_sub<x, e>), dim3(g), dim3(256), 0, s, a, list, n); \
        break;
    switch (X * 64 + E) {
        LV_SUB(4, 1)
        LV_SUB(8, 1)
        LV_SUB(16, 1)
        LV_SUB(4, 2)
        LV_SUB(4, 4)
        LV_SUB(4, 8)
        LV_SUB(8, 2)
        LV_SUB(8, 4)
        LV_SUB(16, 2)
    default:
        hipLaunchKernelGGL((k_move_sub<32, 1>), dim3(g), dim3(256), 0, s, a, list, n);
        break;
    }
#undef LV_SUB
}

void launch_move_wave(int cap, const MoveArgs &a, const int32_t *list, int32_t n, hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = blocks_for(n, 4);
    if (cap == 128)
        hipLaunchKernelGGL((k_move_wave<128, 3, 32>), dim3(blocks_for(n, 8)), dim3(256), 0, s, a,
                           list, n);
    else if (cap == 192)
        hipLaunchKernelGGL((k_move_wave<192, 4, 32>), dim3(blocks_for(n, 8)), dim3(256), 0, s, a,
                           list, n);
    else if (cap == 256)
        hipLaunchKernelGGL((k_move_wave<256, 2>), dim3(g), dim3(256), 0, s, a, list, n);
    else
        hipLaunchKernelGGL((k_move_wave<448, 0>), dim3(g), dim3(256), 0, s, a, list, n);
}

namespace {
// One workgroup tier at `lanes` lanes per row (instances per tier, so that
// each tier is its own kernel in traces).
template <int TIER, bool HINTS>
void launch_wg_tier_h(int lanes, unsigned g, size_t lds, hipStream_t s, const MoveArgs &a,
                      const int32_t *list, int32_t n, int slots, int pct, int32_t *hint,
                      int32_t *done) {
    switch (lanes) {
    case 128:
        hipLaunchKernelGGL((k_move_wg<128, TIER, HINTS>), dim3(g), dim3(128), lds, s, a, list, n,
                           slots, pct, hint, done, 0);
        break;
    case 256:
        hipLaunchKernelGGL((k_move_wg<256, TIER, HINTS>), dim3(g), dim3(256), lds, s, a, list, n,
                           slots, pct, hint, done, 0);
        break;
    case 512:
        hipLaunchKernelGGL((k_move_wg<512, TIER, HINTS>), dim3(g), dim3(512), lds, s, a, list, n,
                           slots, pct, hint, done, 0);
        break;
    default:
        hipLaunchKernelGGL((k_move_wg<1024, TIER, HINTS>), dim3(g), dim3(1024), lds, s, a, list, n,
                           slots, pct, hint, done, 0);
        break;
    }
}

template <int TIER>
void launch_wg_tier(int lanes, unsigned g, size_t lds, hipStream_t s, const MoveArgs &a,
                    const int32_t *list, int32_t n, int slots, int pct, int32_t *hint,
                    int32_t *done) {
    if (hint && done)
        launch_wg_tier_h<TIER, true>(lanes, g, lds, s, a, list, n, slots, pct, hint, done);
    else
        launch_wg_tier_h<TIER, false>(lanes, g, lds, s, a, list, n, slots, pct, nullptr, nullptr);
}
}  // namespace

void launch_move_wg(int tier, int lanes, int slots, int pct, const MoveArgs &a,
                    const int32_t *list, int32_t n, int32_t *hint, int32_t *done, hipStream_t s) {
    if (n <= 0) return;
    const size_t lds = (size_t)12 * (size_t)slots;
    const unsigned g = (unsigned)(n < 1048576 ? n : 1048576);
    switch (tier) {
    case 0: launch_wg_tier<0>(lanes, g, lds, s, a, list, n, slots, pct, hint, done); break;
    case 1: launch_wg_tier<1>(lanes, g, lds, s, a, list, n, slots, pct, hint, done); break;
    case 2: launch_wg_tier<2>(lanes, g, lds, s, a, list, n, slots, pct, hint, done); break;
    default: launch_wg_tier<3>(lanes, g, lds, s, a, list, n, slots, pct, hint, done); break;
    }
}

void launch_move_hinted(const MoveArgs &a, const int32_t *list, int32_t n, int32_t *hint,
                        int32_t *done, int32_t n2, int32_t *hint2, int32_t *done2,
                        hipStream_t s) {
    if (n + n2 <= 0) return;
    const int64_t b = ((int64_t)n + n2 + 3) / 4;
    hipLaunchKernelGGL(k_move_hinted, dim3((unsigned)(b < 16384 ? b : 16384)), dim3(256), 0, s, a,
                       list, n, hint, done, n2, hint2, done2);
}

void launch_heavy_try(int slots, int32_t limit, const MoveArgs &a, const int32_t *list,
                      int32_t n, const ChunkedTier &t, hipStream_t s) {
    if (n <= 0) return;
    if (limit > kHeavyTryLimit) limit = kHeavyTryLimit;
    const size_t lds = (size_t)12 * (size_t)slots;
    const unsigned g = (unsigned)(n < 1048576 ? n : 1048576);
    hipLaunchKernelGGL((k_move_wg<1024, 4>), dim3(g), dim3(1024), lds, s, a, list, n, slots,
                       100, t.hint, t.done, limit);
}

void launch_chunk_aggregate(const MoveArgs &a, const ChunkedTier &t, int32_t nchunks,
                            hipStream_t s) {
    const int32_t g = t.chunk_order ? t.n_chunk_order : nchunks;
    if (g <= 0) return;
    const size_t lds = ((size_t)12 << t.chunk_cap_log2) + sizeof(int32_t) * kMaxParts;
    hipLaunchKernelGGL(k_chunk_aggregate, dim3(g), dim3(256), lds, s, a, t);
}

void launch_chunk_merge(const MoveArgs &a, const ChunkedTier &t, int32_t nitems, hipStream_t s) {
    const int32_t g = t.item_order ? t.n_item_order : nitems;
    if (g <= 0) return;
    hipLaunchKernelGGL(k_chunk_merge, dim3(g), dim3(kMergeThreads), 0, s, a, t);
}

void launch_heavy_final(const MoveArgs &a, const ChunkedTier &t, int32_t nh, hipStream_t s) {
    if (nh <= 0) return;
    hipLaunchKernelGGL(k_heavy_final, dim3(blocks_for(nh, 4)), dim3(256), 0, s, a, nh, t);
    hipLaunchKernelGGL(k_clear_overflow, dim3(nh), dim3(256), 0, s, t, nh);
}

void launch_move_chunked(const MoveArgs &a, const ChunkedTier &t, int32_t nchunks,
                         int32_t nitems, const int32_t *heavy_list, int32_t nh, hipStream_t s) {
    (void)heavy_list;
    if (nh <= 0) return;
    launch_chunk_aggregate(a, t, nchunks, s);
    launch_chunk_merge(a, t, nitems, s);
    launch_heavy_final(a, t, nh, s);
}

void launch_clear_tables(int32_t *keys, double *vals, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_clear_tables, dim3((unsigned)g), dim3(256), 0, s, keys, vals, n);
}

void launch_col_numbering(const int32_t *order, int32_t V, const int32_t *col, int64_t E,
                          int32_t *cpos, int32_t *ccol, hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_col_pos, dim3(blocks_for(V, 256)), dim3(256), 0, s, order, V, cpos);
    if (E > 0)
        hipLaunchKernelGGL(k_col_renumber, dim3((unsigned)std::min<int64_t>(blocks_for(E, 256), 65536)),
                           dim3(256), 0, s, col, E, cpos, ccol);
}

void launch_col_comm(const int32_t *order, int32_t V, const int32_t *comm, int32_t *ccomm,
                     hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_col_comm, dim3(blocks_for(V, 256)), dim3(256), 0, s, order, V, comm, ccomm);
}

void launch_apply(const MoveArgs &a, const int32_t *list, int32_t n, const int32_t *order,
                  const SegTable *gst, const Chunk *chunks, int32_t *comm, int32_t *size,
                  double *sigma, uint64_t *crec, const GroupMarks *gm, uint8_t *hist,
                  unsigned long long *moved, double *own, hipStream_t s, uint32_t *sweep_movers) {
    if (n <= 0 && !gm) return;
    if (n > 0) {
        hipLaunchKernelGGL(k_apply_join, dim3(blocks_for(n, 256)), dim3(256), 0, s, list, n,
                           a.rowptr, a.newc, comm, size, sigma, a.k, gm ? gm->movers : sweep_movers,
                           hist, moved, a.gate, const_cast<int32_t *>(a.ccomm), a.cself);
    }
    const int32_t corr = gm ? gst->total_blocks : 0;
    const int32_t fix = n > 0 ? (int32_t)blocks_for(n, kRowThreads) : 0;
    const int32_t clear = gm ? (int32_t)blocks_for(gm->words, kClearWords) : 0;
    const SegTable empty{};
    if (corr + fix + clear <= 0) return;
    hipLaunchKernelGGL(k_apply_finish, dim3(corr + fix + clear), dim3(kRowThreads), 0, s, a, order,
                       list, n, gm ? *gst : empty, chunks, comm, size, sigma, crec,
                       gm ? gm->movers : nullptr, gm ? gm->clear_movers : nullptr,
                       gm ? gm->words : 0, own, corr, corr + fix, gm ? gm->rounds : 1,
                       gm ? gm->gid : nullptr, gm ? gm->gin : nullptr);
}

void launch_sweep_delta(const int32_t *rowptr, const int32_t *col, const float *w,
                        const int32_t *comm, const int32_t *comm0, const double *dk,
                        const uint32_t *movers, const int32_t *order, const SegTable &ct,
                        const Chunk *chunks, int32_t rounds, double *own_slots, hipStream_t s) {
    if (ct.total_blocks <= 0) return;
    hipLaunchKernelGGL(k_sweep_delta, dim3(ct.total_blocks), dim3(kRowThreads), 0, s, rowptr, col, w,
                       comm, comm0, dk, movers, order, ct, chunks, rounds, own_slots);
}

void launch_modularity_rows(const int32_t *rowptr, const int32_t *col, const float *w,
                            const int32_t *comm, const double *sigma, int32_t V,
                            const int32_t *order, const SegTable &st, const Chunk *chunks,
                            double *partials, hipStream_t s, const int32_t *ccol,
                            const int32_t *ccomm) {
    hipLaunchKernelGGL(k_modularity, dim3(st.total_blocks), dim3(kRowThreads), 0, s, rowptr,
                       ccol ? ccol : col, w, comm, ccol ? ccomm : comm, sigma, V, order, st,
                       chunks, partials);
}

void launch_neighbor_count(const int32_t *rowptr, const int32_t *col, const uint8_t *hist,
                           int32_t V, int32_t max_degree, const int32_t *order,
                           const SegTable &st, const Chunk *chunks, uint32_t *marks,
                           int32_t *nmv, hipStream_t s) {
    if (V <= 0 || st.sq_block_start <= 0) return;
    hipLaunchKernelGGL(k_mark_movers, dim3(blocks_for(V, 256)), dim3(256), 0, s, rowptr, hist, V,
                       max_degree, (V + 31) / 32, marks);
    // every block below sq_block_start: the rows' segments, then the chunks
    hipLaunchKernelGGL(k_neighbor_count, dim3(st.sq_block_start), dim3(kRowThreads), 0, s, rowptr,
                       col, order, st, chunks, marks, nmv);
}

void launch_modularity_finish(const SegTable &st, double M, double *partials, double *q,
                              hipStream_t s) {
    double *partials2 = partials + 2 * (int64_t)st.total_blocks;
    hipLaunchKernelGGL(k_modularity_reduce, dim3(kQReduceBlocks), dim3(256), 0, s, partials,
                       st.total_blocks, st.sq_block_start, st.total_blocks - st.sq_block_start,
                       partials2);
    hipLaunchKernelGGL(k_modularity_final, dim3(1), dim3(256), 0, s, partials2, M, q, nullptr);
}

void launch_modularity(const int32_t *rowptr, const int32_t *col, const float *w,
                       const int32_t *comm, const double *sigma, int32_t V, const int32_t *order,
                       const SegTable &st, const Chunk *chunks, double M, double *partials,
                       double *q, hipStream_t s) {
    launch_modularity_rows(rowptr, col, w, comm, sigma, V, order, st, chunks, partials, s);
    launch_modularity_finish(st, M, partials, q, s);
}

void launch_xcd_splits(const int32_t *rowptr, const int32_t *col, const int32_t *list, int32_t n,
                       int32_t V, int32_t *split, int32_t *unsorted, hipStream_t s) {
    if (n <= 0) return;
    const int32_t blk = (V + 7) / 8;
    hipLaunchKernelGGL(k_xcd_splits, dim3(blocks_for((int64_t)n * 9, 256)), dim3(256), 0, s, rowptr,
                       col, list, n, blk, split);
    hipLaunchKernelGGL(k_rows_sorted, dim3(n), dim3(256), 0, s, rowptr, col, list, n, unsorted);
}

void launch_modularity_xcd(const int32_t *col, const float *w, const int32_t *comm,
                           const int32_t *list, const SegTable &xst, const int32_t *split,
                           double *partials, int32_t base, hipStream_t s) {
    if (xst.total_blocks <= 0) return;
    hipLaunchKernelGGL(k_modularity_xcd, dim3(8 * xst.total_blocks), dim3(kRowThreads), 0, s, col,
                       w, comm, list, xst, split, partials, base);
}

void launch_modularity_reduce(int32_t nb, int32_t sq_lo, int32_t nb_sq, double M,
                              double *partials, double *q, hipStream_t s) {
    double *partials2 = partials + 2 * (int64_t)nb;
    hipLaunchKernelGGL(k_modularity_reduce, dim3(kQReduceBlocks), dim3(256), 0, s, partials, nb,
                       sq_lo, nb_sq, partials2);
    hipLaunchKernelGGL(k_modularity_final, dim3(1), dim3(256), 0, s, partials2, M, q, nullptr);
}

void launch_selfloop_weight(const int32_t *rowptr, const int32_t *col, const float *w, int32_t V,
                            double *partials, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_selfloop_sorted, dim3(kSelfLoopBlocks), dim3(256), 0, s, rowptr, col, w,
                       V, partials);
    hipLaunchKernelGGL(k_selfloop_final, dim3(1), dim3(256), 0, s, partials,
                       (int32_t)kSelfLoopBlocks, out);
}

void launch_modularity_from_own(const double *sigma, int32_t V, double M, double *partials,
                                double *q, double *own_slots, hipStream_t s) {
    const int32_t nb = (int32_t)(V > 0 ? blocks_for(V, kSqItems) : 1);
    hipLaunchKernelGGL(k_sum_squares, dim3(nb), dim3(kRowThreads), 0, s, sigma, V, partials);
    double *partials2 = partials + 2 * (int64_t)nb;
    hipLaunchKernelGGL(k_modularity_reduce, dim3(kQReduceBlocks), dim3(256), 0, s, partials, nb, 0,
                       nb, partials2);
    hipLaunchKernelGGL(k_modularity_final, dim3(1), dim3(256), 0, s, partials2, M, q, own_slots);
}

void launch_compose(int32_t *orig2comm, int32_t V0, const int32_t *map, hipStream_t s) {
    if (V0 <= 0) return;
    hipLaunchKernelGGL(k_compose, dim3(blocks_for(V0, 256)), dim3(256), 0, s, orig2comm, V0, map);
}

namespace {
template <int TIER>
void set_tier_lds_limits(int bytes) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_move_wg<128, TIER>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_move_wg<256, TIER>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_move_wg<512, TIER>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_move_wg<1024, TIER>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_move_wg<128, TIER, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_move_wg<256, TIER, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_move_wg<512, TIER, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_move_wg<1024, TIER, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}
}  // namespace

void set_wg_lds_limits(int bytes) {
    set_tier_lds_limits<0>(bytes);
    set_tier_lds_limits<1>(bytes);
    set_tier_lds_limits<2>(bytes);
    set_tier_lds_limits<3>(bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_move_wg<1024, 4>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace lv
