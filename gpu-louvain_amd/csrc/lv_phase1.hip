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
//   k_move_sub<X>   X = 4,8,16,32 lanes per vertex, one edge per lane;
//                   per-community sums by an in-register all-pairs pass over
//                   the X lanes (row order, so sums are bitwise the
//                   sequential ones), argmax by xor-shuffle of packed keys.
//   k_move_wave<C>  one wavefront per vertex, C-slot LDS hash per wave.
//   k_move_wg<T>    one T-lane workgroup per vertex, LDS hash sized per launch.
//   k_chunk_*       rows beyond one workgroup: stage A aggregates 2048-edge
//                   chunks in LDS and buckets the distinct pairs by hash
//                   partition; stage B merges each (vertex, partition) in
//                   LDS, evaluates gains and combines partitions through a
//                   packed 64-bit atomic max.
#include "lv_phase1.h"

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

// ------------------------------------------------------------------ moves

template <int X>
__global__ __launch_bounds__(256) void k_move_sub(MoveArgs a, const int32_t *__restrict__ list,
                                                  int32_t n) {
    const int lane = threadIdx.x & (X - 1);
    const int32_t slot = (int32_t)((blockIdx.x * 256u + threadIdx.x) / X);
    if (slot >= n) return;  // whole X-lane segment leaves together
    const int32_t v = list[slot];
    const int32_t rs = a.rowptr[v];
    const int32_t deg = a.rowptr[v + 1] - rs;  // <= X by bucket
    const int32_t cur = a.comm[v];
    int32_t c = -1;
    double wv = 0.0;
    if (lane < deg) {
        const int32_t u = a.col[rs + lane];
        const float wf = a.w[rs + lane];
        if (u != v) {
            c = a.comm[u];
            wv = (double)wf;
        }
    }
    // All-pairs pass in row order: K = sum of weights of lanes sharing c;
    // the lowest such lane owns the community.
    double K = 0.0;
    bool owner = c >= 0;
    for (int j = 0; j < deg; ++j) {
        const int32_t cj = __shfl(c, j, X);
        const double wj = __shfl(wv, j, X);
        if (cj == c) {
            K += wj;
            if (j < lane) owner = false;
        }
    }
    double kc = (owner && c == cur) ? K : 0.0;
#pragma unroll
    for (int off = X / 2; off > 0; off >>= 1) kc += __shfl_xor(kc, off, X);

    uint64_t best = pack_best(0.0f, cur);
    if (owner && c != cur) {
        const int32_t size_cur = a.size[cur];
        if (may_move(c, cur, size_cur, a.size)) {
            const float g = gain_of((float)K, a.M32, a.den, (float)a.k[v], (float)a.sigma[cur],
                                    (float)a.sigma[c]);
            const uint64_t key = pack_best(g, c);
            best = key > best ? key : best;
        }
    }
#pragma unroll
    for (int off = X / 2; off > 0; off >>= 1) {
        const uint64_t o = __shfl_xor(best, off, X);
        best = o > best ? o : best;
    }
    if (lane == 0) a.newc[v] = decide_from(best, kc, a.M32, cur);
}

template <int C>
__global__ __launch_bounds__(256) void k_move_wave(MoveArgs a, const int32_t *__restrict__ list,
                                                   int32_t n) {
    constexpr int LOG2C = ilog2_c(C);
    __shared__ int32_t keys_s[4 * C];
    __shared__ double vals_s[4 * C];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int32_t slot = blockIdx.x * 4 + wid;
    const bool active = slot < n;
    int32_t *K = keys_s + wid * C;
    double *W = vals_s + wid * C;
    for (int i = lane; i < C; i += 64) {
        K[i] = -1;
        W[i] = 0.0;
    }
    __syncthreads();
    int32_t v = 0, cur = 0;
    if (active) {
        v = list[slot];
        const int32_t rs = a.rowptr[v], re = a.rowptr[v + 1];
        cur = a.comm[v];
        for (int32_t e = rs + lane; e < re; e += 64) {
            const int32_t u = a.col[e];
            const float wf = a.w[e];
            if (u == v) continue;
            lds_insert(K, W, 32 - LOG2C, C - 1, a.comm[u], (double)wf);
        }
    }
    __syncthreads();
    if (!active) return;  // only wave-level communication below
    const float k32 = (float)a.k[v];
    const float scur = (float)a.sigma[cur];
    const int32_t size_cur = a.size[cur];
    uint64_t best = pack_best(0.0f, cur);
    double kc = 0.0;
    for (int i = lane; i < C; i += 64) {
        const int32_t c = K[i];
        if (c < 0) continue;
        const double Kc = W[i];
        if (c == cur) {
            kc = Kc;
            continue;
        }
        if (!may_move(c, cur, size_cur, a.size)) continue;
        const float g = gain_of((float)Kc, a.M32, a.den, k32, scur, (float)a.sigma[c]);
        const uint64_t key = pack_best(g, c);
        best = key > best ? key : best;
    }
    best = wave_max_u64(best);
    kc = wave_sum_f64(kc);
    if (lane == 0) a.newc[v] = decide_from(best, kc, a.M32, cur);
}

template <int T>
__global__ __launch_bounds__(T) void k_move_wg(MoveArgs a, const int32_t *__restrict__ list,
                                               int32_t n, int cap_log2) {
    extern __shared__ double dyn_lds[];
    const int C = 1 << cap_log2;
    double *W = dyn_lds;
    int32_t *K = reinterpret_cast<int32_t *>(dyn_lds + C);
    __shared__ uint64_t red_best[T / 64];
    __shared__ double red_kc[T / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint32_t shift = 32 - cap_log2, mask = (uint32_t)C - 1;
    for (int32_t vi = blockIdx.x; vi < n; vi += gridDim.x) {
        for (int i = threadIdx.x; i < C; i += T) {
            K[i] = -1;
            W[i] = 0.0;
        }
        __syncthreads();
        const int32_t v = list[vi];
        const int32_t rs = a.rowptr[v], re = a.rowptr[v + 1];
        const int32_t cur = a.comm[v];
        for (int32_t e = rs + threadIdx.x; e < re; e += T) {
            const int32_t u = a.col[e];
            const float wf = a.w[e];
            if (u == v) continue;
            lds_insert(K, W, shift, mask, a.comm[u], (double)wf);
        }
        __syncthreads();
        const float k32 = (float)a.k[v];
        const float scur = (float)a.sigma[cur];
        const int32_t size_cur = a.size[cur];
        uint64_t best = pack_best(0.0f, cur);
        double kc = 0.0;
        for (int i = threadIdx.x; i < C; i += T) {
            const int32_t c = K[i];
            if (c < 0) continue;
            const double Kc = W[i];
            if (c == cur) {
                kc = Kc;
                continue;
            }
            if (!may_move(c, cur, size_cur, a.size)) continue;
            const float g = gain_of((float)Kc, a.M32, a.den, k32, scur, (float)a.sigma[c]);
            const uint64_t key = pack_best(g, c);
            best = key > best ? key : best;
        }
        best = wave_max_u64(best);
        kc = wave_sum_f64(kc);
        if (lane == 0) {
            red_best[wid] = best;
            red_kc[wid] = kc;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t b = red_best[0];
            double kk = red_kc[0];
            for (int w = 1; w < T / 64; ++w) {
                b = red_best[w] > b ? red_best[w] : b;
                kk += red_kc[w];
            }
            a.newc[v] = decide_from(b, kk, a.M32, cur);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------- chunked tier

// Stage A: aggregate one chunk of a heavy row in LDS; write its distinct
// (community, weight) pairs grouped by partition.
__global__ __launch_bounds__(256) void k_chunk_aggregate(MoveArgs a, ChunkedTier t) {
    constexpr int C = 1 << kChunkCapLog2;
    __shared__ int32_t keys[C];
    __shared__ double vals[C];
    __shared__ int32_t pcnt[kMaxParts];
    const Chunk ch = t.chunks[blockIdx.x];
    const int P = ch.P;
    for (int i = threadIdx.x; i < C; i += 256) {
        keys[i] = -1;
        vals[i] = 0.0;
    }
    for (int p = threadIdx.x; p < P; p += 256) pcnt[p] = 0;
    __syncthreads();
    const int32_t v = ch.v;
    for (int64_t e = ch.lo + threadIdx.x; e < ch.hi; e += 256) {
        const int32_t u = a.col[e];
        const float wf = a.w[e];
        if (u == v) continue;
        lds_insert(keys, vals, 32 - kChunkCapLog2, C - 1, a.comm[u], (double)wf);
    }
    __syncthreads();
    for (int s = threadIdx.x; s < C; s += 256) {
        const int32_t k = keys[s];
        if (k >= 0) atomicAdd(&pcnt[part_of(k, (uint32_t)P)], 1);
    }
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
    for (int s = threadIdx.x; s < C; s += 256) {
        const int32_t k = keys[s];
        if (k < 0) continue;
        const int32_t pos = atomicAdd(&pcnt[part_of(k, (uint32_t)P)], 1);
        t.out_key[ch.out + pos] = k;
        t.out_val[ch.out + pos] = vals[s];
    }
}

// Stage B: merge partition p of a heavy vertex and evaluate its candidates.
// A partition holding more pairs than an LDS table takes falls back to the
// vertex's global table (always large enough), so every insert terminates.
constexpr int kMergeThreads = 512;

__global__ __launch_bounds__(kMergeThreads) void k_chunk_merge(MoveArgs a, ChunkedTier t) {
    constexpr int C = 1 << kMergeCapLog2;
    extern __shared__ double merge_lds[];
    double *vals = merge_lds;
    int32_t *keys = reinterpret_cast<int32_t *>(merge_lds + C);
    __shared__ uint64_t red_best[kMergeThreads / 64];
    __shared__ double red_kc[kMergeThreads / 64];
    const MergeItem it = t.items[blockIdx.x];
    const HeavyVertex h = t.hvs[it.hv];
    const int32_t v = h.v;
    const int32_t cur = a.comm[v];
    const float k32 = (float)a.k[v];
    const float scur = (float)a.sigma[cur];
    const int32_t size_cur = a.size[cur];
    const int32_t n = t.ptot[h.ptot_base + it.p];
    uint64_t best = 0;
    double kc = 0.0;
    if (n <= t.merge_limit) {
        for (int i = threadIdx.x; i < C; i += kMergeThreads) {
            keys[i] = -1;
            vals[i] = 0.0;
        }
        __syncthreads();
        for (int32_t c = h.chunk0; c < h.chunk0 + h.nchunks; ++c) {
            const Chunk ch = t.chunks[c];
            const int32_t lo = t.poff[ch.pbase + it.p], hi = t.poff[ch.pbase + it.p + 1];
            for (int32_t i = lo + threadIdx.x; i < hi; i += kMergeThreads)
                lds_insert(keys, vals, 32 - kMergeCapLog2, C - 1, t.out_key[ch.out + i],
                           t.out_val[ch.out + i]);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < C; i += kMergeThreads) {
            const int32_t c = keys[i];
            if (c < 0) continue;
            const double Kc = vals[i];
            if (c == cur) {
                kc = Kc;
                continue;
            }
            if (!may_move(c, cur, size_cur, a.size)) continue;
            const float g = gain_of((float)Kc, a.M32, a.den, k32, scur, (float)a.sigma[c]);
            const uint64_t key = pack_best(g, c);
            best = key > best ? key : best;
        }
    } else {
        int32_t *K = t.gkeys + h.tab_off;
        double *W = t.gvals + h.tab_off;
        const uint32_t cap = 1u << h.tab_lg;
        for (int32_t c = h.chunk0; c < h.chunk0 + h.nchunks; ++c) {
            const Chunk ch = t.chunks[c];
            const int32_t lo = t.poff[ch.pbase + it.p], hi = t.poff[ch.pbase + it.p + 1];
            for (int32_t i = lo + threadIdx.x; i < hi; i += kMergeThreads)
                global_insert(K, W, 32 - h.tab_lg, cap - 1, t.out_key[ch.out + i],
                              t.out_val[ch.out + i]);
        }
        __threadfence();
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < cap; i += kMergeThreads) {
            const int32_t c = __hip_atomic_load(&K[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (c < 0 || part_of(c, (uint32_t)h.P) != (uint32_t)it.p) continue;
            const double Kc = __hip_atomic_load(&W[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            K[i] = -1;  // leave the table clean
            W[i] = 0.0;
            if (c == cur) {
                kc = Kc;
                continue;
            }
            if (!may_move(c, cur, size_cur, a.size)) continue;
            const float g = gain_of((float)Kc, a.M32, a.den, k32, scur, (float)a.sigma[c]);
            const uint64_t key = pack_best(g, c);
            best = key > best ? key : best;
        }
    }
    best = wave_max_u64(best);
    kc = wave_sum_f64(kc);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red_best[wid] = best;
        red_kc[wid] = kc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t b = red_best[0];
        double kk = red_kc[0];
        for (int w = 1; w < kMergeThreads / 64; ++w) {
            b = red_best[w] > b ? red_best[w] : b;
            kk += red_kc[w];
        }
        if (b != 0)
            atomicMax(reinterpret_cast<unsigned long long *>(&t.best[it.hv]),
                      (unsigned long long)b);
        if (part_of(cur, (uint32_t)h.P) == (uint32_t)it.p) t.kcur[it.hv] = kk;
        t.ptot[h.ptot_base + it.p] = 0;  // ready for the next sweep
    }
}

__global__ void k_heavy_final(MoveArgs a, const int32_t *__restrict__ heavy, int32_t nh,
                              ChunkedTier t) {
    const int32_t hv = blockIdx.x * blockDim.x + threadIdx.x;
    if (hv >= nh) return;
    const int32_t v = heavy[hv];
    const int32_t cur = a.comm[v];
    uint64_t best = t.best[hv];
    const uint64_t init = pack_best(0.0f, cur);
    best = best > init ? best : init;
    a.newc[v] = decide_from(best, t.kcur[hv], a.M32, cur);
    t.best[hv] = 0;
    t.kcur[hv] = 0.0;
}

__global__ void k_clear_tables(int32_t *keys, double *vals, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        keys[i] = -1;
        vals[i] = 0.0;
    }
}

// ------------------------------------------------------------------ apply
// updateVertexCommunity (modularity_optimisation.cu:237-249) fused with the
// community-total update (replaces the full rebuild at :328-330 by exact
// FP64 increments).
__global__ __launch_bounds__(256) void k_apply(const int32_t *__restrict__ list, int32_t n,
                                               const int32_t *__restrict__ newc, int32_t *comm,
                                               int32_t *size, double *sigma,
                                               const double *__restrict__ k,
                                               unsigned long long *moved) {
    const int32_t i = blockIdx.x * 256 + threadIdx.x;
    int mv = 0;
    if (i < n) {
        const int32_t v = list[i];
        const int32_t nc = newc[v], oc = comm[v];
        if (nc != oc) {
            comm[v] = nc;
            atomicSub(&size[oc], 1);
            atomicAdd(&size[nc], 1);
            const double kv = k[v];
            atomicAdd(&sigma[oc], -kv);
            atomicAdd(&sigma[nc], kv);
            mv = 1;
        }
    }
    const uint64_t m = __ballot(mv);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(moved, (unsigned long long)__popcll(m));
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
__global__ __launch_bounds__(kRowThreads) void k_strength(const int32_t *__restrict__ rowptr,
                                                          const float *__restrict__ w,
                                                          const int32_t *__restrict__ order,
                                                          SegTable st,
                                                          const Chunk *__restrict__ chunks,
                                                          double *k) {
    __shared__ double lds[kRowThreads / 64];
    const int b = blockIdx.x;
    if (b >= st.chunk_block_start) {
        const Chunk ch = chunks[b - st.chunk_block_start];
        double s = 0.0;
        for (int64_t e = ch.lo + threadIdx.x; e < ch.hi; e += kRowThreads) s += (double)w[e];
        s = block_sum_f64(s, lds);
        if (threadIdx.x == 0) atomicAdd(&k[ch.v], s);
        return;
    }
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
            for (int32_t e = rowptr[v] + lane; e < rowptr[v + 1]; e += X) s += (double)w[e];
        }
        for (int off = X / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0 && v >= 0) k[v] = s;
    } else {
        const int32_t v = order[sg.list_off + (b - sg.block_start)];
        double s = 0.0;
        for (int32_t e = rowptr[v] + threadIdx.x; e < rowptr[v + 1]; e += kRowThreads)
            s += (double)w[e];
        s = block_sum_f64(s, lds);
        if (threadIdx.x == 0) k[v] = s;
    }
}

__global__ void k_level_state(int32_t V, const double *__restrict__ k, int32_t *comm,
                              int32_t *size, double *sigma) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) {
        comm[v] = v;
        size[v] = 1;
        sigma[v] = k[v];
    }
}

// Fused modularity (calculateToOwnCommunity + the two reductions,
// modularity_optimisation.cu:365-401): one launch produces per-block
// partials of sum_v (weight to own community) and sum_c S_c^2; a one-block
// pass adds them in a fixed order (run-to-run deterministic).
__global__ __launch_bounds__(kRowThreads) void k_modularity(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const float *__restrict__ w, const int32_t *__restrict__ comm,
    const double *__restrict__ sigma, int32_t V, const int32_t *__restrict__ order, SegTable st,
    const Chunk *__restrict__ chunks, double *partials) {
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
        own = own_weight(col, w, comm, comm[ch.v], ch.lo + threadIdx.x, ch.hi, kRowThreads);
    } else {
        const Seg sg = st.s[find_seg(st, b)];
        const int lg = sg.lanes_log2;
        if (lg <= 6) {
            const int X = 1 << lg;
            const int32_t idx = (b - sg.block_start) * (kRowThreads >> lg) + (threadIdx.x >> lg);
            const int lane = threadIdx.x & (X - 1);
            if (idx < sg.count) {
                const int32_t v = order[sg.list_off + idx];
                own = own_weight(col, w, comm, comm[v], rowptr[v] + lane, rowptr[v + 1], X);
            }
        } else {
            const int32_t v = order[sg.list_off + (b - sg.block_start)];
            own = own_weight(col, w, comm, comm[v], rowptr[v] + threadIdx.x, rowptr[v + 1],
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

// Two-level fixed-order reduction of the per-block partials: 256 blocks
// each add a contiguous range, then one block adds the 256 results.
constexpr int kQReduceBlocks = 256;

__global__ __launch_bounds__(256) void k_modularity_reduce(const double *__restrict__ partials,
                                                           int32_t nb, double *partials2) {
    __shared__ double lds[4];
    const int64_t lo = (int64_t)nb * blockIdx.x / kQReduceBlocks;
    const int64_t hi = (int64_t)nb * (blockIdx.x + 1) / kQReduceBlocks;
    double own = 0.0, sq = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) {
        own += partials[2 * i];
        sq += partials[2 * i + 1];
    }
    own = block_sum_f64(own, lds);
    sq = block_sum_f64(sq, lds);
    if (threadIdx.x == 0) {
        partials2[2 * blockIdx.x] = own;
        partials2[2 * blockIdx.x + 1] = sq;
    }
}

__global__ __launch_bounds__(256) void k_modularity_final(const double *__restrict__ partials2,
                                                          double M, double *q) {
    __shared__ double lds[4];
    double own = partials2[2 * threadIdx.x];
    double sq = partials2[2 * threadIdx.x + 1];
    own = block_sum_f64(own, lds);
    sq = block_sum_f64(sq, lds);
    if (threadIdx.x == 0) {
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
                     const SegTable &st, const Chunk *chunks, double *k, hipStream_t s) {
    if (st.chunk_block_start + st.nchunks <= 0) return;
    hipLaunchKernelGGL(k_strength, dim3(st.chunk_block_start + st.nchunks), dim3(kRowThreads), 0,
                       s, rowptr, w, order, st, chunks, k);
}

void launch_level_state(int32_t V, const double *k, int32_t *comm, int32_t *size, double *sigma,
                        hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_level_state, dim3(blocks_for(V, 256)), dim3(256), 0, s, V, k, comm, size,
                       sigma);
}

void launch_move_sub(int X, const MoveArgs &a, const int32_t *list, int32_t n, hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = blocks_for((int64_t)n * X, 256);
    switch (X) {
    case 4: hipLaunchKernelGGL(k_move_sub<4>, dim3(g), dim3(256), 0, s, a, list, n); break;
    case 8: hipLaunchKernelGGL(k_move_sub<8>, dim3(g), dim3(256), 0, s, a, list, n); break;
    case 16: hipLaunchKernelGGL(k_move_sub<16>, dim3(g), dim3(256), 0, s, a, list, n); break;
    default: hipLaunchKernelGGL(k_move_sub<32>, dim3(g), dim3(256), 0, s, a, list, n); break;
    }
}

void launch_move_wave(int cap, const MoveArgs &a, const int32_t *list, int32_t n, hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = blocks_for(n, 4);
    if (cap <= 256)
        hipLaunchKernelGGL(k_move_wave<256>, dim3(g), dim3(256), 0, s, a, list, n);
    else
        hipLaunchKernelGGL(k_move_wave<512>, dim3(g), dim3(256), 0, s, a, list, n);
}

void launch_move_wg(int threads, int cap_log2, const MoveArgs &a, const int32_t *list, int32_t n,
                    hipStream_t s) {
    (void)threads;
    if (n <= 0) return;
    const size_t lds = (size_t)12 << cap_log2;
    const unsigned g = (unsigned)(n < 1048576 ? n : 1048576);
    hipLaunchKernelGGL(k_move_wg<256>, dim3(g), dim3(256), lds, s, a, list, n, cap_log2);
}

void launch_move_chunked(const MoveArgs &a, const ChunkedTier &t, int32_t nchunks,
                         int32_t nitems, const int32_t *heavy_list, int32_t nh, hipStream_t s) {
    if (nh <= 0) return;
    hipLaunchKernelGGL(k_chunk_aggregate, dim3(nchunks), dim3(256), 0, s, a, t);
    hipLaunchKernelGGL(k_chunk_merge, dim3(nitems), dim3(kMergeThreads),
                       (size_t)12 << kMergeCapLog2, s, a, t);
    hipLaunchKernelGGL(k_heavy_final, dim3(blocks_for(nh, 256)), dim3(256), 0, s, a, heavy_list,
                       nh, t);
}

void launch_clear_tables(int32_t *keys, double *vals, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_clear_tables, dim3((unsigned)g), dim3(256), 0, s, keys, vals, n);
}

void launch_apply(const int32_t *list, int32_t n, const int32_t *newc, int32_t *comm,
                  int32_t *size, double *sigma, const double *k, unsigned long long *moved,
                  hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_apply, dim3(blocks_for(n, 256)), dim3(256), 0, s, list, n, newc, comm,
                       size, sigma, k, moved);
}

void launch_modularity(const int32_t *rowptr, const int32_t *col, const float *w,
                       const int32_t *comm, const double *sigma, int32_t V, const int32_t *order,
                       const SegTable &st, const Chunk *chunks, double M, double *partials,
                       double *q, hipStream_t s) {
    hipLaunchKernelGGL(k_modularity, dim3(st.total_blocks), dim3(kRowThreads), 0, s, rowptr, col,
                       w, comm, sigma, V, order, st, chunks, partials);
    double *partials2 = partials + 2 * (int64_t)st.total_blocks;
    hipLaunchKernelGGL(k_modularity_reduce, dim3(kQReduceBlocks), dim3(256), 0, s, partials,
                       st.total_blocks, partials2);
    hipLaunchKernelGGL(k_modularity_final, dim3(1), dim3(256), 0, s, partials2, M, q);
}

void launch_compose(int32_t *orig2comm, int32_t V0, const int32_t *map, hipStream_t s) {
    if (V0 <= 0) return;
    hipLaunchKernelGGL(k_compose, dim3(blocks_for(V0, 256)), dim3(256), 0, s, orig2comm, V0, map);
}

void set_wg_lds_limits(int bytes) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_move_wg<256>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_chunk_merge),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace lv
