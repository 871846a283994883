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
//   k_heavy_*       rows beyond the LDS tier: edge chunks from many
//                   workgroups insert into a per-vertex global hash; slot
//                   chunks evaluate gains into a packed 64-bit atomic max.
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

__global__ __launch_bounds__(256) void k_heavy_insert(MoveArgs a, const Chunk *__restrict__ ins,
                                                      HeavyTables t) {
    const Chunk ch = ins[blockIdx.x];
    const int64_t off = t.off[ch.hv];
    const int lg = t.cap_log2[ch.hv];
    int32_t *K = t.keys + off;
    double *W = t.vals + off;
    const int32_t v = ch.v;
    for (int64_t e = ch.lo + threadIdx.x; e < ch.hi; e += 256) {
        const int32_t u = a.col[e];
        const float wf = a.w[e];
        if (u == v) continue;
        global_insert(K, W, 32 - lg, (1u << lg) - 1, a.comm[u], (double)wf);
    }
}

__global__ __launch_bounds__(256) void k_heavy_eval(MoveArgs a, const Chunk *__restrict__ ev,
                                                    HeavyTables t) {
    const Chunk ch = ev[blockIdx.x];
    const int32_t v = ch.v;
    const int32_t cur = a.comm[v];
    const float k32 = (float)a.k[v];
    const float scur = (float)a.sigma[cur];
    const int32_t size_cur = a.size[cur];
    uint64_t best = 0;
    for (int64_t i = ch.lo + threadIdx.x; i < ch.hi; i += 256) {
        const int32_t c = t.keys[i];
        if (c < 0) continue;
        const double Kc = t.vals[i];
        t.keys[i] = -1;  // leave the table clean for the next sweep
        t.vals[i] = 0.0;
        if (c == cur) {
            t.kcur[ch.hv] = Kc;
            continue;
        }
        if (!may_move(c, cur, size_cur, a.size)) continue;
        const float g = gain_of((float)Kc, a.M32, a.den, k32, scur, (float)a.sigma[c]);
        const uint64_t key = pack_best(g, c);
        best = key > best ? key : best;
    }
    best = wave_max_u64(best);
    if ((threadIdx.x & 63) == 0 && best != 0)
        atomicMax(reinterpret_cast<unsigned long long *>(&t.best[ch.hv]),
                  (unsigned long long)best);
}

__global__ void k_heavy_final(MoveArgs a, const int32_t *__restrict__ heavy, int32_t nh,
                              HeavyTables t) {
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
        const int32_t cv = comm[ch.v];
        for (int64_t e = ch.lo + threadIdx.x; e < ch.hi; e += kRowThreads) {
            const int32_t u = col[e];
            const float wf = w[e];
            if (comm[u] == cv) own += (double)wf;
        }
    } else {
        const Seg sg = st.s[find_seg(st, b)];
        const int lg = sg.lanes_log2;
        if (lg <= 6) {
            const int X = 1 << lg;
            const int32_t idx = (b - sg.block_start) * (kRowThreads >> lg) + (threadIdx.x >> lg);
            const int lane = threadIdx.x & (X - 1);
            if (idx < sg.count) {
                const int32_t v = order[sg.list_off + idx];
                const int32_t cv = comm[v];
                for (int32_t e = rowptr[v] + lane; e < rowptr[v + 1]; e += X) {
                    const int32_t u = col[e];
                    const float wf = w[e];
                    if (comm[u] == cv) own += (double)wf;
                }
            }
        } else {
            const int32_t v = order[sg.list_off + (b - sg.block_start)];
            const int32_t cv = comm[v];
            for (int32_t e = rowptr[v] + threadIdx.x; e < rowptr[v + 1]; e += kRowThreads) {
                const int32_t u = col[e];
                const float wf = w[e];
                if (comm[u] == cv) own += (double)wf;
            }
        }
    }
    own = block_sum_f64(own, lds);
    sq = block_sum_f64(sq, lds);
    if (threadIdx.x == 0) {
        partials[2 * b] = own;
        partials[2 * b + 1] = sq;
    }
}

__global__ __launch_bounds__(1024) void k_modularity_final(const double *__restrict__ partials,
                                                           int32_t nb, double M, double *q) {
    __shared__ double lds[16];
    double own = 0.0, sq = 0.0;
    for (int32_t i = threadIdx.x; i < nb; i += 1024) {
        own += partials[2 * i];
        sq += partials[2 * i + 1];
    }
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

int move_wg_max_cap_log2(int max_lds_bytes) {
    int lg = 4;
    while (((size_t)12 << (lg + 1)) + 1024 <= (size_t)max_lds_bytes) ++lg;
    return lg;
}

void launch_move_wg(int threads, int cap_log2, const MoveArgs &a, const int32_t *list, int32_t n,
                    hipStream_t s) {
    if (n <= 0) return;
    const size_t lds = (size_t)12 << cap_log2;
    const unsigned g = (unsigned)(n < 1048576 ? n : 1048576);
    if (threads <= 256)
        hipLaunchKernelGGL(k_move_wg<256>, dim3(g), dim3(256), lds, s, a, list, n, cap_log2);
    else if (threads <= 512)
        hipLaunchKernelGGL(k_move_wg<512>, dim3(g), dim3(512), lds, s, a, list, n, cap_log2);
    else
        hipLaunchKernelGGL(k_move_wg<1024>, dim3(g), dim3(1024), lds, s, a, list, n, cap_log2);
}

void launch_move_heavy(const MoveArgs &a, const int32_t *heavy_list, int32_t nh,
                       const Chunk *ins, int32_t nins, const Chunk *eval, int32_t neval,
                       const HeavyTables &t, hipStream_t s) {
    if (nh <= 0) return;
    hipLaunchKernelGGL(k_heavy_insert, dim3(nins), dim3(256), 0, s, a, ins, t);
    hipLaunchKernelGGL(k_heavy_eval, dim3(neval), dim3(256), 0, s, a, eval, t);
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
    hipLaunchKernelGGL(k_modularity_final, dim3(1), dim3(1024), 0, s, partials, st.total_blocks, M,
                       q);
}

void launch_compose(int32_t *orig2comm, int32_t V0, const int32_t *map, hipStream_t s) {
    if (V0 <= 0) return;
    hipLaunchKernelGGL(k_compose, dim3(blocks_for(V0, 256)), dim3(256), 0, s, orig2comm, V0, map);
}

void set_wg_lds_limits(int bytes) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_move_wg<256>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_move_wg<512>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_move_wg<1024>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace lv
