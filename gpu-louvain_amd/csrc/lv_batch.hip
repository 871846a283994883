// Kernels of the batched (block-diagonal union) path.
#include "lv_batch.h"

namespace lv {

namespace {

inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

__device__ __forceinline__ double row_own(const int32_t *__restrict__ col,
                                          const float *__restrict__ w,
                                          const int32_t *__restrict__ comm, int32_t cv,
                                          int64_t lo, int64_t hi, int st) {
    double own = 0.0;
    for (int64_t e = lo; e < hi; e += st)
        if (comm[col[e]] == cv) own += (double)w[e];
    return own;
}

__device__ __forceinline__ double block_sum(double x, double *lds) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    __syncthreads();
    if (lane == 0) lds[wid] = x;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += lds[i];
    return t;
}

__global__ __launch_bounds__(kRowThreads) void k_vertex_own(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const float *__restrict__ w, const int32_t *__restrict__ comm,
    const int32_t *__restrict__ order, SegTable st, const Chunk *__restrict__ chunks,
    const uint8_t *__restrict__ active, double *own) {
    __shared__ double lds[kRowThreads / 64];
    const int b = blockIdx.x;
    if (b >= st.chunk_block_start) {
        const Chunk ch = chunks[b - st.chunk_block_start];
        if (active && !active[ch.v]) return;
        double x = row_own(col, w, comm, comm[ch.v], ch.lo + threadIdx.x, ch.hi, kRowThreads);
        x = block_sum(x, lds);
        if (threadIdx.x == 0 && x != 0.0) atomicAdd(&own[ch.v], x);
        return;
    }
    int si = 0;
    while (si + 1 < st.nseg && b >= st.s[si + 1].block_start) ++si;
    const Seg sg = st.s[si];
    const int lg = sg.lanes_log2;
    if (lg <= 6) {
        const int X = 1 << lg;
        const int32_t idx = (b - sg.block_start) * (kRowThreads >> lg) + (threadIdx.x >> lg);
        const int lane = threadIdx.x & (X - 1);
        double x = 0.0;
        int32_t v = -1;
        if (idx < sg.count) {
            v = order[sg.list_off + idx];
            x = row_own(col, w, comm, comm[v], rowptr[v] + lane, rowptr[v + 1], X);
        }
        for (int off = X / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0 && v >= 0) own[v] = x;
    } else {
        const int32_t v = order[sg.list_off + (b - sg.block_start)];
        double x = row_own(col, w, comm, comm[v], rowptr[v] + threadIdx.x, rowptr[v + 1],
                           kRowThreads);
        x = block_sum(x, lds);
        if (threadIdx.x == 0) own[v] = x;
    }
}

__global__ __launch_bounds__(256) void k_graph_sums(const int32_t *__restrict__ vptr,
                                                    const double *__restrict__ own,
                                                    const double *__restrict__ sigma,
                                                    const uint8_t *__restrict__ hist,
                                                    const uint8_t *__restrict__ gact,
                                                    double *gin, double *out) {
    __shared__ double lds[4];
    const int32_t g = blockIdx.x;
    if (gact && !gact[g]) return;
    const int32_t lo = vptr[g], hi = vptr[g + 1];
    double o = 0.0, q = 0.0, m = 0.0;
    for (int32_t v = lo + threadIdx.x; v < hi; v += 256) {
        if (own) o += own[v];
        const double x = sigma[v];
        q += x * x;
        m += (double)(hist[v] & 1u);
    }
    if (own) o = block_sum(o, lds);
    q = block_sum(q, lds);
    m = block_sum(m, lds);
    if (threadIdx.x == 0) {
        if (own)
            gin[g] = o;
        else
            o = gin[g];
        out[3 * g] = o;
        out[3 * g + 1] = q;
        out[3 * g + 2] = m;
    }
}

__global__ void k_graph_frontier(const int32_t *__restrict__ rowptr, int32_t V,
                                 const int32_t *__restrict__ gid, const uint8_t *__restrict__ gact,
                                 BucketBounds bb, int8_t *keys, uint8_t *active) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int32_t deg = rowptr[v + 1] - rowptr[v];
    const bool act = deg > 0 && gact[gid[v]];
    active[v] = act ? 1 : 0;
    keys[v] = act ? (int8_t)bucket_of(deg, bb) : (int8_t)-1;
}

__global__ void k_graph_ids(const int32_t *__restrict__ vptr, int32_t *gid) {
    const int32_t g = blockIdx.x;
    for (int32_t v = vptr[g] + threadIdx.x; v < vptr[g + 1]; v += blockDim.x) gid[v] = g;
}

__global__ void k_graph_offsets(const int32_t *__restrict__ vptr, int32_t G, int32_t V,
                                const int32_t *__restrict__ newid, const int32_t *__restrict__ newV,
                                int32_t *vptr_new) {
    const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g > G) return;
    const int32_t p = vptr[g];
    vptr_new[g] = p < V ? newid[p] : *newV;
}

__global__ void k_graph_final(const int32_t *__restrict__ orig2comm, int32_t V0,
                              const int32_t *__restrict__ gid0, const uint8_t *__restrict__ done,
                              const int32_t *__restrict__ vptr, int32_t *final_comm) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= V0) return;
    const int32_t g = gid0[x];
    if (done[g]) final_comm[x] = orig2comm[x] - vptr[g];
}

}  // namespace

void launch_vertex_own(const int32_t *rowptr, const int32_t *col, const float *w,
                       const int32_t *comm, const int32_t *order, const SegTable &st,
                       const Chunk *chunks, const uint8_t *active, double *own, hipStream_t s) {
    const int32_t nb = st.chunk_block_start + st.nchunks;
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_vertex_own, dim3(nb), dim3(kRowThreads), 0, s, rowptr, col, w, comm, order,
                       st, chunks, active, own);
}

void launch_graph_sums(const int32_t *vptr, int32_t G, const double *own, const double *sigma,
                       const uint8_t *hist, const uint8_t *gact, double *gin, double *out,
                       hipStream_t s) {
    if (G <= 0) return;
    hipLaunchKernelGGL(k_graph_sums, dim3(G), dim3(256), 0, s, vptr, own, sigma, hist, gact, gin,
                       out);
}

void launch_graph_frontier(const int32_t *rowptr, int32_t V, const int32_t *gid,
                           const uint8_t *gact, BucketBounds bb, int8_t *keys, uint8_t *active,
                           hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_graph_frontier, dim3(blocks_for(V, 256)), dim3(256), 0, s, rowptr, V, gid,
                       gact, bb, keys, active);
}

void launch_graph_ids(const int32_t *vptr, int32_t G, int32_t *gid, hipStream_t s) {
    if (G <= 0) return;
    hipLaunchKernelGGL(k_graph_ids, dim3(G), dim3(256), 0, s, vptr, gid);
}

void launch_graph_offsets(const int32_t *vptr, int32_t G, int32_t V, const int32_t *newid,
                          const int32_t *newV, int32_t *vptr_new, hipStream_t s) {
    hipLaunchKernelGGL(k_graph_offsets, dim3(blocks_for((int64_t)G + 1, 256)), dim3(256), 0, s, vptr,
                       G, V, newid, newV, vptr_new);
}

void launch_graph_final(const int32_t *orig2comm, int32_t V0, const int32_t *gid0,
                        const uint8_t *done, const int32_t *vptr, int32_t *final_comm,
                        hipStream_t s) {
    if (V0 <= 0) return;
    hipLaunchKernelGGL(k_graph_final, dim3(blocks_for(V0, 256)), dim3(256), 0, s, orig2comm, V0,
                       gid0, done, vptr, final_comm);
}

}  // namespace lv
