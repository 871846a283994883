// Sort-based community aggregation kernels for gfx950.
#include "lv_aggsort.h"
#include "lv_prims.h"

#include <utility>

namespace lv {

namespace {

inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

// ------------------------------------------------------------ radix sort

// Digit of a key whose high word is moved down to bit lo_bits.
__device__ __forceinline__ int radix_digit(uint64_t k, int shift, int lo_bits) {
    const uint64_t vk = ((k >> 32) << lo_bits) | (k & 0xffffffffull);
    return (int)((vk >> shift) & 255u);
}

// Per tile: digit counts, written digit-major (hist[d * ntiles + tile]) so
// that one exclusive scan gives every (digit, tile) its output offset.
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint64_t *__restrict__ keys,
                                                             int64_t n, int shift, int lo_bits,
                                                             int64_t ntiles, int32_t *hist) {
    __shared__ int32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int j = 0; j < kSortItems; ++j) {
        const int64_t i = base + (int64_t)j * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&cnt[radix_digit(keys[i], shift, lo_bits)], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Stable scatter of one tile.  Wave w owns items [w*512, w*512+512) of the
// tile, taken as 8 rounds of 64 consecutive items; a lane's rank among the
// equal digits of its round comes from 8 ballots (one per digit bit), and
// per-wave digit counters carry the ranks across rounds.  The tile is then
// reordered by digit in LDS, so that the stores of consecutive lanes go to
// consecutive addresses of each digit's output range (offset by the scanned
// histogram).
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint64_t *__restrict__ keys, const float *__restrict__ vals, int64_t n, int shift,
    int lo_bits, int64_t ntiles, const int32_t *__restrict__ offs, uint64_t *okeys,
    float *ovals) {
    constexpr int NW = kSortThreads / 64;
    __shared__ int32_t wcnt[NW][256];
    __shared__ int32_t tstart[256];     // first tile-local slot of each digit
    __shared__ int32_t gbase[256];      // output offset of each digit minus tstart
    __shared__ uint64_t lk[kSortTile];
    __shared__ float lv[kSortTile];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int d = lane; d < 256; d += 64) wcnt[wid][d] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int64_t tile0 = (int64_t)blockIdx.x * kSortTile;
    const int64_t base = tile0 + (int64_t)wid * (kSortItems * 64);
    uint64_t k[kSortItems];
    float v[kSortItems];
    int dg[kSortItems], rk[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const int64_t i = base + (int64_t)j * 64 + lane;
        const bool ok = i < n;
        k[j] = ok ? keys[i] : 0ull;
        v[j] = ok ? vals[i] : 0.0f;
        const int d = ok ? radix_digit(k[j], shift, lo_bits) : -1;
        dg[j] = d;
        // lanes of this round with the same digit
        uint64_t same = __ballot(ok);
        if (!ok) same = 0;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bit = __ballot(ok && ((d >> b) & 1));
            same &= ((d >> b) & 1) ? bit : ~bit;
        }
        int r = 0;
        if (ok) r = wcnt[wid][d] + __popcll(same & lt);
        rk[j] = r;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // the lowest lane of each digit group advances the counter
        if (ok && (same & lt) == 0) wcnt[wid][d] += __popcll(same);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __syncthreads();
    // per digit: tile count, tile-local start (block exclusive scan), and
    // the wave offsets within the digit (counts of the earlier waves)
    {
        const int d = threadIdx.x;  // 256 threads = 256 digits
        int32_t c = 0;
        for (int w2 = 0; w2 < NW; ++w2) {
            const int32_t x = wcnt[w2][d];
            wcnt[w2][d] = c;  // becomes the wave offset
            c += x;
        }
        // exclusive scan of c over the 256 digits
        int32_t inc = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int32_t y = __shfl_up(inc, off, 64);
            if (lane >= off) inc += y;
        }
        __shared__ int32_t wsum[NW];
        if (lane == 63) wsum[wid] = inc;
        __syncthreads();
        int32_t pre = 0;
        for (int w2 = 0; w2 < wid; ++w2) pre += wsum[w2];
        const int32_t start = pre + inc - c;
        tstart[d] = start;
        gbase[d] = offs[(int64_t)d * ntiles + blockIdx.x] - start;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const int d = dg[j];
        if (d < 0) continue;
        const int slot = tstart[d] + wcnt[wid][d] + rk[j];
        lk[slot] = k[j];
        lv[slot] = v[j];
    }
    __syncthreads();
    const int64_t cnt = n - tile0 < kSortTile ? n - tile0 : kSortTile;
    for (int slot = threadIdx.x; slot < cnt; slot += kSortThreads) {
        const uint64_t key = lk[slot];
        const int d = radix_digit(key, shift, lo_bits);
        const int64_t o = (int64_t)gbase[d] + slot;
        okeys[o] = key;
        ovals[o] = lv[slot];
    }
}

// ------------------------------------------------------------ aggregation

// cn[v] = newid[comm[v]]: the new id of each vertex's community, so that a
// row entry costs one gather.
// cn[v] = newid[comm[v]]; with cpos also ccn[cpos[v]] = cn[v] (the same in
// the renumbered column order).
__global__ void k_comm_newid(const int32_t *__restrict__ comm, const int32_t *__restrict__ newid,
                             int32_t V, int32_t *cn, const int32_t *__restrict__ cpos,
                             int32_t *ccn) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) {
        const int32_t x = newid[comm[v]];
        cn[v] = x;
        if (cpos) ccn[cpos[v]] = x;
    }
}

__device__ __forceinline__ void emit_row(const int32_t *__restrict__ col,
                                         const float *__restrict__ w,
                                         const int32_t *__restrict__ cn, uint64_t src,
                                         int64_t lo, int64_t hi, int st, uint64_t *keys,
                                         float *vals) {
    for (int64_t e = lo; e < hi; e += st) {
        const uint32_t dst = (uint32_t)cn[col[e]];
        keys[e] = (src << 32) | dst;
        vals[e] = w[e];
    }
}

// Row entries name their columns through gcol and gcn: the renumbered
// columns and cn in that order, or col and cn; the row's own id is cn[v].
__global__ __launch_bounds__(kRowThreads) void k_agg_pairs(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ gcol,
    const float *__restrict__ w, const int32_t *__restrict__ cn,
    const int32_t *__restrict__ gcn, const int32_t *__restrict__ order, SegTable st,
    const Chunk *__restrict__ chunks, uint64_t *keys, float *vals) {
    const int b = blockIdx.x;
    if (b >= st.chunk_block_start) {
        const Chunk ch = chunks[b - st.chunk_block_start];
        const uint64_t src = (uint32_t)cn[ch.v];
        emit_row(gcol, w, gcn, src, ch.lo + threadIdx.x, ch.hi, kRowThreads, keys, vals);
        return;
    }
    int si = 0;
    while (si + 1 < st.nseg && b >= st.s[si + 1].block_start) ++si;
    const Seg sg = st.s[si];
    const int lg = sg.lanes_log2;
    if (lg <= 6) {
        const int32_t idx = (b - sg.block_start) * (kRowThreads >> lg) + (threadIdx.x >> lg);
        if (idx < sg.count) {
            const int32_t v = order[sg.list_off + idx];
            const uint64_t src = (uint32_t)cn[v];
            emit_row(gcol, w, gcn, src, rowptr[v] + (threadIdx.x & ((1 << lg) - 1)),
                     rowptr[v + 1], 1 << lg, keys, vals);
        }
    } else {
        const int32_t v = order[sg.list_off + (b - sg.block_start)];
        const uint64_t src = (uint32_t)cn[v];
        emit_row(gcol, w, gcn, src, rowptr[v] + threadIdx.x, rowptr[v + 1], kRowThreads, keys,
                 vals);
    }
}

__global__ void k_run_heads(const uint64_t *__restrict__ keys, int64_t n, int32_t *head) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// FP64 run sums: a wave adds its lanes' values per run (runs are contiguous
// in sorted order, so a segmented shuffle scan by run position suffices),
// then one atomic per run piece of the wave.  Sums of weights on a dyadic
// grid are exact, hence independent of the order of the atomics.
__global__ __launch_bounds__(256) void k_run_reduce(const uint64_t *__restrict__ keys,
                                                    const float *__restrict__ vals,
                                                    const int32_t *__restrict__ head,
                                                    const int32_t *__restrict__ pos, int64_t n,
                                                    uint64_t *ukey, double *acc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool ok = i < n;
    const int32_t p = ok ? pos[i] + head[i] - 1 : -1;  // output slot of i's run
    double x = ok ? (double)vals[i] : 0.0;
    if (ok && head[i]) ukey[p] = keys[i];
    // inclusive segmented scan over lanes with the same p (contiguous)
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double y = __shfl_up(x, off, 64);
        const int32_t q = __shfl_up(p, off, 64);
        if (lane >= off && q == p) x += y;
    }
    // the last lane of each run piece in the wave adds the piece
    const int32_t pn = __shfl_down(p, 1, 64);
    const bool last = lane == 63 || pn != p;
    if (ok && last) atomicAdd(&acc[p], x);
}

__global__ void k_unique_rows(const uint64_t *__restrict__ ukey, const double *__restrict__ acc,
                              const int32_t *__restrict__ U, int32_t *cnt, int32_t *col,
                              float *w) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *U) return;
    const uint64_t key = ukey[k];
    const int32_t src = (int32_t)(key >> 32);
    col[k] = (int32_t)(uint32_t)key;
    w[k] = (float)acc[k];
    // entries of row src = end - start, added by the row's last and first
    // entry (commuting atomics; both may be the same entry)
    const bool first = k == 0 || (int32_t)(ukey[k - 1] >> 32) != src;
    const bool lastk = k + 1 == *U || (int32_t)(ukey[k + 1] >> 32) != src;
    if (lastk) atomicAdd(&cnt[src], (int32_t)(k + 1));
    if (first) atomicAdd(&cnt[src], -(int32_t)k);
}

}  // namespace

int64_t radix_scratch_elems(int64_t n) {
    const int64_t nt = (n + kSortTile - 1) / kSortTile;
    const int64_t m = 256 * (nt > 0 ? nt : 1);
    return 2 * m + scan_scratch_elems(m) + 16;
}

bool radix_sort_pairs(uint64_t *keys, float *vals, uint64_t *tk, float *tv, int64_t n,
                      const int *shifts, int npasses, int lo_bits, int32_t *scratch,
                      hipStream_t s) {
    if (n <= 0) return false;
    const int64_t nt = (n + kSortTile - 1) / kSortTile;
    const int64_t m = 256 * nt;
    int32_t *hist = scratch;
    int32_t *offs = scratch + m;
    int32_t *sscr = scratch + 2 * m;
    uint64_t *ik = keys, *ok = tk;
    float *iv = vals, *ov = tv;
    for (int p = 0; p < npasses; ++p) {
        hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)nt), dim3(kSortThreads), 0, s, ik, n,
                           shifts[p], lo_bits, nt, hist);
        exclusive_scan(hist, offs, m, nullptr, sscr, s);
        hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)nt), dim3(kSortThreads), 0, s, ik, iv,
                           n, shifts[p], lo_bits, nt, offs, ok, ov);
        std::swap(ik, ok);
        std::swap(iv, ov);
    }
    return (npasses & 1) != 0;
}

void launch_agg_pairs(const int32_t *rowptr, const int32_t *col, const float *w,
                      const int32_t *comm, const int32_t *newid, int32_t V, const int32_t *order,
                      const SegTable &st, const Chunk *chunks, int32_t *cn, uint64_t *keys,
                      float *vals, hipStream_t s, const int32_t *ccol, const int32_t *cpos,
                      int32_t *ccn) {
    const int32_t nb = st.chunk_block_start + st.nchunks;
    if (nb <= 0 || V <= 0) return;
    hipLaunchKernelGGL(k_comm_newid, dim3(blocks_for(V, 256)), dim3(256), 0, s, comm, newid, V, cn,
                       ccol ? cpos : nullptr, ccn);
    hipLaunchKernelGGL(k_agg_pairs, dim3(nb), dim3(kRowThreads), 0, s, rowptr, ccol ? ccol : col,
                       w, cn, ccol ? ccn : cn, order, st, chunks, keys, vals);
}

void launch_run_heads(const uint64_t *keys, int64_t n, int32_t *head, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_run_heads, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n, head);
}

void launch_run_reduce(const uint64_t *keys, const float *vals, const int32_t *head,
                       const int32_t *pos, int64_t n, uint64_t *ukey, double *acc, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_run_reduce, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, vals, head,
                       pos, n, ukey, acc);
}

void launch_unique_rows(const uint64_t *ukey, const double *acc, const int32_t *U, int32_t *cnt,
                        int32_t *col, float *w, int64_t n_max, hipStream_t s) {
    if (n_max <= 0) return;
    hipLaunchKernelGGL(k_unique_rows, dim3(blocks_for(n_max, 256)), dim3(256), 0, s, ukey, acc, U,
                       cnt, col, w);
}

}  // namespace lv
