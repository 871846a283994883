// Device-wide exclusive scan (reduce-then-scan over 2048-item tiles) and a
// stable key partition built on it.  Wavefront scans are 64-lane
// shuffle-up ladders; workgroup scans combine 4 wavefronts through LDS.
#include "lv_prims.h"

namespace lv {

namespace {

__device__ __forceinline__ int32_t wave_incl_scan(int32_t x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

// Exclusive scan of one value per thread across the workgroup (NW waves).
template <int NW>
__device__ __forceinline__ int32_t block_excl_scan(int32_t x, int32_t *lds, int32_t &total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int32_t inc = wave_incl_scan(x);
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    int32_t woff = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        int32_t t = lds[i];
        woff += (i < wid) ? t : 0;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return inc - x + woff;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int32_t *__restrict__ in,
                                                              int64_t n,
                                                              int32_t *__restrict__ partial) {
    __shared__ int32_t lds[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j)
        if (base + j < n) s += in[base + j];
    int32_t tot;
    block_excl_scan<kScanThreads / 64>(s, lds, tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_partials(int32_t *partial, int64_t nt,
                                                        int32_t *total) {
    __shared__ int32_t lds[16];
    int32_t carry = 0;
    for (int64_t b = 0; b < nt; b += 1024) {
        int64_t i = b + threadIdx.x;
        int32_t x = i < nt ? partial[i] : 0;
        int32_t tot;
        int32_t ex = block_excl_scan<16>(x, lds, tot);
        if (i < nt) partial[i] = ex + carry;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down(const int32_t *in, int32_t *out,
                                                            int64_t n,
                                                            const int32_t *__restrict__ partial) {
    __shared__ int32_t lds[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int32_t v[kScanItems];
    int32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = (base + j < n) ? in[base + j] : 0;
        s += v[j];
    }
    int32_t tot;
    int32_t run = block_excl_scan<kScanThreads / 64>(s, lds, tot) + partial[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
}

__global__ __launch_bounds__(kPartTile) void k_part_count(const int8_t *__restrict__ keys,
                                                          int32_t n, int nk, int32_t ntiles,
                                                          int32_t *__restrict__ cnt) {
    __shared__ int32_t c[64];
    if (threadIdx.x < 64) c[threadIdx.x] = 0;
    __syncthreads();
    const int32_t i = blockIdx.x * kPartTile + threadIdx.x;
    if (i < n) {
        int k = keys[i];
        if (k >= 0) atomicAdd(&c[k], 1);
    }
    __syncthreads();
    if (threadIdx.x < nk) cnt[(int64_t)threadIdx.x * ntiles + blockIdx.x] = c[threadIdx.x];
}

__global__ __launch_bounds__(kPartTile) void k_part_scatter(const int8_t *__restrict__ keys,
                                                            int32_t n, int nk, int32_t ntiles,
                                                            const int32_t *__restrict__ pos,
                                                            int32_t *__restrict__ items) {
    constexpr int NW = kPartTile / 64;
    __shared__ int32_t wc[NW][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int32_t i = blockIdx.x * kPartTile + threadIdx.x;
    const int key = (i < n) ? keys[i] : -1;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int rank = 0;
    for (int k = 0; k < nk; ++k) {
        uint64_t m = __ballot(key == k);
        if (lane == 0) wc[wid][k] = __popcll(m);
        if (key == k) rank = __popcll(m & lt);
    }
    __syncthreads();
    if (key >= 0) {
        int32_t off = pos[(int64_t)key * ntiles + blockIdx.x];
        for (int w = 0; w < wid; ++w) off += wc[w][key];
        items[off + rank] = i;
    }
}

__global__ void k_part_offsets(const int32_t *pos, int nk, int32_t ntiles, const int32_t *total,
                               int32_t *offsets) {
    int k = threadIdx.x;
    if (k < nk) offsets[k] = pos[(int64_t)k * ntiles];
    if (k == nk) offsets[nk] = *total;
}

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u8(uint8_t *p, uint8_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f64(double *p, double v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_iota_i32(int32_t *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = (int32_t)i;
}

inline unsigned grid_for(int64_t n, int threads) {
    int64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

}  // namespace

int64_t scan_scratch_elems(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void exclusive_scan(const int32_t *in, int32_t *out, int64_t n, int32_t *total, int32_t *scratch,
                    hipStream_t s) {
    if (n <= 0) {
        if (total) fill_i32(total, 0, 1, s);
        return;
    }
    const int64_t nt = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nt), dim3(kScanThreads), 0, s, in, n, scratch);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, scratch, nt, total);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nt), dim3(kScanThreads), 0, s, in, out, n,
                       scratch);
}

int64_t partition_scratch_elems(int64_t n, int nk) {
    const int64_t ntiles = (n + kPartTile - 1) / kPartTile + 1;
    const int64_t m = ntiles * nk;
    return m + scan_scratch_elems(m) + 2;
}

void stable_partition(const int8_t *keys, int32_t n, int nk, int32_t *items, int32_t *offsets,
                      int32_t *scratch, hipStream_t s) {
    const int32_t ntiles = (n + kPartTile - 1) / kPartTile;
    if (n <= 0) {
        fill_i32(offsets, 0, nk + 1, s);
        return;
    }
    const int64_t m = (int64_t)ntiles * nk;
    int32_t *cnt = scratch;
    int32_t *total = scratch + m;
    int32_t *sscr = scratch + m + 2;
    hipLaunchKernelGGL(k_part_count, dim3(ntiles), dim3(kPartTile), 0, s, keys, n, nk, ntiles, cnt);
    exclusive_scan(cnt, cnt, m, total, sscr, s);
    hipLaunchKernelGGL(k_part_scatter, dim3(ntiles), dim3(kPartTile), 0, s, keys, n, nk, ntiles,
                       cnt, items);
    hipLaunchKernelGGL(k_part_offsets, dim3(1), dim3(64), 0, s, cnt, nk, ntiles, total, offsets);
}

void fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
}
void fill_u8(uint8_t *p, uint8_t v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
}
void fill_f64(double *p, double v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
}
void iota_i32(int32_t *p, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_iota_i32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n);
}

}  // namespace lv
