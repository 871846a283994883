// Modularity of a partition over caller-owned device arrays, with no device
// allocation: for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The same quantity as lv_modularity (calculateModularity,
// modularity_optimisation.cu:365-401), FP64:
//   Q = own / 2M - (sum_c S_c^2) / (2M)^2,
// own = the weight of the entries whose two ends share a community (a
// self-loop entry counts once), S_c = the sum of the strengths of c's
// members, 2M = the sum of all CSR weights.
//
// Scratch lives in one file-scope device array (`scratch`), one copy per
// GPU, held until the process ends; its address on each device is looked up
// once (hipGetSymbolAddress) and only that address is handed to kernels.
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// race.  The kernels clear what must start at zero themselves (no
// hipMemsetAsync), so a call may be captured in a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "../../include/louvain.h"
#include "lv_error.h"

namespace {

constexpr int32_t kNaMaxV = LV_NOALLOC_MAX_V;
constexpr int kNaBlocks = 1024;   // fixed grid: partials stay in place
constexpr int kNaThreads = 256;
constexpr int kMaxDevices = 64;
// Layout (doubles): S_c [kNaMaxV] | vertex-pass partials (own, 2M) [2 x
// kNaBlocks] | square-pass partials [kNaBlocks] | flag word.
constexpr int64_t kOffPartV = kNaMaxV;
constexpr int64_t kOffPartS = kOffPartV + 2 * kNaBlocks;
constexpr int64_t kOffFlag = kOffPartS + kNaBlocks;
constexpr int64_t kNaDoubles = kOffFlag + 32;

static __device__ __align__(256) double scratch[kNaDoubles];

double *g_scratch[kMaxDevices] = {nullptr};
std::mutex g_scratch_mu;

lv_status na_fail(lv_status st, const char *m) {
    lv::set_error(m);
    return st;
}

__device__ __forceinline__ double na_block_sum(double x, double *lds) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    __syncthreads();
    if (lane == 0) lds[wid] = x;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < kNaThreads / 64; ++i) t += lds[i];
    __syncthreads();
    return t;
}

// S_c = 0 for c < V and the flag cleared.
__global__ __launch_bounds__(kNaThreads) void k_na_clear(double *sigma, int32_t V, int32_t *flag) {
    for (int64_t c = (int64_t)blockIdx.x * kNaThreads + threadIdx.x; c < V;
         c += (int64_t)gridDim.x * kNaThreads)
        sigma[c] = 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0;
}

// One wave per vertex (grid-stride): its strength goes to S_comm[v]; its
// entries inside its community and all its weight go to the block partials.
// A row, column or community outside its range raises the flag and is
// skipped (nothing out of bounds is read or written).
__global__ __launch_bounds__(kNaThreads) void k_na_vertices(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const float *__restrict__ w, const int32_t *__restrict__ comm, int32_t V, int64_t E,
    double *sigma, double *part, int32_t *flag) {
    __shared__ double lds[kNaThreads / 64];
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (kNaThreads / 64);
    double own = 0.0, total = 0.0;
    bool bad = false;
    for (int64_t v = (int64_t)blockIdx.x * (kNaThreads / 64) + (threadIdx.x >> 6); v < V;
         v += waves) {
        const int32_t lo = rowptr[v], hi = rowptr[v + 1];
        const int32_t cv = comm[v];
        if (lo < 0 || hi < lo || (int64_t)hi > E || cv < 0 || cv >= V) {
            bad = true;
            continue;
        }
        double kv = 0.0;
        for (int32_t e = lo + lane; e < hi; e += 64) {
            const int32_t u = col[e];
            const double x = (double)w[e];
            kv += x;
            if (u < 0 || u >= V) {
                bad = true;
                continue;
            }
            if (comm[u] == cv) own += x;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) kv += __shfl_xor(kv, off, 64);
        if (lane == 0 && kv != 0.0) atomicAdd(&sigma[cv], kv);
        total += lane == 0 ? kv : 0.0;
    }
    if (__ballot(bad) && lane == 0) atomicOr(flag, 1);
    own = na_block_sum(own, lds);
    total = na_block_sum(total, lds);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = own;
        part[2 * blockIdx.x + 1] = total;
    }
}

__global__ __launch_bounds__(kNaThreads) void k_na_squares(const double *__restrict__ sigma,
                                                           int32_t V, double *part) {
    __shared__ double lds[kNaThreads / 64];
    double sq = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * kNaThreads + threadIdx.x; c < V;
         c += (int64_t)gridDim.x * kNaThreads) {
        const double x = sigma[c];
        sq += x * x;
    }
    sq = na_block_sum(sq, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = sq;
}

// Fixed-order sums of the partials; q = NaN when the flag is raised.
__global__ __launch_bounds__(kNaThreads) void k_na_final(const double *__restrict__ partv,
                                                         const double *__restrict__ parts,
                                                         int32_t nv, int32_t ns,
                                                         const int32_t *__restrict__ flag,
                                                         double *q) {
    __shared__ double lds[kNaThreads / 64];
    double own = 0.0, total = 0.0, sq = 0.0;
    for (int i = threadIdx.x; i < nv; i += kNaThreads) {
        own += partv[2 * i];
        total += partv[2 * i + 1];
    }
    for (int i = threadIdx.x; i < ns; i += kNaThreads) sq += parts[i];
    own = na_block_sum(own, lds);
    total = na_block_sum(total, lds);
    sq = na_block_sum(sq, lds);
    if (threadIdx.x == 0)
        *q = *flag ? __builtin_nan("") : own / total - sq / (total * total);
}

// This device's address of `scratch`, looked up once per device with that
// device current.
lv_status scratch_of(int32_t device, double **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return na_fail(LV_ERR_HIP, "hipGetSymbolAddress(scratch) failed");
        g_scratch[device] = static_cast<double *>(p);
    }
    *out = g_scratch[device];
    return LV_OK;
}

// Restores the caller's current device on every return path.
struct DeviceGuard {
    int prev = -1;
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

}  // namespace

extern "C" {

lv_status lv_modularity_noalloc(int32_t device, void *stream, const int32_t *rowptr,
                                const int32_t *col, const float *w, const int32_t *comm,
                                int32_t V, int64_t E, double *q) {
    if (V < 0 || V > kNaMaxV)
        return na_fail(LV_ERR_ARG, "lv_modularity_noalloc: V exceeds LV_NOALLOC_MAX_V");
    if (E < 0 || E > 0x7fffffffLL) return na_fail(LV_ERR_ARG, "lv_modularity_noalloc: E out of range");
    if (!rowptr || !q || (V > 0 && !comm) || (E > 0 && (!col || !w)))
        return na_fail(LV_ERR_ARG, "lv_modularity_noalloc: null array");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return na_fail(LV_ERR_NODEVICE, "no HIP device visible");
    if (device < 0 || device >= ndev || device >= kMaxDevices)
        return na_fail(LV_ERR_ARG, "device ordinal out of range");
    DeviceGuard guard;
    if (hipGetDevice(&guard.prev) != hipSuccess) guard.prev = -1;
    if (hipSetDevice(device) != hipSuccess) return na_fail(LV_ERR_HIP, "hipSetDevice failed");
    double *base = nullptr;
    const lv_status st = scratch_of(device, &base);
    if (st != LV_OK) return st;
    double *sigma = base;
    double *partv = base + kOffPartV;
    double *parts = base + kOffPartS;
    int32_t *flag = reinterpret_cast<int32_t *>(base + kOffFlag);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t want_v = ((int64_t)V + kNaThreads / 64 - 1) / (kNaThreads / 64);
    const int64_t want_c = ((int64_t)V + kNaThreads - 1) / kNaThreads;
    const int nv = (int)std::max<int64_t>(1, std::min<int64_t>(kNaBlocks, want_v));
    const int ns = (int)std::max<int64_t>(1, std::min<int64_t>(kNaBlocks, want_c));
    hipLaunchKernelGGL(k_na_clear, dim3(ns), dim3(kNaThreads), 0, s, sigma, V, flag);
    hipLaunchKernelGGL(k_na_vertices, dim3(nv), dim3(kNaThreads), 0, s, rowptr, col, w, comm, V, E,
                       sigma, partv, flag);
    hipLaunchKernelGGL(k_na_squares, dim3(ns), dim3(kNaThreads), 0, s, sigma, V, parts);
    hipLaunchKernelGGL(k_na_final, dim3(1), dim3(kNaThreads), 0, s, partv, parts, nv, ns, flag, q);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return na_fail(LV_ERR_HIP, hipGetErrorString(e));
    return LV_OK;
}

}  // extern "C"
