// Phase 2 (community aggregation) kernels for gfx950.
//
// Restates aggregateCommunities (community_aggregation.cu:248-347): the new
// vertex of community c is newid[c] (order-preserving compaction of the
// non-empty communities, :267-273); its row lists each distinct newid[comm[u]]
// over the row entries (u,w) of all members of c, with summed weights
// (mergeCommunity, :76-163).  Where the reference emits in hash-claim order,
// rows here come out in ascending new id and sums are FP64 rounded once to
// fp32, so the CSR is canonical.
//
// Members of a community form one flat range of row entries (scan of member
// degrees), so work splits evenly whatever the mix of member degrees:
//   k_merge_lds<T,..>  one workgroup per community; each lane walks a
//                      contiguous slice of the flat range; LDS hash, then
//                      compaction and an LDS bitonic sort of (key, slot).
//   k_agg_insert/extract/sort  communities beyond the LDS tier: flat-range
//                      chunks from many workgroups into a global hash,
//                      extraction, then a one-workgroup sort (LDS bitonic up
//                      to 8192 entries, merge-path passes beyond).
#include "lv_phase2.h"

namespace lv {

namespace {

constexpr int ilog2_c(int x) { return x <= 1 ? 0 : 1 + ilog2_c(x >> 1); }

__device__ __forceinline__ uint64_t pack_kw(int32_t key, float w) {
    return ((uint64_t)(uint32_t)key << 32) | (uint64_t)__float_as_uint(w);
}

// Bitonic sort of buf[0..P) ascending, P a power of two; whole workgroup.
__device__ void bitonic_sort(uint64_t *buf, int P) {
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t x = buf[i], y = buf[ixj];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        buf[i] = y;
                        buf[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ int pow2_at_least(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

__global__ void k_present(const int32_t *__restrict__ size, int32_t V, int32_t *flag) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < V) flag[c] = size[c] > 0 ? 1 : 0;
}

__global__ void k_has_degree(const int32_t *__restrict__ rowptr, const int32_t *__restrict__ comm,
                             int32_t V, int32_t *flag) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V && rowptr[v + 1] > rowptr[v]) flag[comm[v]] = 1;
}

__global__ void k_present_compact(const int32_t *__restrict__ size, int32_t V, int32_t *flag) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < V) flag[c] = (size[c] > 0 && flag[c]) ? 1 : 0;
}

__global__ void k_ref_present(const int32_t *__restrict__ size, const int32_t *__restrict__ ref_of,
                              int32_t V, int32_t *present_ref) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < V) present_ref[ref_of[c]] = size[c] > 0 ? 1 : 0;
}

__global__ void k_next_ref_of(const int32_t *__restrict__ keep, const int32_t *__restrict__ newid,
                              const int32_t *__restrict__ ref_of,
                              const int32_t *__restrict__ newid_ref, int32_t V,
                              int32_t *ref_of_next) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < V && keep[c]) ref_of_next[newid[c]] = newid_ref[ref_of[c]];
}

__global__ void k_cmp_of(const int32_t *__restrict__ ref_of, int32_t V, int32_t *cmp_of) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) cmp_of[ref_of[i]] = i;
}

__global__ void k_compose_ref(int32_t *orig2comm, int32_t Vorig, const int32_t *__restrict__ cmp_of,
                              const int32_t *__restrict__ ref_of,
                              const int32_t *__restrict__ comm) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= Vorig) return;
    const int32_t i = cmp_of[orig2comm[x]];
    if (i >= 0) orig2comm[x] = ref_of[comm[i]];
}

__global__ void k_members(const int32_t *__restrict__ comm, int32_t V,
                          const int32_t *__restrict__ mstart, int32_t *fillc, int32_t *members) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) {
        const int32_t c = comm[v];
        members[mstart[c] + atomicAdd(&fillc[c], 1)] = v;
    }
}

__global__ void k_member_degrees(const int32_t *__restrict__ rowptr,
                                 const int32_t *__restrict__ members, int32_t V, int32_t *mdeg) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) {
        const int32_t v = members[i];
        mdeg[i] = rowptr[v + 1] - rowptr[v];
    }
}

__global__ void k_agg_keys(const int32_t *__restrict__ size, const int32_t *__restrict__ mstart,
                           const int32_t *__restrict__ mscan, int32_t V, AggBounds bounds,
                           int8_t *keys) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= V) return;
    const int32_t sz = size[c];
    int8_t key = -1;
    if (sz > 0) {
        const int32_t m0 = mstart[c];
        const int32_t D = mscan[m0 + sz] - mscan[m0];
        // The LDS tiers stage the member-degree scan too, so tier by the
        // larger of D and the member count (members of degree 0 are
        // possible only in externally supplied states).
        const int32_t need = D > sz ? D : sz;
        if (D > 0) {
            key = kAggTiers - 1;
            for (int t = 0; t < kAggTiers - 1; ++t)
                if (need <= bounds.hi[t]) {
                    key = (int8_t)t;
                    break;
                }
        }
    }
    keys[c] = key;
}

// Largest i in [0, n) with seg[i] <= x (seg ascending, seg[0] <= x).
__device__ __forceinline__ int32_t upper_index(const int32_t *seg, int32_t n, int32_t x) {
    int32_t a = 0, b = n - 1;
    while (a < b) {
        const int32_t mid = (a + b + 1) >> 1;
        if (seg[mid] <= x) a = mid;
        else b = mid - 1;
    }
    return a;
}

template <int T, int CAP_LOG2, int PMAX>
__global__ __launch_bounds__(T) void k_merge_lds(AggArgs a, const int32_t *__restrict__ list,
                                                 int32_t n) {
    constexpr int C = 1 << CAP_LOG2;
    extern __shared__ double agg_lds[];
    double *vals = agg_lds;                                             // [C]
    uint64_t *sbuf = reinterpret_cast<uint64_t *>(vals + C);            // [PMAX]
    int32_t *keys = reinterpret_cast<int32_t *>(sbuf + PMAX);           // [C]
    int32_t *mseg = keys + C;                                           // [PMAX+1]
    int32_t *counter = mseg + PMAX + 1;                                 // [1]
    for (int32_t li = blockIdx.x; li < n; li += gridDim.x) {
        const int32_t c = list[li];
        const int32_t m0 = a.mstart[c];
        const int32_t msz = a.size[c];
        const int32_t f0 = a.mscan[m0];
        for (int i = threadIdx.x; i <= msz; i += T) mseg[i] = a.mscan[m0 + i] - f0;
        // table of the community: smallest power of two >= 2*D (load <= 1/2),
        // at least 64 slots, at most the allocation
        const int32_t Dc = a.mscan[m0 + msz] - f0;
        const int cl = min(CAP_LOG2, max(6, Dc > 1 ? 32 - __clz(2 * Dc - 1) : 1));
        const int Cv = 1 << cl;
        const uint32_t shift = 32 - cl;
        for (int i = threadIdx.x; i < Cv; i += T) {
            keys[i] = -1;
            vals[i] = 0.0;
        }
        if (threadIdx.x == 0) *counter = 0;
        __syncthreads();
        const int32_t D = mseg[msz];
        const int32_t per = (D + T - 1) / T;
        const int32_t lo = threadIdx.x * per;
        const int32_t hi = lo + per < D ? lo + per : D;
        if (lo < hi) {
            int32_t i = upper_index(mseg, msz, lo);
            int32_t v = a.members[m0 + i];
            int32_t base = a.rowptr[v] - mseg[i];
            for (int32_t f = lo; f < hi; ++f) {
                if (mseg[i + 1] <= f) {
                    do {
                        ++i;
                    } while (mseg[i + 1] <= f);
                    v = a.members[m0 + i];
                    base = a.rowptr[v] - mseg[i];
                }
                const int32_t e = base + f;
                const int32_t key = a.newid[a.comm[a.col[e]]];
                lds_insert(keys, vals, shift, (uint32_t)Cv - 1, key, (double)a.w[e]);
            }
        }
        __syncthreads();
        for (int s = threadIdx.x; s < Cv; s += T) {
            const int32_t k = keys[s];
            if (k >= 0) sbuf[atomicAdd(counter, 1)] = ((uint64_t)(uint32_t)k << 32) | (uint32_t)s;
        }
        __syncthreads();
        const int32_t cnt = *counter;
        const int P = pow2_at_least(cnt);
        for (int i = cnt + threadIdx.x; i < P; i += T) sbuf[i] = ~0ull;
        __syncthreads();
        bitonic_sort(sbuf, P);
        for (int i = threadIdx.x; i < cnt; i += T) {
            const uint64_t x = sbuf[i];
            a.tmp_col[f0 + i] = (int32_t)(x >> 32);
            a.tmp_w[f0 + i] = (float)vals[(uint32_t)x];
        }
        if (threadIdx.x == 0) a.cnt[c] = cnt;
        __syncthreads();
    }
}

constexpr int kAggChunkThreads = 256;

__global__ __launch_bounds__(kAggChunkThreads) void k_agg_insert(AggArgs a,
                                                                 const AggChunk *__restrict__ ins,
                                                                 const int64_t *__restrict__ off,
                                                                 const int32_t *__restrict__ lg,
                                                                 int32_t *keys, double *vals) {
    const AggChunk ch = ins[blockIdx.x];
    const int32_t c = ch.c;
    const int32_t m0 = a.mstart[c];
    const int32_t msz = a.size[c];
    const int32_t *mseg = a.mscan + m0;  // absolute flat offsets
    const int32_t f0 = mseg[0];
    const int64_t len = ch.hi - ch.lo;
    const int64_t per = (len + kAggChunkThreads - 1) / kAggChunkThreads;
    const int64_t lo64 = ch.lo + (int64_t)threadIdx.x * per;
    const int64_t hi64 = lo64 + per < ch.hi ? lo64 + per : ch.hi;
    if (lo64 >= hi64) return;
    const int32_t lo = f0 + (int32_t)lo64, hi = f0 + (int32_t)hi64;
    const int lgc = lg[ch.ti];
    int32_t *K = keys + off[ch.ti];
    double *W = vals + off[ch.ti];
    int32_t i = upper_index(mseg, msz, lo);
    int32_t v = a.members[m0 + i];
    int32_t base = a.rowptr[v] - mseg[i];
    for (int32_t f = lo; f < hi; ++f) {
        if (mseg[i + 1] <= f) {
            do {
                ++i;
            } while (mseg[i + 1] <= f);
            v = a.members[m0 + i];
            base = a.rowptr[v] - mseg[i];
        }
        const int32_t e = base + f;
        const int32_t key = a.newid[a.comm[a.col[e]]];
        global_insert(K, W, 32 - lgc, (1u << lgc) - 1, key, (double)a.w[e]);
    }
}

__global__ __launch_bounds__(kAggChunkThreads) void k_agg_extract(AggArgs a,
                                                                  const AggChunk *__restrict__ ext,
                                                                  int32_t *keys, double *vals) {
    const AggChunk ch = ext[blockIdx.x];
    const int32_t c = ch.c;
    const int32_t f0 = a.mscan[a.mstart[c]];
    for (int64_t s = ch.lo + threadIdx.x; s < ch.hi; s += kAggChunkThreads) {
        const int32_t k = keys[s];
        if (k < 0) continue;
        const double x = vals[s];
        keys[s] = -1;  // leave the arena clean
        vals[s] = 0.0;
        const int32_t idx = atomicAdd(&a.cnt[c], 1);
        a.tmp_col[f0 + idx] = k;
        a.tmp_w[f0 + idx] = (float)x;
    }
}

constexpr int kSortThreads = 1024;
constexpr int kSortRun = 8192;

__global__ __launch_bounds__(kSortThreads) void k_agg_sort(AggArgs a,
                                                           const int32_t *__restrict__ glist,
                                                           uint64_t *A, uint64_t *B,
                                                           const int64_t *__restrict__ sort_off) {
    extern __shared__ uint64_t sort_lds[];  // [kSortRun]
    const int32_t ti = blockIdx.x;
    const int32_t c = glist[ti];
    const int32_t n = a.cnt[c];
    const int32_t f0 = a.mscan[a.mstart[c]];
    int32_t *kc = a.tmp_col + f0;
    float *kw = a.tmp_w + f0;
    if (n <= kSortRun) {
        const int P = pow2_at_least(n > 0 ? n : 1);
        for (int i = threadIdx.x; i < P; i += kSortThreads)
            sort_lds[i] = i < n ? pack_kw(kc[i], kw[i]) : ~0ull;
        __syncthreads();
        bitonic_sort(sort_lds, P);
        for (int i = threadIdx.x; i < n; i += kSortThreads) {
            const uint64_t x = sort_lds[i];
            kc[i] = (int32_t)(x >> 32);
            kw[i] = __uint_as_float((uint32_t)x);
        }
        return;
    }
    uint64_t *src = A + sort_off[ti];
    uint64_t *dst = B + sort_off[ti];
    for (int32_t r0 = 0; r0 < n; r0 += kSortRun) {
        const int32_t len = n - r0 < kSortRun ? n - r0 : kSortRun;
        for (int i = threadIdx.x; i < kSortRun; i += kSortThreads)
            sort_lds[i] = i < len ? pack_kw(kc[r0 + i], kw[r0 + i]) : ~0ull;
        __syncthreads();
        bitonic_sort(sort_lds, kSortRun);
        for (int i = threadIdx.x; i < len; i += kSortThreads) src[r0 + i] = sort_lds[i];
        __syncthreads();
    }
    for (int64_t L = kSortRun; L < n; L <<= 1) {
        for (int64_t p0 = 0; p0 < n; p0 += 2 * L) {
            const int64_t alen = n - p0 < L ? n - p0 : L;
            const int64_t b0 = p0 + alen;
            const int64_t blen = n - b0 < L ? (n - b0 > 0 ? n - b0 : 0) : L;
            const int64_t total = alen + blen;
            const int64_t per = (total + kSortThreads - 1) / kSortThreads;
            const int64_t d = (int64_t)threadIdx.x * per;
            if (d >= total) continue;
            const int64_t dend = d + per < total ? d + per : total;
            const uint64_t *Ap = src + p0;
            const uint64_t *Bp = src + b0;
            int64_t lo = d - blen > 0 ? d - blen : 0;
            int64_t hi = d < alen ? d : alen;
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (Ap[mid] < Bp[d - mid - 1]) lo = mid + 1;
                else hi = mid;
            }
            int64_t i = lo, j = d - lo;
            for (int64_t o = d; o < dend; ++o) {
                const bool takeA = j >= blen || (i < alen && Ap[i] < Bp[j]);
                dst[p0 + o] = takeA ? Ap[i++] : Bp[j++];
            }
        }
        __syncthreads();
        uint64_t *t = src;
        src = dst;
        dst = t;
    }
    for (int32_t i = threadIdx.x; i < n; i += kSortThreads) {
        const uint64_t x = src[i];
        kc[i] = (int32_t)(x >> 32);
        kw[i] = __uint_as_float((uint32_t)x);
    }
}

__global__ void k_new_rowptr(const int32_t *__restrict__ size, const int32_t *__restrict__ newid,
                             const int32_t *__restrict__ cstart, int32_t V,
                             const int32_t *__restrict__ newV, const int32_t *__restrict__ total,
                             int32_t *rowptr2) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < V && size[c] > 0) rowptr2[newid[c]] = cstart[c];
    if (c == 0) rowptr2[*newV] = *total;
}

template <int LANES>
__global__ __launch_bounds__(256) void k_copy_rows(const int32_t *__restrict__ list, int32_t n,
                                                   const int32_t *__restrict__ mstart,
                                                   const int32_t *__restrict__ mscan,
                                                   const int32_t *__restrict__ cnt,
                                                   const int32_t *__restrict__ cstart,
                                                   const int32_t *__restrict__ tmp_col,
                                                   const float *__restrict__ tmp_w,
                                                   int32_t *col2, float *w2) {
    constexpr int PER = 256 / LANES;
    const int32_t li = blockIdx.x * PER + threadIdx.x / LANES;
    if (li >= n) return;
    const int lane = threadIdx.x & (LANES - 1);
    const int32_t c = list[li];
    const int32_t f0 = mscan[mstart[c]];
    const int32_t o0 = cstart[c];
    const int32_t m = cnt[c];
    for (int32_t i = lane; i < m; i += LANES) {
        col2[o0 + i] = tmp_col[f0 + i];
        w2[o0 + i] = tmp_w[f0 + i];
    }
}

inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

template <int T, int CAP_LOG2, int PMAX>
constexpr int merge_bytes() {
    return (1 << CAP_LOG2) * 12 + PMAX * 8 + (PMAX + 2) * 4;
}

}  // namespace

void launch_present(const int32_t *size, int32_t V, int32_t *flag, hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_present, dim3(blocks_for(V, 256)), dim3(256), 0, s, size, V, flag);
}

void launch_has_degree(const int32_t *rowptr, const int32_t *comm, int32_t V, int32_t *flag,
                       hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_has_degree, dim3(blocks_for(V, 256)), dim3(256), 0, s, rowptr, comm, V,
                       flag);
}

void launch_present_compact(const int32_t *size, int32_t V, int32_t *flag, hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_present_compact, dim3(blocks_for(V, 256)), dim3(256), 0, s, size, V, flag);
}

void launch_ref_present(const int32_t *size, const int32_t *ref_of, int32_t V,
                        int32_t *present_ref, hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_ref_present, dim3(blocks_for(V, 256)), dim3(256), 0, s, size, ref_of, V,
                       present_ref);
}

void launch_next_ref_of(const int32_t *keep, const int32_t *newid, const int32_t *ref_of,
                        const int32_t *newid_ref, int32_t V, int32_t *ref_of_next, hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_next_ref_of, dim3(blocks_for(V, 256)), dim3(256), 0, s, keep, newid, ref_of,
                       newid_ref, V, ref_of_next);
}

void launch_cmp_of(const int32_t *ref_of, int32_t V, int32_t *cmp_of, hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_cmp_of, dim3(blocks_for(V, 256)), dim3(256), 0, s, ref_of, V, cmp_of);
}

void launch_compose_ref(int32_t *orig2comm, int32_t Vorig, const int32_t *cmp_of,
                        const int32_t *ref_of, const int32_t *comm, hipStream_t s) {
    if (Vorig <= 0) return;
    hipLaunchKernelGGL(k_compose_ref, dim3(blocks_for(Vorig, 256)), dim3(256), 0, s, orig2comm,
                       Vorig, cmp_of, ref_of, comm);
}

void launch_members(const int32_t *comm, int32_t V, const int32_t *mstart, int32_t *fillc,
                    int32_t *members, hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_members, dim3(blocks_for(V, 256)), dim3(256), 0, s, comm, V, mstart, fillc,
                       members);
}

void launch_member_degrees(const int32_t *rowptr, const int32_t *members, int32_t V,
                           int32_t *mdeg, hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_member_degrees, dim3(blocks_for(V, 256)), dim3(256), 0, s, rowptr, members,
                       V, mdeg);
}

void launch_agg_keys(const int32_t *size, const int32_t *mstart, const int32_t *mscan, int32_t V,
                     AggBounds bounds, int8_t *keys, hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_agg_keys, dim3(blocks_for(V, 256)), dim3(256), 0, s, size, mstart, mscan,
                       V, bounds, keys);
}

int merge_lds_bytes(int tier) {
    switch (tier) {
    case 0: return merge_bytes<64, 8, 128>();
    case 1: return merge_bytes<256, 11, 1024>();
    default: return merge_bytes<512, 13, 4096>();
    }
}

void set_merge_lds_limits(int bytes) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_merge_lds<64, 8, 128>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_merge_lds<256, 11, 1024>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_merge_lds<512, 13, 4096>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_agg_sort),
                        hipFuncAttributeMaxDynamicSharedMemorySize, kSortRun * 8);
}

void launch_merge_lds(int tier, const AggArgs &a, const int32_t *list, int32_t n, hipStream_t s) {
    if (n <= 0) return;
    const unsigned g = (unsigned)(n < 1048576 ? n : 1048576);
    switch (tier) {
    case 0:
        k_merge_lds<64, 8, 128><<<g, 64, merge_bytes<64, 8, 128>(), s>>>(a, list, n);
        break;
    case 1:
        k_merge_lds<256, 11, 1024><<<g, 256, merge_bytes<256, 11, 1024>(), s>>>(a, list, n);
        break;
    default:
        k_merge_lds<512, 13, 4096><<<g, 512, merge_bytes<512, 13, 4096>(), s>>>(a, list, n);
        break;
    }
}

void launch_agg_global(const AggArgs &a, const AggChunk *ins, int32_t nins,
                       const AggChunk *ext, int32_t next, const int32_t *glist, int32_t ng,
                       const int64_t *tab_off, const int32_t *cap_log2, int32_t *keys,
                       double *vals, uint64_t *sort_a, uint64_t *sort_b,
                       const int64_t *sort_off, hipStream_t s) {
    if (ng <= 0) return;
    hipLaunchKernelGGL(k_agg_insert, dim3(nins), dim3(kAggChunkThreads), 0, s, a, ins, tab_off,
                       cap_log2, keys, vals);
    hipLaunchKernelGGL(k_agg_extract, dim3(next), dim3(kAggChunkThreads), 0, s, a, ext, keys,
                       vals);
    hipLaunchKernelGGL(k_agg_sort, dim3(ng), dim3(kSortThreads), kSortRun * 8, s, a, glist, sort_a,
                       sort_b, sort_off);
}

void launch_new_rowptr(const int32_t *size, const int32_t *newid, const int32_t *cstart,
                       int32_t V, const int32_t *newV, const int32_t *total, int32_t *rowptr2,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_new_rowptr, dim3(blocks_for(V > 0 ? V : 1, 256)), dim3(256), 0, s, size,
                       newid, cstart, V, newV, total, rowptr2);
}

void launch_copy_rows(const int32_t *list, int32_t n, bool wide, const int32_t *mstart,
                      const int32_t *mscan, const int32_t *cnt, const int32_t *cstart,
                      const int32_t *tmp_col, const float *tmp_w, int32_t *col2, float *w2,
                      hipStream_t s) {
    if (n <= 0) return;
    if (wide)
        hipLaunchKernelGGL(k_copy_rows<256>, dim3(n), dim3(256), 0, s, list, n, mstart, mscan, cnt,
                           cstart, tmp_col, tmp_w, col2, w2);
    else
        hipLaunchKernelGGL(k_copy_rows<64>, dim3(blocks_for(n, 4)), dim3(256), 0, s, list, n,
                           mstart, mscan, cnt, cstart, tmp_col, tmp_w, col2, w2);
}

}  // namespace lv
