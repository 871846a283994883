// Learned frontier kernels for gfx950.
//
// Features (8 fp32 per vertex, from the state after the last sweep):
//   x0 log2(1+deg)/16                        own-degree scale
//   x1 K[cur]/k at the last evaluation        own-community weight fraction
//   x2 clamp(margin*M/k, -4, 4)/4             last best-gain margin, in units of k/M
//   x3 moves in the last 4 sweeps / 4
//   x4 log2(1+size(cur))/24                   community size
//   x5 log2(1 + V*S_cur/(2M))/24              community total relative to the mean vertex
//   x6 min(1, nmv/deg)                        share of row entries whose neighbour moved
//                                             (neighbours of degree <= kNmvMaxDegree)
//   x7 min(sweeps since the last move, 8)/8
// where margin = bestGain - K[cur]/M of the vertex's last evaluation (> 0 iff
// it moved then) and S_cur, size(cur) are read from the community record.
//
// Inference runs the two hidden layers on MFMA: v_mfma_f32_16x16x32_bf16
// with A = 16 vertices x 32 inputs (features zero-padded from 8) and
// B = weights^T; lane l holds A[l%16][8*(l/16)+j] and B[8*(l/16)+j][l%16]
// (j < 8) and receives D[4*(l/16)+r][l%16] (r < 4).  Inputs, W1, the layer-1
// activations and W2 are rounded to bf16; accumulation, biases and the
// output layer (32 -> 1, a 16-lane reduction) are fp32.
#include "lv_learn.h"

namespace lv {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

__device__ __forceinline__ void features_from(int32_t deg, float kv, float scur, int32_t sz,
                                              uint32_t h, float kown, float margin, int32_t nmv,
                                              float M, float vscale, float *x) {
    x[0] = log2f(1.0f + (float)deg) * (1.0f / 16.0f);
    x[1] = kv > 0.0f ? kown / kv : 0.0f;
    x[2] = kv > 0.0f ? fminf(fmaxf(margin * M / kv, -4.0f), 4.0f) * 0.25f : 0.0f;
    x[3] = (float)__popc(h & 0xfu) * 0.25f;
    x[4] = log2f(1.0f + (float)sz) * (1.0f / 24.0f);
    x[5] = log2f(1.0f + scur * vscale) * (1.0f / 24.0f);
    x[6] = deg > 0 ? fminf(1.0f, (float)nmv / (float)deg) : 0.0f;
    x[7] = (float)min(__ffs(h | 0x100u) - 1, 8) * 0.125f;
}

__device__ __forceinline__ void features_of(const FeatureArgs &f, int32_t v, float *x) {
    const uint64_t r = f.crec[f.comm[v]];
    float M = f.M, vscale = f.vscale;
    if (f.gid) {
        const int32_t g = f.gid[v];
        M = f.gM[g];
        vscale = f.gvscale[g];
    }
    features_from(f.rowptr[v + 1] - f.rowptr[v], (float)f.k[v], rec_sigma(r), rec_size(r),
                  f.hist[v], f.kown[v], f.margin[v], f.nmv[v], M, vscale, x);
}

__global__ void k_graph_model_frontier(const int32_t *__restrict__ rowptr,
                                       const int32_t *__restrict__ list, int32_t n,
                                       const int32_t *__restrict__ gid,
                                       const uint8_t *__restrict__ gact,
                                       const uint8_t *__restrict__ gfull, BucketBounds bb,
                                       int8_t *keys, uint8_t *active) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t v = list[i];
    const int32_t g = gid[v];
    const bool act = gact[g] && (gfull[g] || active[v]);
    active[v] = act ? 1 : 0;
    keys[v] = act ? (int8_t)bucket_of(rowptr[v + 1] - rowptr[v], bb) : (int8_t)-1;
}

// Features of a vertex without edges (a singleton of total 0 that never
// moved and was never evaluated), in out[0..8).
__global__ void k_isolated_features(float M, float vscale, float *out) {
    float x[kFeat];
    features_from(0, 0.0f, 0.0f, 1, 0u, 0.0f, 0.0f, 0, M, vscale, x);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < kFeat; ++j) out[j] = x[j];
    }
}

__global__ void k_hist_shift(uint8_t *hist, int32_t V, const int32_t *__restrict__ gate) {
    if (gated_off(gate)) return;
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) hist[v] = (uint8_t)(hist[v] << 1);
}

// Rows of last-sweep movers: one count per row entry at the neighbour.  The
// wave's mover rows are laid end to end (inclusive scan of their degrees) and
// its 64 lanes take 64 consecutive entries at a time; a lane finds the row of
// its entry by a binary search over the scan (shuffles), so a wave of short
// rows costs about (sum of degrees) / 64 steps instead of one per mover.
__global__ __launch_bounds__(256) void k_neighbor_moves(const int32_t *__restrict__ rowptr,
                                                        const int32_t *__restrict__ col,
                                                        const int32_t *__restrict__ list,
                                                        int32_t n,
                                                        const uint8_t *__restrict__ hist,
                                                        int32_t *nmv) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int32_t lo = 0, deg = 0;
    if (i < n) {
        const int32_t v = list[i];
        if (hist[v] & 1u) {
            lo = rowptr[v];
            const int32_t d = rowptr[v + 1] - lo;
            deg = d <= kNmvMaxDegree ? d : 0;
        }
    }
    int32_t inc = deg;  // at most 64 * kNmvMaxDegree
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t t = __shfl_up(inc, off, 64);
        if (lane >= off) inc += t;
    }
    const int32_t total = __shfl(inc, 63, 64);
    const int32_t pre = inc - deg;  // nondecreasing over the lanes
    for (int32_t base = 0; base < total; base += 64) {
        const int32_t t = base + lane;
        // the last lane whose rows start at or before entry t owns it
        int src = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
            const int c = src + step;
            if (__shfl(pre, c, 64) <= t) src = c;
        }
        const int32_t slo = __shfl(lo, src, 64), spre = __shfl(pre, src, 64);
        if (t < total) atomicAdd(&nmv[col[slo + (t - spre)]], 1);
    }
}

__global__ void k_features(FeatureArgs f, float *feat) {
    const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= f.V) return;
    float x[kFeat];
    features_of(f, v, x);
#pragma unroll
    for (int j = 0; j < kFeat; ++j) feat[(int64_t)v * kFeat + j] = x[j];
}

// Per-lane MLP operands (B fragments of W1^T and W2^T, biases, output layer).
struct MlpLane {
    bf16x8 w1b[2], w2b[2];
    float bias1[2], bias2[2], w3v[2];
    float b3;
};

__device__ __forceinline__ void load_mlp(const float *__restrict__ P, int col, int grp,
                                         MlpLane &m) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int n = col + 16 * t;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int kk = 8 * grp + j;
            m.w1b[t][j] = (__bf16)(kk < kFeat ? P[kOffW1 + n * kFeat + kk] : 0.0f);
            m.w2b[t][j] = (__bf16)P[kOffW2 + n * kHid + kk];
        }
        m.bias1[t] = P[kOffB1 + n];
        m.bias2[t] = P[kOffB2 + n];
        m.w3v[t] = P[kOffW3 + n];
    }
    m.b3 = P[kOffB3];
}

// Logits of the 16 vertices whose features are xs[row][0..8) (row < 16) of
// this wave; on return lanes with l%16 == 0 hold z of rows 4*(l/16) + r in
// z[r].  h1s is the wave's 16 x 33 transpose buffer.  Wave-synchronous: the
// LDS exchanges are ordered by the wave's own program order and a
// wave-level fence.
__device__ __forceinline__ void mlp_tile(const float (*xs)[kFeat + 1], float (*h1s)[kHid + 1],
                                         const MlpLane &m, int col, int grp, float z[4]) {
    bf16x8 a1;
#pragma unroll
    for (int j = 0; j < 8; ++j) a1[j] = (__bf16)(grp == 0 ? xs[col][j] : 0.0f);
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    const f32x4 c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, m.w1b[0], zero, 0, 0, 0);
    const f32x4 c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, m.w1b[1], zero, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        h1s[4 * grp + r][col] = fmaxf(c0[r] + m.bias1[0], 0.0f);
        h1s[4 * grp + r][col + 16] = fmaxf(c1[r] + m.bias1[1], 0.0f);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    bf16x8 a2;
#pragma unroll
    for (int j = 0; j < 8; ++j) a2[j] = (__bf16)h1s[col][8 * grp + j];
    const f32x4 d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, m.w2b[0], zero, 0, 0, 0);
    const f32x4 d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, m.w2b[1], zero, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r)
        z[r] = fmaxf(d0[r] + m.bias2[0], 0.0f) * m.w3v[0] + fmaxf(d1[r] + m.bias2[1], 0.0f) * m.w3v[1];
#pragma unroll
    for (int off = 1; off < 16; off <<= 1)
#pragma unroll
        for (int r = 0; r < 4; ++r) z[r] += __shfl_xor(z[r], off, 64);
#pragma unroll
    for (int r = 0; r < 4; ++r) z[r] += m.b3;
    // h1s is rewritten by the next tile: keep every lane's reads before that
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

// Each wave handles 64 vertices per iteration: features to LDS, then four
// 16-vertex MFMA tiles.
template <bool FROM_FEAT>
__global__ __launch_bounds__(256) void k_mlp(FeatureArgs f, const int32_t *__restrict__ list,
                                             const float *__restrict__ feat,
                                             int64_t n, const float *__restrict__ P,
                                             float threshold, float heavy_threshold,
                                             int32_t heavy_degree, int32_t and_nbr,
                                             BucketBounds bb, int8_t *keys,
                                             uint8_t *active, float *prob) {
    __shared__ float xs[4][64][kFeat + 1];
    __shared__ float h1s[4][16][kHid + 1];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int col = lane & 15, grp = lane >> 4;
    MlpLane m;
    load_mlp(P, col, grp, m);
    const int64_t wstride = (int64_t)gridDim.x * 256;
    for (int64_t base = (int64_t)blockIdx.x * 256 + 64 * wid; base < n; base += wstride) {
        const int64_t vl = base + lane;
        float x[kFeat];
        if (vl < n) {
            if (FROM_FEAT) {
#pragma unroll
                for (int j = 0; j < kFeat; ++j) x[j] = feat[vl * kFeat + j];
            } else {
                features_of(f, list ? list[vl] : (int32_t)vl, x);
            }
        } else {
#pragma unroll
            for (int j = 0; j < kFeat; ++j) x[j] = 0.0f;
        }
#pragma unroll
        for (int j = 0; j < kFeat; ++j) xs[wid][lane][j] = x[j];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int q = 0; q < 4; ++q) {
            float z[4];
            mlp_tile(&xs[wid][16 * q], h1s[wid], m, col, grp, z);
            if (col == 0) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t i = base + 16 * q + 4 * grp + r;
                    if (i >= n) continue;
                    const float p = sigmoid(z[r]);
                    const int64_t v = (!FROM_FEAT && list) ? (int64_t)list[i] : i;
                    if (prob) prob[v] = p;
                    if (!FROM_FEAT) {
                        const int32_t deg = f.rowptr[v + 1] - f.rowptr[v];
                        const float thr = heavy_degree > 0 && deg > heavy_degree
                                              ? heavy_threshold
                                              : threshold;
                        const bool act = deg > 0 && p > thr &&
                                         (!and_nbr || (f.hist[v] & 1u) || f.nmv[v] > 0);
                        active[v] = act ? 1 : 0;
                        keys[v] = act ? (int8_t)bucket_of(deg, bb) : (int8_t)-1;
                    }
                }
            }
        }
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

__global__ void k_frontier_rule(int kind, FeatureArgs f, const int32_t *__restrict__ list,
                                int32_t n, float p, uint64_t seed, BucketBounds bb, int8_t *keys,
                                uint8_t *active) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t v = list ? list[i] : i;
    const int32_t deg = f.rowptr[v + 1] - f.rowptr[v];
    bool act = deg > 0;
    if (kind == kFrontierNeighbor) {
        act = act && ((f.hist[v] & 1u) || f.nmv[v] > 0);
    } else if (kind == kFrontierRandom) {
        // drawn per vertex id of the caller's numbering
        const int32_t id = f.ref_of ? f.ref_of[v] : v;
        const uint64_t r = mix64(seed ^ mix64((uint64_t)(uint32_t)id));
        act = act && (float)(r >> 40) * (1.0f / 16777216.0f) < p;
    }
    active[v] = act ? 1 : 0;
    keys[v] = act ? (int8_t)bucket_of(deg, bb) : (int8_t)-1;
}

// ------------------------------------------------------------- training

constexpr int kTrainBlock = 128;

__device__ __forceinline__ float softplus(float x) {
    return fmaxf(x, 0.0f) + log1pf(expf(-fabsf(x)));
}

// Forward + backward of weighted binary cross-entropy for 128 samples per
// block (fp32, the master weights); the block's gradient sums and loss sum
// go to its partial row, so the reduction order is fixed.
__global__ __launch_bounds__(kTrainBlock) void k_train_grad(const float *__restrict__ feat,
                                                            const float *__restrict__ labels,
                                                            int64_t n,
                                                            const float *__restrict__ params,
                                                            float pw, float *partials) {
    __shared__ float P[kParams];
    __shared__ float X[kTrainBlock][kFeat + 1];
    __shared__ float H1[kTrainBlock][kHid + 1];
    __shared__ float H2[kTrainBlock][kHid + 1];
    __shared__ float D1[kTrainBlock][kHid + 1];
    __shared__ float D2[kTrainBlock][kHid + 1];
    __shared__ float DZ[kTrainBlock];
    __shared__ float LS[kTrainBlock];
    const int t = threadIdx.x;
    for (int i = t; i < kParams; i += kTrainBlock) P[i] = params[i];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * kTrainBlock + t;
    const bool valid = i < n;
    float x[kFeat];
#pragma unroll
    for (int j = 0; j < kFeat; ++j) x[j] = valid ? feat[i * kFeat + j] : 0.0f;
    for (int a = 0; a < kHid; ++a) {
        float s = P[kOffB1 + a];
#pragma unroll
        for (int j = 0; j < kFeat; ++j) s += P[kOffW1 + a * kFeat + j] * x[j];
        H1[t][a] = fmaxf(s, 0.0f);
    }
    float z = P[kOffB3];
    for (int a = 0; a < kHid; ++a) {
        float s = P[kOffB2 + a];
        for (int j = 0; j < kHid; ++j) s += P[kOffW2 + a * kHid + j] * H1[t][j];
        const float h2 = fmaxf(s, 0.0f);
        H2[t][a] = h2;
        z += P[kOffW3 + a] * h2;
    }
    const float y = valid ? labels[i] : 0.0f;
    const float p = 1.0f / (1.0f + expf(-z));
    // L = -(pw*y*log p + (1-y)*log(1-p)),  dL/dz = pw*y*(p-1) + (1-y)*p
    const float loss = valid ? pw * y * softplus(-z) + (1.0f - y) * softplus(z) : 0.0f;
    const float dz = valid ? pw * y * (p - 1.0f) + (1.0f - y) * p : 0.0f;
    for (int a = 0; a < kHid; ++a) D2[t][a] = H2[t][a] > 0.0f ? dz * P[kOffW3 + a] : 0.0f;
    for (int j = 0; j < kHid; ++j) {
        float s = 0.0f;
        for (int a = 0; a < kHid; ++a) s += P[kOffW2 + a * kHid + j] * D2[t][a];
        D1[t][j] = H1[t][j] > 0.0f ? s : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < kFeat; ++j) X[t][j] = x[j];
    DZ[t] = dz;
    LS[t] = loss;
    __syncthreads();
    float *out = partials + (int64_t)blockIdx.x * (kParams + 1);
    for (int pi = t; pi <= kParams; pi += kTrainBlock) {
        float s = 0.0f;
        if (pi < kOffB1) {
            const int a = (pi - kOffW1) / kFeat, j = (pi - kOffW1) % kFeat;
            for (int q = 0; q < kTrainBlock; ++q) s += D1[q][a] * X[q][j];
        } else if (pi < kOffW2) {
            const int a = pi - kOffB1;
            for (int q = 0; q < kTrainBlock; ++q) s += D1[q][a];
        } else if (pi < kOffB2) {
            const int a = (pi - kOffW2) / kHid, j = (pi - kOffW2) % kHid;
            for (int q = 0; q < kTrainBlock; ++q) s += D2[q][a] * H1[q][j];
        } else if (pi < kOffW3) {
            const int a = pi - kOffB2;
            for (int q = 0; q < kTrainBlock; ++q) s += D2[q][a];
        } else if (pi < kOffB3) {
            const int a = pi - kOffW3;
            for (int q = 0; q < kTrainBlock; ++q) s += DZ[q] * H2[q][a];
        } else if (pi == kOffB3) {
            for (int q = 0; q < kTrainBlock; ++q) s += DZ[q];
        } else {
            for (int q = 0; q < kTrainBlock; ++q) s += LS[q];
        }
        out[pi] = s;
    }
}

// Fixed-order FP64 sum of the block partials, divided by n.
__global__ void k_train_reduce(const float *__restrict__ partials, int64_t nblocks, int64_t n,
                               float *grads) {
    const int pi = blockIdx.x * blockDim.x + threadIdx.x;
    if (pi > kParams) return;
    double s = 0.0;
    for (int64_t b = 0; b < nblocks; ++b) s += (double)partials[b * (kParams + 1) + pi];
    grads[pi] = (float)(s / (double)(n > 0 ? n : 1));
}

// Adam (Kingma & Ba 2015) with bias correction; t counts from 1.
__global__ void k_adam(float *params, const float *__restrict__ grads, float *m, float *v,
                       int32_t t, float lr, float beta1, float beta2, float eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const float g = grads[i];
    const float mi = beta1 * m[i] + (1.0f - beta1) * g;
    const float vi = beta2 * v[i] + (1.0f - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - powf(beta1, (float)t));
    const float vh = vi / (1.0f - powf(beta2, (float)t));
    params[i] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void k_accumulate_loss(const float *__restrict__ grads, double weight, double *acc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *acc += weight * (double)grads[kParams];
}

}  // namespace

void launch_accumulate_loss(const float *grads, double weight, double *acc, hipStream_t s) {
    hipLaunchKernelGGL(k_accumulate_loss, dim3(1), dim3(64), 0, s, grads, weight, acc);
}

void launch_hist_shift(uint8_t *hist, int32_t V, hipStream_t s, const int32_t *gate) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_hist_shift, dim3(blocks_for(V, 256)), dim3(256), 0, s, hist, V, gate);
}

void launch_neighbor_moves(const int32_t *rowptr, const int32_t *col, const int32_t *list,
                           int32_t n, const uint8_t *hist, int32_t *nmv, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_neighbor_moves, dim3(blocks_for(n, 256)), dim3(256), 0, s, rowptr, col,
                       list, n, hist, nmv);
}

void launch_isolated_features(const FeatureArgs &f, float *out, hipStream_t s) {
    hipLaunchKernelGGL(k_isolated_features, dim3(1), dim3(64), 0, s, f.M, f.vscale, out);
}

void launch_features(const FeatureArgs &f, float *feat, hipStream_t s) {
    if (f.V <= 0) return;
    hipLaunchKernelGGL(k_features, dim3(blocks_for(f.V, 256)), dim3(256), 0, s, f, feat);
}

void launch_frontier(const FeatureArgs &f, const int32_t *list, int32_t n, const float *params,
                     float threshold, float heavy_threshold, int32_t heavy_degree,
                     BucketBounds bb, int8_t *keys, uint8_t *active, float *prob, hipStream_t s,
                     int32_t and_neighbor) {
    if (!list) n = f.V;
    if (n <= 0) return;
    int64_t g = blocks_for(n, 256);
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_mlp<false>, dim3((unsigned)g), dim3(256), 0, s, f, list, nullptr,
                       (int64_t)n, params, threshold, heavy_threshold, heavy_degree, and_neighbor,
                       bb, keys, active, prob);
}

void launch_graph_model_frontier(const int32_t *rowptr, const int32_t *list, int32_t n,
                                 const int32_t *gid, const uint8_t *gact, const uint8_t *gfull,
                                 BucketBounds bb, int8_t *keys, uint8_t *active, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_graph_model_frontier, dim3(blocks_for(n, 256)), dim3(256), 0, s, rowptr,
                       list, n, gid, gact, gfull, bb, keys, active);
}

void launch_frontier_rule(int kind, const FeatureArgs &f, const int32_t *list, int32_t n, float p,
                          uint64_t seed, BucketBounds bb, int8_t *keys, uint8_t *active,
                          hipStream_t s) {
    if (!list) n = f.V;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_frontier_rule, dim3(blocks_for(n, 256)), dim3(256), 0, s, kind, f, list, n,
                       p, seed, bb, keys, active);
}

void launch_mlp_forward(const float *feat, int64_t n, const float *params, float *prob,
                        hipStream_t s) {
    if (n <= 0) return;
    int64_t g = blocks_for(n, 256);
    if (g > 8192) g = 8192;
    FeatureArgs f{};
    hipLaunchKernelGGL(k_mlp<true>, dim3((unsigned)g), dim3(256), 0, s, f, nullptr, feat, n, params,
                       0.0f, 0.0f, 0, 0, BucketBounds{}, nullptr, nullptr, prob);
}

int64_t grad_blocks(int64_t n) { return (n + kTrainBlock - 1) / kTrainBlock; }

void launch_train_grad(const float *feat, const float *labels, int64_t n, const float *params,
                       float pos_weight, float *partials, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_train_grad, dim3((unsigned)grad_blocks(n)), dim3(kTrainBlock), 0, s, feat,
                       labels, n, params, pos_weight, partials);
}

void launch_train_reduce(const float *partials, int64_t nblocks, int64_t n, float *grads,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_train_reduce, dim3(blocks_for(kParams + 1, 256)), dim3(256), 0, s, partials,
                       nblocks, n, grads);
}

void launch_adam(float *params, const float *grads, float *m, float *v, int32_t t, float lr,
                 float beta1, float beta2, float eps, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(kParams, 256)), dim3(256), 0, s, params, grads, m, v,
                       t, lr, beta1, beta2, eps);
}

}  // namespace lv
