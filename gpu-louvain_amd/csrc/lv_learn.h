// Learned frontier: per-vertex features, the BF16-MFMA inference of the
// move predictor (8 -> 32 -> 32 -> 1, ReLU, sigmoid), rule-based frontiers
// for comparison, and the training kernels (forward + backward with a
// fixed-order gradient reduction, Adam).
#pragma once

#include "lv_common.h"
#include "lv_phase1.h"

namespace lv {

constexpr int kFeat = 8;
constexpr int kHid = 32;
// Parameter layout (fp32 master copy): W1[32][8], b1[32], W2[32][32], b2[32], w3[32], b3.
// W[a][j] multiplies input j of output unit a.
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * kFeat;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid;
constexpr int kParams = kOffB3 + 1;  // 1377

// Inputs of the feature function (state after the last sweep).
struct FeatureArgs {
    const int32_t *rowptr;
    const double *k;
    const int32_t *comm;
    const uint64_t *crec;
    const float *margin;  // last evaluation: bestGain - K[cur]/M
    const float *kown;    // last evaluation: K[cur]
    const uint8_t *hist;  // move history, bit 0 = most recent sweep
    const int32_t *nmv;   // row entries whose neighbour moved in the last sweep
    int32_t V;
    float M;
    float vscale;         // V / (2M), V counting the vertices without edges
    const int32_t *ref_of;  // stored -> caller's vertex id (null: the same)
    // batches: gid[v] = graph of v, whose M (fp32) and V/(2M) replace M and
    // vscale (null for a single graph)
    const int32_t *gid;
    const float *gM;
    const float *gvscale;
};

// Frontier selection rules.
enum FrontierKind { kFrontierModel = 0, kFrontierNeighbor = 1, kFrontierRandom = 2, kFrontierAll = 3 };

// Movers of degree above this do not count in nmv: a hub's move says little
// about any one neighbour and its row would dominate the pass.
constexpr int32_t kNmvMaxDegree = 1024;

// hist[v] <<= 1 for every vertex (start of a sweep).
void launch_hist_shift(uint8_t *hist, int32_t V, hipStream_t s, const int32_t *gate = nullptr);
// nmv[u] += 1 for every row entry (u, v) of a vertex v that moved in the last
// sweep (hist bit 0) and has degree <= kNmvMaxDegree; nmv must be zero
// before.  The n vertices of list are scanned 64 per wave; the wave then
// walks each mover's row in turn, one entry per lane.
void launch_neighbor_moves(const int32_t *rowptr, const int32_t *col, const int32_t *list,
                           int32_t n, const uint8_t *hist, int32_t *nmv, hipStream_t s);
// out[0..8) = the features of a vertex without edges (the same for all such
// vertices of a level).
void launch_isolated_features(const FeatureArgs &f, float *out, hipStream_t s);
// feat[v*8 .. v*8+8) for every vertex.
void launch_features(const FeatureArgs &f, float *feat, hipStream_t s);
// Frontier from the model: active[v] = deg > 0 and sigmoid(MLP(features)) >
// threshold; keys[v] = execution bucket of an active vertex, -1 otherwise;
// prob (optional) = the predicted probability.  Vertices of degree above
// heavy_degree (when positive) use heavy_threshold instead.  Over the n
// vertices of list, or over all f.V vertices when list is null (entries of
// vertices outside the list are left as they are).
// and_neighbor: a vertex is active only when the neighbour rule also selects
// it (it or a neighbour moved in the last sweep).
void launch_frontier(const FeatureArgs &f, const int32_t *list, int32_t n, const float *params,
                     float threshold, float heavy_threshold, int32_t heavy_degree,
                     BucketBounds bb, int8_t *keys, uint8_t *active, float *prob, hipStream_t s,
                     int32_t and_neighbor = 0);
// Batches: the vertices of list[0..n) of the active graphs (gact) keep the
// frontier flags the model set when their graph's sweep is pruned, and are
// all in it when their graph's sweep is full (gfull); the others leave it.
void launch_graph_model_frontier(const int32_t *rowptr, const int32_t *list, int32_t n,
                                 const int32_t *gid, const uint8_t *gact, const uint8_t *gfull,
                                 BucketBounds bb, int8_t *keys, uint8_t *active, hipStream_t s);
// Frontier from a rule: kFrontierNeighbor = moved in the last sweep or a
// neighbour did; kFrontierRandom = seeded Bernoulli(p) per vertex and sweep;
// kFrontierAll = every vertex of positive degree.  Over the n vertices of
// list (null: all f.V vertices).
void launch_frontier_rule(int kind, const FeatureArgs &f, const int32_t *list, int32_t n, float p,
                          uint64_t seed, BucketBounds bb, int8_t *keys, uint8_t *active,
                          hipStream_t s);
// prob[i] = sigmoid(MLP(feat[i*8 .. i*8+8))) on MFMA (bf16 operands, fp32 accumulation).
void launch_mlp_forward(const float *feat, int64_t n, const float *params, float *prob,
                        hipStream_t s);

// Training.  partials: grad_blocks(n) x (kParams + 1) floats (last = loss).
int64_t grad_blocks(int64_t n);
void launch_train_grad(const float *feat, const float *labels, int64_t n, const float *params,
                       float pos_weight, float *partials, hipStream_t s);
// grads[0..kParams) = mean gradient, grads[kParams] = mean loss (fixed order).
void launch_train_reduce(const float *partials, int64_t nblocks, int64_t n, float *grads,
                         hipStream_t s);
// *acc += weight * grads[kParams] (loss bookkeeping of an epoch).
void launch_accumulate_loss(const float *grads, double weight, double *acc, hipStream_t s);
void launch_adam(float *params, const float *grads, float *m, float *v, int32_t t, float lr,
                 float beta1, float beta2, float eps, hipStream_t s);

}  // namespace lv
