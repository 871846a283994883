// Host-side input conditioning of the Louvain library: the weight grid and
// the symmetry check (OpenMP; no device code).
//
// Weight grid.  The reference sums fp32 weights with fp32 atomics
// (modularity_optimisation.cu:144-153, community_aggregation.cu:146-159), so
// its sums depend on the order in which threads land.  This library sums in
// FP64, and makes every such sum exact, hence independent of order, by first
// rounding each weight to a grid of 2^-s:
//   T = sum of |w| over the entries, FP64, in chunks of kGridChunk entries
//       (sequential inside a chunk, chunk sums added in chunk order), so T
//       does not depend on the thread count;
//   c = the smallest integer with 2^c >= T * (1 + 2^-20);
//   s = 50 - c;
//   q(w) = rint(w * 2^s) * 2^-s (FP64, ties to even), stored as fp32.
// q(w) is exact in fp32: below 2^24 grid steps it has at most 24 significant
// bits, and above it w's own ulp is already a multiple of 2^-s, so q(w) = w.
// Every quantity the engine accumulates is a signed sum of rounded weights
// whose partial sums stay below 8T (vertex strengths, community totals while
// joins and leaves land, candidate sums, merged weights, the weight inside
// communities and its per-sweep corrections), i.e. an integer multiple of
// 2^-s below 2^53 of them, which FP64 represents exactly.  Merged weights
// rounded to fp32 stay on the grid, so later levels need no new rounding.
// Integer weights and weights on a coarser grid (the generators' 2^-24 up to
// T = 2^26) are left unchanged.
//
// Symmetry.  The engine needs every entry (v, u, w) matched by an entry
// (u, v, w): candidate bounds use S_c >= K_c and the incremental modularity
// counts each entry from both ends.  The reference's reader always
// symmetrises (utils.cu:32-45).  The check compares the multisets of
// (row, column, weight bits) and (column, row, weight bits) through a sum of
// 64-bit mixes of each triple (commutative, so the row order does not
// matter); an asymmetric input passes only on a 64-bit hash collision.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

#include "../../include/louvain.h"
#include "lv_error.h"
#include "lv_weights.h"

namespace {

constexpr int64_t kGridChunk = 65536;

lv_status wfail(lv_status st, const char *m) {
    lv::set_error(m);
    return st;
}

inline uint64_t mix64(uint64_t z) {
    z ^= z >> 33;
    z *= 0xff51afd7ed558ccdull;
    z ^= z >> 33;
    z *= 0xc4ceb9fe1a85ec53ull;
    z ^= z >> 33;
    return z;
}

inline uint64_t triple_hash(uint32_t a, uint32_t b, uint32_t wbits) {
    const uint64_t k = ((uint64_t)a << 32) | b;
    return mix64(mix64(k ^ 0x9e3779b97f4a7c15ull) ^ ((uint64_t)wbits * 0x2545f4914f6cdd1dull));
}

}  // namespace

namespace lv {

double weight_total(const float *w, int64_t n) {
    const int64_t nc = (n + kGridChunk - 1) / kGridChunk;
    std::vector<double> part((size_t)std::max<int64_t>(nc, 1), 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c) {
        double s = 0.0;
        const int64_t hi = std::min(n, (c + 1) * kGridChunk);
        for (int64_t e = c * kGridChunk; e < hi; ++e) s += std::fabs((double)w[e]);
        part[(size_t)c] = s;
    }
    double t = 0.0;
    for (int64_t c = 0; c < nc; ++c) t += part[(size_t)c];
    return t;
}

int32_t grid_shift(double T) {
    if (!(T > 0.0)) return 0;
    const double t1 = T * (1.0 + 0x1p-20);
    int e = 0;
    const double m = std::frexp(t1, &e);  // t1 = m * 2^e, m in [0.5, 1)
    const int c = (m == 0.5) ? e - 1 : e;
    return 50 - c;
}

float quantize(float w, int32_t shift) {
    const double x = std::ldexp((double)w, shift);
    return (float)(std::ldexp(std::rint(x), -shift));
}

int64_t quantize_into(const float *w, int64_t n, int32_t shift, float *out) {
    int64_t changed = 0;
#pragma omp parallel for schedule(static) reduction(+ : changed)
    for (int64_t e = 0; e < n; ++e) {
        const float q = quantize(w[e], shift);
        uint32_t a, b;
        std::memcpy(&a, &w[e], 4);
        std::memcpy(&b, &q, 4);
        changed += a != b ? 1 : 0;
        if (out) out[e] = q;
    }
    return changed;
}

bool csr_symmetric(const int32_t *rowptr, const int32_t *col, const float *w, int32_t V) {
    uint64_t fwd = 0, rev = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : fwd, rev)
    for (int32_t v = 0; v < V; ++v) {
        for (int32_t e = rowptr[v]; e < rowptr[v + 1]; ++e) {
            uint32_t wb;
            std::memcpy(&wb, &w[e], 4);
            fwd += triple_hash((uint32_t)v, (uint32_t)col[e], wb);
            rev += triple_hash((uint32_t)col[e], (uint32_t)v, wb);
        }
    }
    return fwd == rev;
}

}  // namespace lv

extern "C" {

lv_status lv_weight_grid(const float *w, int64_t n, int32_t *shift, int64_t *changed) {
    if (n < 0 || (n > 0 && !w) || !shift) return wfail(LV_ERR_ARG, "bad argument");
    const double T = lv::weight_total(w, n);
    if (!std::isfinite(T)) return wfail(LV_ERR_ARG, "a weight is not finite");
    *shift = lv::grid_shift(T);
    if (changed) *changed = T > 0.0 ? lv::quantize_into(w, n, *shift, nullptr) : 0;
    return LV_OK;
}

lv_status lv_quantize_weights(const float *w, int64_t n, int32_t shift, float *out) {
    if (n < 0 || (n > 0 && (!w || !out))) return wfail(LV_ERR_ARG, "bad argument");
    if (shift < -1100 || shift > 1100) return wfail(LV_ERR_ARG, "shift out of range");
    (void)lv::quantize_into(w, n, shift, out);
    return LV_OK;
}

}  // extern "C"
