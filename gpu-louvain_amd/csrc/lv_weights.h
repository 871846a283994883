// Host-side weight grid and symmetry check (lv_weights.cpp).
#pragma once

#include <stdint.h>

namespace lv {
// Sum of |w| in fixed chunks (independent of the thread count).
double weight_total(const float *w, int64_t n);
// Grid exponent s for a total T (0 when T is 0).
int32_t grid_shift(double T);
// Rounds n weights to the 2^-shift grid into out (may be null); returns how
// many changed.
int64_t quantize_into(const float *w, int64_t n, int32_t shift, float *out);
// Every entry (v, u, w) has a matching (u, v, w) (multiset hash).
bool csr_symmetric(const int32_t *rowptr, const int32_t *col, const float *w, int32_t V);
}  // namespace lv
