// Device-wide primitives written for 64-lane wavefronts: exclusive scan and
// stable key partition (used for degree bucketing, renumbering and CSR
// offsets).  Host launchers enqueue on the given stream and never sync.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lv {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048 items per tile

// Scratch (in int32 elements) needed by exclusive_scan for n items.
int64_t scan_scratch_elems(int64_t n);
// out[i] = sum_{j<i} in[i]; *total = sum of all (device pointer, may be null).
// in == out allowed.  scratch: scan_scratch_elems(n) int32 elements.
void exclusive_scan(const int32_t *in, int32_t *out, int64_t n, int32_t *total,
                    int32_t *scratch, hipStream_t s);

// Stable partition of the indices 0..n-1 by key (keys in [0,nk) kept, keys
// < 0 dropped): items[] lists indices grouped by key in ascending key, and in
// ascending index within a key; offsets[0..nk] (device) the group starts.
// scratch: partition_scratch_elems(n, nk) int32 elements.
constexpr int kPartTile = 256;
int64_t partition_scratch_elems(int64_t n, int nk);
void stable_partition(const int8_t *keys, int32_t n, int nk, int32_t *items, int32_t *offsets,
                      int32_t *scratch, hipStream_t s);

// Fill helpers.
void fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s);
void fill_u8(uint8_t *p, uint8_t v, int64_t n, hipStream_t s);
void fill_f64(double *p, double v, int64_t n, hipStream_t s);
void iota_i32(int32_t *p, int64_t n, hipStream_t s);

}  // namespace lv
