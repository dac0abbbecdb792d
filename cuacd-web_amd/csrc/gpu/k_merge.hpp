// k_merge.hpp — the merge pass's cheap pair rejections over every pair of several parts
// lists in one launch (src/merge.ts:192-204); the matching stays on the host.
#pragma once
#include <hip/hip_runtime.h>

#include "../host/cost.hpp"

namespace dk {

// a workgroup per list: its threads take rows i (every pair (i, j > i) of the row), writing one byte
// per pair at the list's base plus merge_row(n, i) + j - i - 1, the host's order
__global__ __launch_bounds__(256) void k_merge_pairs(const cuacd::MergePart* __restrict__ parts,
                                                     const uint32_t* __restrict__ off,
                                                     const unsigned long long* __restrict__ pairBase, double threshold,
                                                     double halfT2, uint8_t* __restrict__ pass) {
  const uint32_t l = blockIdx.x;
  const uint32_t p0 = off[l], n = off[l + 1] - p0;
  uint8_t* out = pass + pairBase[l];
  for (uint32_t i = threadIdx.x; i + 1 < n; i += blockDim.x) {
    const cuacd::MergePart a = parts[p0 + i];
    uint8_t* row = out + cuacd::merge_row(n, i) - i - 1;
    for (uint32_t j = i + 1; j < n; j++) row[j] = cuacd::merge_pair_passes(a, parts[p0 + j], threshold, halfT2) ? 1 : 0;
  }
}

}  // namespace dk
