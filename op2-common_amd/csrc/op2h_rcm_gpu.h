// Reverse Cuthill-McKee on the GPU (internal): the ordering of op2h_rcm.h,
// bit for bit, built level by level on `stream`.  Neighbour lists must be
// sorted by index and hold no duplicates.  two_pass forces the two-pass level
// sort that is otherwise taken only when the packed 64-bit key is too narrow.
// Returns new_of_old.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace op2h {
std::vector<int> rcm_order_gpu(int n, const std::vector<int> &off,
                               const std::vector<int> &adj, hipStream_t stream,
                               bool two_pass = false);
}
