// Finishing a gbl reduction from per-workgroup partials, shared by the eager
// finish kernels and the graph's history kernel so that both produce the same
// bits.  One group of kRedThreads (4 waves) reduces one gbl:
//   thread t combines partials t, t + 256, t + 512, ... in that order
//   (independent loads, issued together), then a wave tree by xor shuffles,
//   then thread 0 of the group combines its 4 waves in order.
// Kinds: 1 sum of doubles; 2 max and 3 min of (value, loc) pairs, equal
// values keeping the lower loc (the first element reaching the extremum).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace op2h_red {

constexpr int kRedThreads = 256;
constexpr int kRedMaxLoads = 16;  // partials per thread (4096 per gbl)

struct ValLoc {
  double v;
  int loc;
};

__device__ __forceinline__ ValLoc combine(ValLoc a, ValLoc b, int kind) {
  if (kind == 1) return ValLoc{a.v + b.v, 0};
  const bool better = kind == 2 ? b.v > a.v : b.v < a.v;
  if (better || (b.v == a.v && b.loc < a.loc)) return b;
  return a;
}

__device__ __forceinline__ ValLoc identity(int kind) {
  return ValLoc{kind == 1 ? 0.0 : kind == 2 ? -INFINITY : INFINITY, 0x7fffffff};
}

// Reduces `n` partials (vals, and locs for kinds 2/3) with the group whose
// threads have local index `t` (0..kRedThreads-1) and wave slots s_v/s_l
// [4] in LDS.  All threads of the group must call it; the result is returned
// to t == 0.  The caller synchronises the workgroup between uses of s_v/s_l.
__device__ __forceinline__ ValLoc group_reduce(const double *vals, const int *locs, int n,
                                               int kind, int t, double *s_v, int *s_l) {
  ValLoc m = identity(kind);
  double v[kRedMaxLoads];
  int l[kRedMaxLoads];
#pragma unroll
  for (int i = 0; i < kRedMaxLoads; ++i) {
    const int p = t + i * kRedThreads;
    v[i] = p < n ? vals[p] : identity(kind).v;
    l[i] = (kind != 1 && p < n) ? locs[p] : 0x7fffffff;
  }
#pragma unroll
  for (int i = 0; i < kRedMaxLoads; ++i) m = combine(m, ValLoc{v[i], l[i]}, kind);
  for (int o = 32; o > 0; o >>= 1) {
    ValLoc x{__shfl_xor(m.v, o, 64), __shfl_xor(m.loc, o, 64)};
    m = combine(m, x, kind);
  }
  const int lane = t & 63, wave = t >> 6;
  if (lane == 0) {
    s_v[wave] = m.v;
    s_l[wave] = m.loc;
  }
  __syncthreads();
  ValLoc r = identity(kind);
  if (t == 0)
    for (int w = 0; w < kRedThreads / 64; ++w) r = combine(r, ValLoc{s_v[w], s_l[w]}, kind);
  return r;
}

}  // namespace op2h_red
