// Device-side plan colouring.
//
// Flat element colouring and block colouring: speculative first-fit in rounds.
// Each round every uncoloured item takes the lowest colour absent from the
// 64-bit colour masks of its increment targets, claims each target with the
// key (colour << 32 | h(item)), h a bijective 32-bit hash, by atomicMin, and
// commits only if it holds every one of its targets.  Two items committed in
// one round therefore never share a target, and the item with the smallest key
// always commits, so every round makes progress; hashed priorities keep the
// number of rounds logarithmic on meshes numbered in any order.  Colours are
// limited to 64 per plan (the host builder is used beyond that).
//
// Thread colouring inside plan blocks: one workgroup per block, one lane walks
// the block's elements in order with the colour masks of the block's local
// targets in LDS — the reference's greedy first-fit
// (op2/src/core/op_rt_support.cpp:684-754), so thread colours equal the host
// builder's whenever a block needs at most 32 colours.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "op2h_plan_build.h"

#define GCHECK(x)                                                            \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      printf("op2hip: %s failed at %s:%d: %s\n", #x, __FILE__, __LINE__,     \
             hipGetErrorString(e_));                                         \
      exit(-1);                                                              \
    }                                                                        \
  } while (0)

namespace op2h {
namespace {

constexpr int kMaxConflictArgs = 8;
constexpr int kMaxRounds = 4096;
typedef unsigned long long u64;

struct ConflictArgs {
  const int *map[kMaxConflictArgs];
  int dim[kMaxConflictArgs];
  int idx[kMaxConflictArgs];
  u64 *mask[kMaxConflictArgs];  // per arg: masks of its dat's targets
  u64 *win[kMaxConflictArgs];   // per arg: claims on its dat's targets
  int n;
};

__device__ __forceinline__ unsigned mix32(unsigned x) {
  // xorshift-multiply rounds: each step is invertible, so the hash is a
  // permutation of 32-bit integers and keys of distinct items never tie
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

__global__ void k_fill64(u64 *p, size_t n, u64 v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill32(int *p, size_t n, int v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---- flat element colouring ----

__global__ void k_elem_tentative(int n, ConflictArgs A, const int *col, int *tent,
                                 int *err) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n || col[e] >= 0) return;
  u64 m = 0;
  for (int a = 0; a < A.n; ++a) m |= A.mask[a][A.map[a][(size_t)e * A.dim[a] + A.idx[a]]];
  int c = __builtin_ffsll((long long)~m) - 1;
  if (c < 0) atomicOr(err, 1);
  tent[e] = c;
}

__global__ void k_elem_claim(int n, ConflictArgs A, const int *col, const int *tent) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n || col[e] >= 0 || tent[e] < 0) return;
  u64 key = ((u64)tent[e] << 32) | mix32((unsigned)e);
  for (int a = 0; a < A.n; ++a)
    atomicMin(&A.win[a][A.map[a][(size_t)e * A.dim[a] + A.idx[a]]], key);
}

__global__ void k_elem_commit(int n, ConflictArgs A, int *col, const int *tent,
                              int *remaining, int *ncolors) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n || col[e] >= 0) return;
  int c = tent[e];
  bool ok = c >= 0;
  if (ok) {
    u64 key = ((u64)c << 32) | mix32((unsigned)e);
    for (int a = 0; a < A.n; ++a)
      ok = ok && A.win[a][A.map[a][(size_t)e * A.dim[a] + A.idx[a]]] == key;
  }
  if (!ok) {
    atomicAdd(remaining, 1);
    return;
  }
  col[e] = c;
  atomicMax(ncolors, c + 1);
  for (int a = 0; a < A.n; ++a)
    atomicOr(&A.mask[a][A.map[a][(size_t)e * A.dim[a] + A.idx[a]]], 1ull << c);
}

// ---- per-block thread colouring (sequential greedy, one lane per block) ----

struct LocalArgs {
  const unsigned short *loc[kMaxConflictArgs];
  int lds_off[kMaxConflictArgs];  // per arg: offset of its dat's masks in LDS
  int n;
  const int *ind_sizes[kMaxConflictArgs];  // per arg: unique targets per block
};

__global__ void k_thread_colours(const int *offset, const int *nelems, LocalArgs L,
                                 int lds_words, unsigned char *thrcol,
                                 int *nthrcol, int *err) {
  extern __shared__ u64 lm[];
  const int b = blockIdx.x;
  for (int k = threadIdx.x; k < lds_words; k += blockDim.x) lm[k] = 0;
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int lo = offset[b], hi = lo + nelems[b];
  int ncol = 0;
  for (int e = lo; e < hi; ++e) {
    u64 m = 0;
    for (int a = 0; a < L.n; ++a) m |= lm[L.lds_off[a] + L.loc[a][e]];
    int c = __builtin_ffsll((long long)~m) - 1;
    if (c < 0 || c > 254) {
      atomicOr(err, 1);
      c = 0;
    }
    thrcol[e] = (unsigned char)c;
    ncol = c + 1 > ncol ? c + 1 : ncol;
    for (int a = 0; a < L.n; ++a) lm[L.lds_off[a] + L.loc[a][e]] |= 1ull << c;
  }
  nthrcol[b] = ncol;
}

// ---- block colouring over the blocks' unique target lists ----

struct BlockTargets {
  const int *ind_map[kMaxConflictArgs];
  const int *ind_offs[kMaxConflictArgs];
  const int *ind_sizes[kMaxConflictArgs];
  u64 *mask[kMaxConflictArgs];
  u64 *win[kMaxConflictArgs];
  int n;
};

__device__ __forceinline__ u64 wg_or(u64 v, u64 *sh) {
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  u64 r = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r |= sh[i];
  __syncthreads();
  return r;
}

__global__ void k_blk_tentative(BlockTargets T, const int *blkcol, int *tent,
                                int *err) {
  __shared__ u64 sh[16];
  const int b = blockIdx.x;
  if (blkcol[b] >= 0) return;  // uniform across the workgroup
  u64 m = 0;
  for (int a = 0; a < T.n; ++a) {
    int base = T.ind_offs[a][b], cnt = T.ind_sizes[a][b];
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) m |= T.mask[a][T.ind_map[a][base + k]];
  }
  m = wg_or(m, sh);
  if (threadIdx.x == 0) {
    int c = __builtin_ffsll((long long)~m) - 1;
    if (c < 0) atomicOr(err, 1);
    tent[b] = c;
  }
}

__global__ void k_blk_claim(BlockTargets T, const int *blkcol, const int *tent) {
  const int b = blockIdx.x;
  if (blkcol[b] >= 0 || tent[b] < 0) return;
  u64 key = ((u64)tent[b] << 32) | mix32((unsigned)b);
  for (int a = 0; a < T.n; ++a) {
    int base = T.ind_offs[a][b], cnt = T.ind_sizes[a][b];
    for (int k = threadIdx.x; k < cnt; k += blockDim.x)
      atomicMin(&T.win[a][T.ind_map[a][base + k]], key);
  }
}

__global__ void k_blk_commit(BlockTargets T, int *blkcol, const int *tent,
                             int *remaining, int *ncolors) {
  __shared__ u64 sh[16];
  const int b = blockIdx.x;
  if (blkcol[b] >= 0) return;
  const int c = tent[b];
  u64 lost = c < 0 ? 1 : 0;
  if (c >= 0) {
    u64 key = ((u64)c << 32) | mix32((unsigned)b);
    for (int a = 0; a < T.n; ++a) {
      int base = T.ind_offs[a][b], cnt = T.ind_sizes[a][b];
      for (int k = threadIdx.x; k < cnt; k += blockDim.x)
        if (T.win[a][T.ind_map[a][base + k]] != key) lost = 1;
    }
  }
  lost = wg_or(lost, sh);
  if (lost) {
    if (threadIdx.x == 0) atomicAdd(remaining, 1);
    return;
  }
  for (int a = 0; a < T.n; ++a) {
    int base = T.ind_offs[a][b], cnt = T.ind_sizes[a][b];
    for (int k = threadIdx.x; k < cnt; k += blockDim.x)
      atomicOr(&T.mask[a][T.ind_map[a][base + k]], 1ull << c);
  }
  if (threadIdx.x == 0) {
    blkcol[b] = c;
    atomicMax(ncolors, c + 1);
  }
}

int grid(size_t n, int bs) { return (int)((n + bs - 1) / bs); }

struct Scratch {
  std::vector<void *> p;
  template <class T> T *get(size_t n) {
    void *q = nullptr;
    GCHECK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    p.push_back(q);
    return (T *)q;
  }
  ~Scratch() {
    for (void *q : p) (void)hipFree(q);
  }
};

// Conflict args of the loop; one mask/claim array per conflicting dat.
bool conflict_args(const LoopShape &sh, std::vector<int> &args,
                   std::vector<int> &dat_of_arg) {
  args.clear();
  dat_of_arg.clear();
  std::vector<std::pair<int, op_map>> dat_map;  // a dat reached through two maps
  for (int m = 0; m < sh.nargs; ++m) {
    if (!sh.conflict[m]) continue;
    args.push_back(m);
    dat_of_arg.push_back(sh.dat_slot[m]);
    for (auto &dm : dat_map)
      if (dm.first == sh.dat_slot[m] && dm.second != sh.arg_map[m].first) return false;
    dat_map.push_back({sh.dat_slot[m], sh.arg_map[m].first});
  }
  return (int)args.size() <= kMaxConflictArgs;
}

}  // namespace

int gpu_colour_elements(const LoopShape &sh, int n, int *col_d,
                        hipStream_t stream) {
  std::vector<int> cargs, cdat;
  if (!conflict_args(sh, cargs, cdat)) return -1;
  Scratch S;
  ConflictArgs A{};
  A.n = (int)cargs.size();
  std::vector<u64 *> masks(sh.dat_to_size.size(), nullptr), wins(sh.dat_to_size.size(), nullptr);
  for (int i = 0; i < A.n; ++i) {
    int m = cargs[i], d = cdat[i];
    if (!masks[d]) {
      masks[d] = S.get<u64>(sh.dat_to_size[d]);
      wins[d] = S.get<u64>(sh.dat_to_size[d]);
      k_fill64<<<grid(sh.dat_to_size[d], 256), 256, 0, stream>>>(masks[d], sh.dat_to_size[d], 0ull);
    }
    A.map[i] = sh.arg_map[m].first->map_d;
    A.dim[i] = sh.arg_map[m].first->dim;
    A.idx[i] = sh.arg_map[m].second;
    A.mask[i] = masks[d];
    A.win[i] = wins[d];
  }
  int *tent = S.get<int>(n), *cnt = S.get<int>(3);  // remaining, ncolors, err
  k_fill32<<<grid(n, 256), 256, 0, stream>>>(col_d, n, -1);
  k_fill32<<<1, 64, 0, stream>>>(cnt, 3, 0);
  int h[3] = {0, 0, 0};
  for (int round = 0; round < kMaxRounds; ++round) {
    for (size_t d = 0; d < wins.size(); ++d)
      if (wins[d])
        k_fill64<<<grid(sh.dat_to_size[d], 256), 256, 0, stream>>>(wins[d], sh.dat_to_size[d], ~0ull);
    k_fill32<<<1, 64, 0, stream>>>(cnt, 1, 0);
    k_elem_tentative<<<grid(n, 256), 256, 0, stream>>>(n, A, col_d, tent, cnt + 2);
    k_elem_claim<<<grid(n, 256), 256, 0, stream>>>(n, A, col_d, tent);
    k_elem_commit<<<grid(n, 256), 256, 0, stream>>>(n, A, col_d, tent, cnt, cnt + 1);
    GCHECK(hipMemcpyAsync(h, cnt, 3 * sizeof(int), hipMemcpyDeviceToHost, stream));
    GCHECK(hipStreamSynchronize(stream));
    if (h[2]) return -1;  // more than 64 colours needed
    if (h[0] == 0) return h[1];
  }
  return -1;
}

int gpu_colour_blocks_and_threads(const LoopShape &sh, int nblocks, int part_size,
                                  const op2h_staged_plan &P, int *blkcol_d,
                                  unsigned char *thrcol_d, int *nthrcol_d,
                                  hipStream_t stream) {
  (void)part_size;
  std::vector<int> cargs, cdat;
  if (!conflict_args(sh, cargs, cdat)) return -1;
  Scratch S;
  int *cnt = S.get<int>(3);
  k_fill32<<<1, 64, 0, stream>>>(cnt, 3, 0);
  // thread colours: LDS masks per conflicting dat, indexed by local target id
  LocalArgs L{};
  L.n = (int)cargs.size();
  std::vector<int> dat_lds(sh.dat_to_size.size(), -1);
  int words = 0;
  for (int i = 0; i < L.n; ++i) {
    int m = cargs[i], d = cdat[i];
    if (dat_lds[d] < 0) {
      dat_lds[d] = words;
      words += P.max_ind_h[sh.map_slot[m]];
    }
    L.loc[i] = P.loc_map[sh.loc_slot[m]];
    L.lds_off[i] = dat_lds[d];
  }
  if (words * 8 > 64 * 1024) return -1;
  if (nblocks > 0)
    k_thread_colours<<<nblocks, 64, words * 8, stream>>>(P.offset, P.nelems, L, words,
                                                         thrcol_d, nthrcol_d, cnt + 2);
  // block colours
  BlockTargets T{};
  T.n = 0;
  std::vector<int> seen_dat;
  std::vector<u64 *> masks, wins;
  for (int i = 0; i < (int)cargs.size(); ++i) {
    int m = cargs[i], d = cdat[i];
    if (std::find(seen_dat.begin(), seen_dat.end(), d) != seen_dat.end()) continue;
    seen_dat.push_back(d);
    int s = sh.map_slot[m];
    T.ind_map[T.n] = P.ind_map[s];
    T.ind_offs[T.n] = P.ind_offs[s];
    T.ind_sizes[T.n] = P.ind_sizes[s];
    T.mask[T.n] = S.get<u64>(sh.dat_to_size[d]);
    T.win[T.n] = S.get<u64>(sh.dat_to_size[d]);
    k_fill64<<<grid(sh.dat_to_size[d], 256), 256, 0, stream>>>(T.mask[T.n], sh.dat_to_size[d], 0ull);
    masks.push_back(T.mask[T.n]);
    wins.push_back(T.win[T.n]);
    T.n++;
  }
  int *tent = S.get<int>(nblocks);
  k_fill32<<<grid(nblocks, 256), 256, 0, stream>>>(blkcol_d, nblocks, -1);
  int h[3] = {0, 0, 0};
  for (int round = 0; round < kMaxRounds; ++round) {
    for (int i = 0; i < T.n; ++i)
      k_fill64<<<grid(sh.dat_to_size[seen_dat[i]], 256), 256, 0, stream>>>(
          wins[i], sh.dat_to_size[seen_dat[i]], ~0ull);
    k_fill32<<<1, 64, 0, stream>>>(cnt, 1, 0);
    if (nblocks > 0) {
      k_blk_tentative<<<nblocks, 256, 0, stream>>>(T, blkcol_d, tent, cnt + 2);
      k_blk_claim<<<nblocks, 256, 0, stream>>>(T, blkcol_d, tent);
      k_blk_commit<<<nblocks, 256, 0, stream>>>(T, blkcol_d, tent, cnt, cnt + 1);
    }
    GCHECK(hipMemcpyAsync(h, cnt, 3 * sizeof(int), hipMemcpyDeviceToHost, stream));
    GCHECK(hipStreamSynchronize(stream));
    if (h[2]) return -1;
    if (h[0] == 0) return h[1];
  }
  return -1;
}

}  // namespace op2h
