// Reverse Cuthill-McKee on the GPU, with the rule of op2h_rcm.h bit for bit.
//
// The host version is a queue: each dequeued vertex appends its unvisited
// neighbours sorted by (degree, index).  The vertices of BFS level L+1 are
// exactly those appended while level L is dequeued, each by the first (lowest
// Cuthill-McKee position) level-L vertex adjacent to it, so the queue order
// equals a level-synchronous order: level L+1 sorted by (position of that
// first parent, degree, index).  Here every level is one round of kernels:
// claim (atomicMin of the parent position), collect, three stable radix sorts
// (index, degree, parent position: least significant first), assign.  The
// start vertex of each component comes from the same George-Liu
// pseudo-peripheral search, on GPU breadth-first sweeps.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <climits>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "op2h_rcm.h"
#include "op2h_rcm_gpu.h"

#define RCHECK(x)                                                            \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      printf("op2hip rcm: %s failed at %s:%d: %s\n", #x, __FILE__, __LINE__, \
             hipGetErrorString(e_));                                         \
      exit(-1);                                                              \
    }                                                                        \
  } while (0)

namespace op2h {
namespace {

typedef unsigned long long u64;

int grid(long long n, int bs) { return n > 0 ? (int)((n + bs - 1) / bs) : 1; }

__global__ void k_fill(int *p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_degree(const int *off, int n, int *deg) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) deg[i] = off[i + 1] - off[i];
}

// lowest (degree, index) key over `list`
__global__ void k_min_key(const int *list, int m, const int *deg, u64 *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int v = list[i];
  atomicMin(out, ((u64)(unsigned)deg[v] << 32) | (unsigned)v);
}

// (degree, index) key of every vertex, for the order in which roots are tried
__global__ void k_root_keys(const int *deg, int n, int idx_bits, u64 *key, int *vert) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = ((u64)(unsigned)deg[i] << idx_bits) | (unsigned)i;
  vert[i] = i;
}

// first position in [lo, hi) of `order` whose vertex is still unplaced
__global__ void k_first_unplaced(const int *order, int lo, int hi, const int *pos, u64 *out) {
  int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < hi && pos[order[i]] < 0) atomicMin(out, (u64)i);
}

// vertices without neighbours lead the order, by index, one component each
__global__ void k_place_isolated(const int *order, int z, int *pos) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < z) pos[order[i]] = i;
}

// one breadth-first level over unplaced vertices: level[w] = depth + 1
__global__ void k_bfs(const int *frontier, int nf, const int *off, const int *adj,
                      const int *pos, int *level, int depth, int *next, int *nnext) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  int v = frontier[i];
  for (int k = off[v]; k < off[v + 1]; ++k) {
    int w = adj[k];
    if (pos[w] < 0 && atomicCAS(&level[w], -1, depth + 1) == -1) next[atomicAdd(nnext, 1)] = w;
  }
}

__global__ void k_reset_level(const int *list, int m, int *level) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) level[list[i]] = -1;
}

// first parent of every unplaced neighbour of the frontier
__global__ void k_claim(const int *frontier, int nf, const int *off, const int *adj,
                        const int *pos, int *parent) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  int v = frontier[i];
  for (int k = off[v]; k < off[v + 1]; ++k) {
    int w = adj[k];
    if (pos[w] < 0) atomicMin(&parent[w], pos[v]);
  }
}

// each newly reached vertex once (duplicate edges are skipped by `taken`)
__global__ void k_collect(const int *frontier, int nf, const int *off, const int *adj,
                          const int *pos, const int *parent, int *taken, int *next,
                          int *nnext) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  int v = frontier[i];
  for (int k = off[v]; k < off[v + 1]; ++k) {
    int w = adj[k];
    if (pos[w] < 0 && parent[w] == pos[v] && atomicCAS(&taken[w], 0, 1) == 0)
      next[atomicAdd(nnext, 1)] = w;
  }
}

// sort key of each newly reached vertex: (parent position relative to the
// level's first, degree, index) packed from the top; with par_shift < 0 only
// (degree, index) is packed and the parent goes to pkey for a second pass
__global__ void k_keys(const int *list, int m, const int *parent, const int *deg,
                       int par_base, int par_shift, int deg_shift, u64 *key,
                       unsigned *pkey) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int w = list[i];
  const u64 k = ((u64)(unsigned)deg[w] << deg_shift) | (unsigned)w;
  const unsigned par = (unsigned)(parent[w] - par_base);
  if (par_shift >= 0) {
    key[i] = k | ((u64)par << par_shift);
  } else {
    key[i] = k;
    pkey[i] = par;
  }
}

__global__ void k_parent_keys(const int *list, int m, const int *parent, int par_base,
                              unsigned *pkey) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) pkey[i] = (unsigned)(parent[list[i]] - par_base);
}

__global__ void k_assign(const int *list, int m, int base, int *pos) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) pos[list[i]] = base + i;
}

__global__ void k_place_root(int root, int at, int *pos) { pos[root] = at; }

struct Dev {
  std::vector<void *> ptrs;
  template <class T> T *alloc(size_t n) {
    void *p = nullptr;
    RCHECK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
    ptrs.push_back(p);
    return (T *)p;
  }
  ~Dev() {
    for (void *p : ptrs) (void)hipFree(p);
  }
};

}  // namespace

std::vector<int> rcm_order_gpu(int n, const std::vector<int> &off_h,
                               const std::vector<int> &adj_h, hipStream_t st,
                               bool two_pass) {
  std::vector<int> new_of_old(n);
  if (n == 0) return new_of_old;
  const int bs = 256, nadj = (int)adj_h.size();
  Dev d;
  int *off = d.alloc<int>(n + 1), *adj = d.alloc<int>(nadj), *deg = d.alloc<int>(n);
  int *pos = d.alloc<int>(n), *level = d.alloc<int>(n), *parent = d.alloc<int>(n);
  int *taken = d.alloc<int>(n), *bufa = d.alloc<int>(n), *bufb = d.alloc<int>(n);
  int *touched = d.alloc<int>(n), *cnt = d.alloc<int>(1);
  u64 *key = d.alloc<u64>(1), *keys_in = d.alloc<u64>(n), *keys_out = d.alloc<u64>(n);
  unsigned *pkeys_in = d.alloc<unsigned>(n), *pkeys_out = d.alloc<unsigned>(n);
  RCHECK(hipMemcpyAsync(off, off_h.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, st));
  if (nadj)
    RCHECK(hipMemcpyAsync(adj, adj_h.data(), nadj * sizeof(int), hipMemcpyHostToDevice, st));
  k_degree<<<grid(n, bs), bs, 0, st>>>(off, n, deg);
  k_fill<<<grid(n, bs), bs, 0, st>>>(pos, n, -1);
  k_fill<<<grid(n, bs), bs, 0, st>>>(level, n, -1);
  k_fill<<<grid(n, bs), bs, 0, st>>>(parent, n, INT_MAX);
  k_fill<<<grid(n, bs), bs, 0, st>>>(taken, n, 0);
  // key widths: index and degree bits are fixed, the parent's depend on the
  // width of the level it comes from
  auto bits = [](long long v) {
    int b = 0;
    while (v > 0) ++b, v >>= 1;
    return b;
  };
  int max_deg = 0;
  for (int v = 0; v < n; ++v) max_deg = std::max(max_deg, off_h[v + 1] - off_h[v]);
  const int idx_bits = bits(n - 1), deg_bits = bits(max_deg);
  // radix-sort scratch, sized once for the largest possible level
  size_t tmp_bytes = 0, tmp2 = 0;
  RCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys_in, keys_out, bufa, bufb,
                                            n, 0, 64, st));
  RCHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp2, pkeys_in, pkeys_out, bufa, bufb,
                                            n, 0, 32, st));
  tmp_bytes = std::max(tmp_bytes, tmp2);
  void *tmp = d.alloc<char>(tmp_bytes);

  auto read_key = [&]() {
    u64 h = 0;
    RCHECK(hipMemcpyAsync(&h, key, sizeof(u64), hipMemcpyDeviceToHost, st));
    RCHECK(hipStreamSynchronize(st));
    return h;
  };
  auto clear_key = [&]() {
    const u64 init = ~0ull;
    RCHECK(hipMemcpyAsync(key, &init, sizeof(u64), hipMemcpyHostToDevice, st));
  };
  auto min_key = [&](const int *list, int m) {
    clear_key();
    k_min_key<<<grid(m, bs), bs, 0, st>>>(list, m, deg, key);
    return (int)(read_key() & 0xffffffffu);
  };
  // roots are tried in (degree, index) order: sort once, then scan forward
  int *order = d.alloc<int>(n);
  k_root_keys<<<grid(n, bs), bs, 0, st>>>(deg, n, idx_bits, keys_in, bufa);
  RCHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys_in, keys_out, bufa, order, n,
                                            0, std::max(deg_bits + idx_bits, 1), st));
  int zero_deg = 0;
  for (int v = 0; v < n; ++v) zero_deg += off_h[v + 1] == off_h[v];
  k_place_isolated<<<grid(zero_deg, bs), bs, 0, st>>>(order, zero_deg, pos);
  int scan = zero_deg;  // every vertex before order[scan] is placed
  const int window = 1 << 16;
  auto next_root = [&]() {
    while (true) {
      const int hi = std::min(n, scan + window);
      clear_key();
      k_first_unplaced<<<grid(hi - scan, bs), bs, 0, st>>>(order, scan, hi, pos, key);
      const u64 h = read_key();
      if (h != ~0ull) {
        scan = (int)h;
        int v = 0;
        RCHECK(hipMemcpyAsync(&v, order + scan, sizeof(int), hipMemcpyDeviceToHost, st));
        RCHECK(hipStreamSynchronize(st));
        return v;
      }
      scan = hi;
    }
  };
  auto read_cnt = [&]() {
    int h = 0;
    RCHECK(hipMemcpyAsync(&h, cnt, sizeof(int), hipMemcpyDeviceToHost, st));
    RCHECK(hipStreamSynchronize(st));
    return h;
  };
  auto zero_cnt = [&]() { RCHECK(hipMemsetAsync(cnt, 0, sizeof(int), st)); };
  // eccentricity of `root` among unplaced vertices; the last level is left in
  // bufa[0, *nlast) and levels are reset afterwards
  auto bfs = [&](int root, int *nlast) {
    int h_root = root;
    RCHECK(hipMemcpyAsync(bufa, &h_root, sizeof(int), hipMemcpyHostToDevice, st));
    RCHECK(hipMemcpyAsync(touched, &h_root, sizeof(int), hipMemcpyHostToDevice, st));
    k_place_root<<<1, 1, 0, st>>>(root, 0, level);
    int nf = 1, ntouched = 1, depth = 0;
    int *cur = bufa, *nxt = bufb;
    while (true) {
      zero_cnt();
      k_bfs<<<grid(nf, bs), bs, 0, st>>>(cur, nf, off, adj, pos, level, depth, nxt, cnt);
      const int nn = read_cnt();
      if (nn == 0) break;
      RCHECK(hipMemcpyAsync(touched + ntouched, nxt, nn * sizeof(int), hipMemcpyDeviceToDevice, st));
      ntouched += nn;
      std::swap(cur, nxt);
      nf = nn;
      ++depth;
    }
    if (cur != bufa) RCHECK(hipMemcpyAsync(bufa, cur, nf * sizeof(int), hipMemcpyDeviceToDevice, st));
    *nlast = nf;
    k_reset_level<<<grid(ntouched, bs), bs, 0, st>>>(touched, ntouched, level);
    return depth;
  };

  int placed = zero_deg;
  while (placed < n) {
    int root = next_root();
    int nlast = 0;
    int ecc = bfs(root, &nlast);
    for (int it = 0; it < 16; ++it) {
      const int cand = min_key(bufa, nlast);
      int nlast2 = 0;
      const int ecc2 = bfs(cand, &nlast2);
      if (ecc2 <= ecc) break;
      root = cand;
      ecc = ecc2;
      nlast = nlast2;
    }
    // Cuthill-McKee from root, one level per round
    k_place_root<<<1, 1, 0, st>>>(root, placed, pos);
    int h_root = root;
    RCHECK(hipMemcpyAsync(bufa, &h_root, sizeof(int), hipMemcpyHostToDevice, st));
    int *cur = bufa, *nxt = bufb;
    int nf = 1;
    ++placed;
    while (nf > 0) {
      k_claim<<<grid(nf, bs), bs, 0, st>>>(cur, nf, off, adj, pos, parent);
      zero_cnt();
      k_collect<<<grid(nf, bs), bs, 0, st>>>(cur, nf, off, adj, pos, parent, taken, nxt, cnt);
      const int nn = read_cnt();
      if (nn == 0) break;
      // level order: (first parent's position, degree, index), one radix
      // sort when the packed key fits 64 bits, else a stable second pass on
      // the parent; the spent frontier's buffer takes each pass's output
      const int par_bits = bits(nf - 1), low_bits = deg_bits + idx_bits;
      const bool one_pass = !two_pass && par_bits + low_bits <= 64;
      k_keys<<<grid(nn, bs), bs, 0, st>>>(nxt, nn, parent, deg, placed - nf,
                                          one_pass ? low_bits : -1, idx_bits, keys_in,
                                          pkeys_in);
      RCHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys_in, keys_out, nxt, cur, nn,
                                                0, one_pass ? par_bits + low_bits : low_bits,
                                                st));
      std::swap(cur, nxt);
      if (!one_pass) {
        // the parent keys follow the first pass's permutation
        k_parent_keys<<<grid(nn, bs), bs, 0, st>>>(nxt, nn, parent, placed - nf, pkeys_in);
        RCHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, pkeys_in, pkeys_out, nxt, cur,
                                                  nn, 0, std::max(par_bits, 1), st));
        std::swap(cur, nxt);
      }
      k_assign<<<grid(nn, bs), bs, 0, st>>>(nxt, nn, placed, pos);
      placed += nn;
      std::swap(cur, nxt);
      nf = nn;
    }
  }
  std::vector<int> pos_h(n);
  RCHECK(hipMemcpyAsync(pos_h.data(), pos, n * sizeof(int), hipMemcpyDeviceToHost, st));
  RCHECK(hipStreamSynchronize(st));
  for (int v = 0; v < n; ++v) new_of_old[v] = n - 1 - pos_h[v];
  return new_of_old;
}

}  // namespace op2h

extern "C" int op2h_rcm_order(int n, const int *off, const int *adj, int on_gpu,
                              int *new_of_old) {
  if (n < 0 || on_gpu < 0 || on_gpu > 2 || !off || (n && (!new_of_old || off[0] != 0)))
    return -1;
  const int nadj = n ? off[n] : 0;
  if (nadj < 0 || (nadj && !adj)) return -1;
  for (int v = 0; v < n; ++v) {
    if (off[v + 1] < off[v]) return -1;
    for (int k = off[v]; k < off[v + 1]; ++k)
      if (adj[k] < 0 || adj[k] >= n || (k > off[v] && adj[k] <= adj[k - 1])) return -1;
  }
  std::vector<int> o(off, off + n + 1), a(adj, adj + nadj), r;
  if (on_gpu) {
    hipStream_t st;
    RCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    r = op2h::rcm_order_gpu(n, o, a, st, on_gpu == 2);
    RCHECK(hipStreamDestroy(st));
  } else {
    r = op2h::rcm_order(n, o, a);
  }
  std::copy(r.begin(), r.end(), new_of_old);
  return 0;
}
