// Owner-compute partition plan (OP2H_PLAN_OWNER).
//
// The reference's plans execute increments either through a two-level
// colouring of contiguous element blocks (op_plan_core with OP_STAGE_ALL,
// op2/src/core/op_rt_support.cpp:426-450 block partition, :609-682 staging
// lists, :684-884 colouring) or with atomics.  This plan partitions the
// *target* set instead: each block owns a compact patch of targets, evaluates
// every element touching the patch once (boundary elements are evaluated by
// both neighbouring blocks, as the reference's MPI backend does for its exec
// halo, docs/developer_guide.rst:254-333), and each owned target sums its own
// increments in ascending element order.  No colours, no atomics, and the sum
// order is the sequential executor's (op_seq.h:918-1073), so results are
// bit-identical to it.
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "op2h_plan_build.h"

#define OCHECK(x)                                                            \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      printf("op2hip: %s failed at %s:%d: %s\n", #x, __FILE__, __LINE__,     \
             hipGetErrorString(e_));                                         \
      exit(-1);                                                              \
    }                                                                        \
  } while (0)

namespace op2h {

// Breadth-first growth of blocks over the target graph (targets joined when
// one element increments both).  Seeds are taken in ascending target order;
// when a block's search runs dry before block_cells targets (a fragment cut
// off by earlier blocks, or no adjacency at all) it continues from the next
// unassigned seed, so every block but the last holds exactly block_cells
// targets.  When the numbering is local (the targets of one element lie
// within a bandwidth w much below 2^15) a block only takes targets within
// 65535 - 2w - 1024 of its first seed, so that every target its elements
// reach lies within 16 bits of the block's lowest (stubs then read 16-bit
// offsets instead of full indices): the window cuts a block short only when
// the growth would have left it.  Returns the block count.
static int grow_blocks(int nto, int nel, int dim, const int *map, int ninc,
                       const int *inc_idx, int block_cells,
                       std::vector<int> &block_of) {
  std::vector<int> deg(nto + 1, 0);
  std::vector<char> used(nto, 0);
  for (int e = 0; e < nel; ++e)
    for (int k = 0; k < ninc; ++k) {
      int t = map[(size_t)e * dim + inc_idx[k]];
      used[t] = 1;
      for (int l = 0; l < ninc; ++l) {
        int u = map[(size_t)e * dim + inc_idx[l]];
        if (l != k && u != t) deg[t + 1]++;
      }
    }
  for (int t = 0; t < nto; ++t) deg[t + 1] += deg[t];
  std::vector<int> adj(deg[nto]), fill(deg.begin(), deg.end() - 1);
  for (int e = 0; e < nel; ++e)
    for (int k = 0; k < ninc; ++k) {
      int t = map[(size_t)e * dim + inc_idx[k]];
      for (int l = 0; l < ninc; ++l) {
        int u = map[(size_t)e * dim + inc_idx[l]];
        if (l != k && u != t) adj[fill[t]++] = u;
      }
    }
  // the bandwidth of the numbering: the widest spread of one element's targets
  long long w = 0;
  for (int e = 0; e < nel; ++e) {
    int lo = INT_MAX, hi = -1;
    for (int k = 0; k < ninc; ++k) {
      const int t = map[(size_t)e * dim + inc_idx[k]];
      lo = std::min(lo, t);
      hi = std::max(hi, t);
    }
    w = std::max(w, (long long)hi - lo);
  }
  const long long window = 65535 - 2 * w - 1024;
  const long long span = window >= 32768 ? window : 0;  // 0: no window
  block_of.assign(nto, -1);
  std::vector<int> mark(nto, -1), queue;
  queue.reserve(4 * (size_t)block_cells + 16);
  int nb = 0, seed = 0;
  auto next_seed = [&]() {
    // (called only with an empty or fully drained queue, so no unassigned
    // target is waiting in it)
    while (seed < nto && (block_of[seed] >= 0 || !used[seed])) ++seed;
    return seed < nto;
  };
  while (next_seed()) {
    queue.clear();
    size_t head = 0;
    int count = 0;
    const long long s0 = seed;  // the block's lowest target
    while (count < block_cells) {
      if (head == queue.size()) {
        if (!next_seed()) break;
        if (span && seed - s0 > span) break;
        mark[seed] = nb;
        queue.push_back(seed);
      }
      int c = queue[head++];
      block_of[c] = nb;
      ++count;
      for (int j = deg[c]; j < deg[c + 1]; ++j) {
        int u = adj[j];
        if (block_of[u] < 0 && mark[u] != nb && (!span || u - s0 <= span)) {
          mark[u] = nb;
          queue.push_back(u);
        }
      }
    }
    ++nb;
  }
  return nb;
}

bool owner_lists(int nto, int nel, int dim, const int *map, int ninc,
                 const int *inc_idx, int block_cells, OwnerLists &L,
                 std::string &err, const int *elem_key, bool key_instances) {
  std::vector<int> block_of;
  const int nb = grow_blocks(nto, nel, dim, map, ninc, inc_idx, block_cells, block_of);
  L.nblocks = nb;

  // owned targets, ascending within each block
  std::vector<int> &owned_offs = L.owned_offs;
  owned_offs.assign(nb + 1, 0);
  for (int t = 0; t < nto; ++t)
    if (block_of[t] >= 0) owned_offs[block_of[t] + 1]++;
  for (int b = 0; b < nb; ++b) owned_offs[b + 1] += owned_offs[b];
  std::vector<int> &owned = L.owned;
  owned.assign(owned_offs[nb], 0);
  {
    std::vector<int> f(owned_offs.begin(), owned_offs.end() - 1);
    for (int t = 0; t < nto; ++t)
      if (block_of[t] >= 0) owned[f[block_of[t]]++] = t;
  }
  // instances: each element once per distinct block among its targets
  auto blocks_of_elem = [&](int e, int *out) {
    int n = 0;
    for (int k = 0; k < ninc; ++k) {
      int b = block_of[map[(size_t)e * dim + inc_idx[k]]];
      bool seen = false;
      for (int j = 0; j < n; ++j) seen |= out[j] == b;
      if (!seen) out[n++] = b;
    }
    return n;
  };
  std::vector<int> &elem_offs = L.elem_offs, tmp(ninc);
  elem_offs.assign(nb + 1, 0);
  for (int e = 0; e < nel; ++e) {
    int n = blocks_of_elem(e, tmp.data());
    for (int j = 0; j < n; ++j) elem_offs[tmp[j] + 1]++;
  }
  for (int b = 0; b < nb; ++b) elem_offs[b + 1] += elem_offs[b];
  const int total = elem_offs[nb];
  std::vector<int> &elems = L.elems;
  elems.assign(total, 0);
  {
    // instances in element order: ascending element, or ascending key (the
    // application's element order) when a key is given, so that a block's
    // instances and each target's increments follow the same order
    std::vector<int> order(nel);
    for (int e = 0; e < nel; ++e) order[e] = e;
    if (elem_key && key_instances)
      std::sort(order.begin(), order.end(),
                [&](int a, int b) { return elem_key[a] < elem_key[b]; });
    std::vector<int> f(elem_offs.begin(), elem_offs.end() - 1);
    for (int e : order) {
      int n = blocks_of_elem(e, tmp.data());
      for (int j = 0; j < n; ++j) elems[f[tmp[j]]++] = e;
    }
  }
  L.max_elems = L.max_owned = 0;
  for (int b = 0; b < nb; ++b) {
    L.max_elems = std::max(L.max_elems, elem_offs[b + 1] - elem_offs[b]);
    L.max_owned = std::max(L.max_owned, owned_offs[b + 1] - owned_offs[b]);
  }
  if ((long long)L.max_elems * ninc > 65535) {
    err = "owner plan: a block has too many elements for 16-bit local ids";
    return false;
  }

  // increments of each owned target, in (instance, k) order
  std::vector<int> ownpos(nto, -1);
  for (int i = 0; i < (int)owned.size(); ++i) ownpos[owned[i]] = i;
  std::vector<int> &csr_offs = L.csr_offs;
  csr_offs.assign(owned.size() + 1, 0);
  for (int b = 0; b < nb; ++b)
    for (int j = elem_offs[b]; j < elem_offs[b + 1]; ++j)
      for (int k = 0; k < ninc; ++k) {
        int t = map[(size_t)elems[j] * dim + inc_idx[k]];
        if (block_of[t] == b) csr_offs[ownpos[t] + 1]++;
      }
  for (size_t i = 0; i < owned.size(); ++i) csr_offs[i + 1] += csr_offs[i];
  std::vector<unsigned short> &csr = L.csr;
  csr.assign(csr_offs.back(), 0);
  {
    std::vector<int> f(csr_offs.begin(), csr_offs.end() - 1);
    for (int b = 0; b < nb; ++b)
      for (int j = elem_offs[b]; j < elem_offs[b + 1]; ++j)
        for (int k = 0; k < ninc; ++k) {
          int t = map[(size_t)elems[j] * dim + inc_idx[k]];
          if (block_of[t] == b)
            csr[f[ownpos[t]]++] = (unsigned short)((j - elem_offs[b]) * ninc + k);
        }
  }
  if (elem_key) {  // each target's increments in (key of the element, k) order
    for (int b = 0; b < nb; ++b)
      for (int i = owned_offs[b]; i < owned_offs[b + 1]; ++i)
        std::sort(csr.begin() + csr_offs[i], csr.begin() + csr_offs[i + 1],
                  [&](unsigned short x, unsigned short y) {
                    const int kx = elem_key[elems[elem_offs[b] + x / ninc]];
                    const int ky = elem_key[elems[elem_offs[b] + y / ninc]];
                    return kx != ky ? kx < ky : x % ninc < y % ninc;
                  });
  }
  L.max_incs = 0;
  for (size_t i = 0; i < owned.size(); ++i)
    L.max_incs = std::max(L.max_incs, csr_offs[i + 1] - csr_offs[i]);
  return true;
}

std::unique_ptr<OwnerPlanRec> build_owner_plan(op_set set, int block_cells,
                                               int nargs, const op_arg *args,
                                               hipStream_t stream,
                                               std::string &err, bool key_instances) {
  auto rec = std::make_unique<OwnerPlanRec>();
  op2h_owner_plan &P = rec->plan;
  // the increment args: all through one map into one dat
  op_map own = nullptr;
  op_dat owndat = nullptr;
  std::vector<int> inc_idx;
  for (int m = 0; m < nargs; ++m) P.inc_of_arg[m] = -1;
  for (int m = 0; m < nargs; ++m) {
    const op_arg &a = args[m];
    if (a.argtype != OP_ARG_DAT || !a.opt || !a.map || a.idx < 0) continue;
    if (a.acc == OP_RW || a.acc == OP_WRITE) {
      err = "owner plan: indirect OP_RW/OP_WRITE args are not supported";
      return nullptr;
    }
    if (a.acc != OP_INC) continue;
    if ((own && a.map != own) || (owndat && a.dat != owndat)) {
      err = "owner plan: increments must all go through one map into one dat";
      return nullptr;
    }
    own = a.map;
    owndat = a.dat;
    P.inc_of_arg[m] = (int)inc_idx.size();
    inc_idx.push_back(a.idx);
  }
  if (!own) {
    err = "owner plan: the loop has no indirect increments";
    return nullptr;
  }
  // Elements on a block boundary are evaluated by both blocks: every other
  // effect of an element must be one that a second evaluation repeats
  // harmlessly, and no element may read the dat being incremented.
  for (int m = 0; m < nargs; ++m) {
    const op_arg &a = args[m];
    if (!a.opt) continue;
    if (a.argtype == OP_ARG_GBL && a.acc != OP_READ) {
      err = "owner plan: global reductions are not supported (elements on block "
            "boundaries are evaluated twice)";
      return nullptr;
    }
    if (a.argtype == OP_ARG_INFO) {
      err = "owner plan: op_arg_info is not supported";
      return nullptr;
    }
    if (a.argtype != OP_ARG_DAT) continue;
    if (a.dat == owndat && a.acc != OP_INC) {
      err = "owner plan: the incremented dat is also read by the loop";
      return nullptr;
    }
    if (!a.map && (a.acc == OP_RW || a.acc == OP_INC)) {
      err = "owner plan: direct OP_RW/OP_INC args are not supported (elements on "
            "block boundaries are evaluated twice)";
      return nullptr;
    }
  }
  if (block_cells < 1 || block_cells > 4096) {
    err = "owner plan: block size must be in [1, 4096]";
    return nullptr;
  }
  const int nel = set->size, nto = own->to->size, dim = own->dim;
  const int ninc = (int)inc_idx.size();
  OwnerLists L;
  // under transparent renumbering the application's element order fixes the
  // order of each target's increments
  if (!owner_lists(nto, nel, dim, host_map(own), ninc, inc_idx.data(), block_cells, L, err,
                   old_index_host(set), key_instances))
    return nullptr;
  rec->key_instances = key_instances;
  P.set_size = nel;
  P.to_size = nto;
  P.block_cells = block_cells;
  P.nblocks = L.nblocks;
  P.ninc = ninc;
  P.total = (int)L.elems.size();
  P.redundancy = nel ? (double)P.total / nel : 1.0;
  P.max_elems = L.max_elems;
  P.max_owned = L.max_owned;
  const std::vector<int> &owned_offs = L.owned_offs, &owned = L.owned,
                         &elem_offs = L.elem_offs, &elems = L.elems, &csr_offs = L.csr_offs;
  const std::vector<unsigned short> &csr = L.csr;
  P.owned_offs = rec->mem.upload(owned_offs, stream);
  P.owned = rec->mem.upload(owned, stream);
  P.elem_offs = rec->mem.upload(elem_offs, stream);
  P.elems = rec->mem.upload(elems, stream);
  P.csr_offs = rec->mem.upload(csr_offs, stream);
  P.csr = rec->mem.upload(csr, stream);
  rec->n_owned = (int)owned.size();
  rec->n_csr = (int)csr.size();
  P.n_owned = rec->n_owned;
  P.max_incs = L.max_incs;
  P.csr4 = nullptr;
  if (L.max_incs <= 4) {
    std::vector<unsigned short> c4(4 * owned.size(), 0xFFFF);
    for (size_t i = 0; i < owned.size(); ++i)
      for (int k = csr_offs[i]; k < csr_offs[i + 1]; ++k) c4[4 * i + (k - csr_offs[i])] = csr[k];
    P.csr4 = rec->mem.upload(c4, stream);
    OCHECK(hipStreamSynchronize(stream));  // c4 is a local
  }
  OCHECK(hipStreamSynchronize(stream));
  return rec;
}

int download_owner(const OwnerPlanRec &r, int *owned_offs, int *owned,
                   int *elem_offs, int *elems, int *csr_offs,
                   unsigned short *csr) {
  const op2h_owner_plan &p = r.plan;
  OCHECK(hipDeviceSynchronize());
  auto get = [](void *dst, const void *src, size_t bytes) {
    if (dst && bytes) OCHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  };
  get(owned_offs, p.owned_offs, sizeof(int) * (p.nblocks + 1));
  get(owned, p.owned, sizeof(int) * r.n_owned);
  get(elem_offs, p.elem_offs, sizeof(int) * (p.nblocks + 1));
  get(elems, p.elems, sizeof(int) * p.total);
  get(csr_offs, p.csr_offs, sizeof(int) * (r.n_owned + 1));
  get(csr, p.csr, sizeof(unsigned short) * r.n_csr);
  return 0;
}

}  // namespace op2h

extern "C" int op2h_owner_partition(int nto, int nel, int dim, const int *map,
                                    int ninc, const int *inc_idx,
                                    int block_cells, int *block_of) {
  if (nto < 0 || nel < 0 || dim < 1 || ninc < 1 || !map || !inc_idx ||
      block_cells < 1 || !block_of)
    return -1;
  for (int k = 0; k < ninc; ++k)
    if (inc_idx[k] < 0 || inc_idx[k] >= dim) return -1;
  for (size_t i = 0; i < (size_t)nel * dim; ++i)
    if (map[i] < 0 || map[i] >= nto) return -1;
  std::vector<int> b;
  int nb = op2h::grow_blocks(nto, nel, dim, map, ninc, inc_idx, block_cells, b);
  std::copy(b.begin(), b.end(), block_of);
  return nb;
}
