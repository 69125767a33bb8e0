// Plan construction: block partition, staging lists, two-level colouring
// (op_plan_core(OP_STAGE_ALL), op2/src/core/op_rt_support.cpp:301-1126), flat
// colouring (OP_COLOR2, :380-382,445-450,757-796) and the owner-compute inverse
// map.  Colourings are built with the reference's greedy first-fit rule on the
// host, or on the GPU by op2h_plan_gpu.hip.
#include "op2h_plan_build.h"

#include <strings.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>

#define PCHECK(x)                                                            \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      printf("op2hip: %s failed at %s:%d: %s\n", #x, __FILE__, __LINE__,     \
             hipGetErrorString(e_));                                         \
      exit(-1);                                                              \
    }                                                                        \
  } while (0)

namespace op2h {

template <class T>
T *DeviceAllocs::upload(const std::vector<T> &v, hipStream_t s) {
  T *p = alloc<T>(v.size());
  if (!v.empty())
    PCHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T),
                          hipMemcpyHostToDevice, s));
  return p;
}

template <class T> T *DeviceAllocs::alloc(size_t n) {
  void *p = nullptr;
  PCHECK(hipMalloc(&p, std::max(n, (size_t)1) * sizeof(T)));
  ptrs.push_back(p);
  return (T *)p;
}

DeviceAllocs::~DeviceAllocs() {
  for (void *p : ptrs) (void)hipFree(p);
}

LoopShape loop_shape(op_set set, int nargs, const op_arg *args) {
  LoopShape sh;
  sh.set_size = set->size;
  sh.nargs = nargs;
  sh.map_slot.assign(nargs, -1);
  sh.loc_slot.assign(nargs, -1);
  sh.dat_slot.assign(nargs, -1);
  sh.conflict.assign(nargs, false);
  sh.arg_map.assign(nargs, {nullptr, -1});
  std::vector<op_dat> dats;
  for (int m = 0; m < nargs; ++m) {
    const op_arg &a = args[m];
    if (a.argtype != OP_ARG_DAT || !a.opt || !a.map || a.idx < 0) continue;
    sh.arg_map[m] = {a.map, a.idx};
    auto mi = std::find(sh.maps.begin(), sh.maps.end(), a.map);
    if (mi == sh.maps.end()) {
      sh.maps.push_back(a.map);
      mi = sh.maps.end() - 1;
    }
    sh.map_slot[m] = (int)(mi - sh.maps.begin());
    std::pair<op_map, int> key{a.map, a.idx};
    auto li = std::find(sh.locs.begin(), sh.locs.end(), key);
    if (li == sh.locs.end()) {
      sh.locs.push_back(key);
      li = sh.locs.end() - 1;
    }
    sh.loc_slot[m] = (int)(li - sh.locs.begin());
    auto di = std::find(dats.begin(), dats.end(), a.dat);
    if (di == dats.end()) {
      dats.push_back(a.dat);
      sh.dat_to_size.push_back(a.dat->set->size);
      di = dats.end() - 1;
    }
    sh.dat_slot[m] = (int)(di - dats.begin());
    sh.conflict[m] = a.acc == OP_INC || a.acc == OP_RW;
  }
  if ((int)sh.maps.size() > OP2H_MAX_ARGS || (int)sh.locs.size() > OP2H_MAX_ARGS) {
    printf("op2hip: too many indirect maps in one loop\n");
    exit(-1);
  }
  return sh;
}

namespace {
inline int mapv(const LoopShape &sh, int m, int e) {
  const auto &am = sh.arg_map[m];
  return host_map(am.first)[(size_t)e * am.first->dim + am.second];
}
}  // namespace

int greedy_block_thread_colours(const LoopShape &sh, int lo, int hi,
                                std::vector<std::vector<unsigned>> &work,
                                int *thrcol) {
  for (int e = lo; e < hi; ++e) thrcol[e - lo] = -1;
  int repeat = 1, ncolor = 0, ncolors = 0;
  while (repeat) {
    repeat = 0;
    for (int m = 0; m < sh.nargs; ++m)
      if (sh.dat_slot[m] >= 0)
        for (int e = lo; e < hi; ++e) work[sh.dat_slot[m]][mapv(sh, m, e)] = 0;
    for (int e = lo; e < hi; ++e) {
      if (thrcol[e - lo] != -1) continue;
      unsigned mask = 0;
      for (int m = 0; m < sh.nargs; ++m)
        if (sh.conflict[m]) mask |= work[sh.dat_slot[m]][mapv(sh, m, e)];
      int color = ffs((int)~mask) - 1;
      if (color == -1) {
        repeat = 1;
        continue;
      }
      thrcol[e - lo] = ncolor + color;
      ncolors = std::max(ncolors, ncolor + color + 1);
      for (int m = 0; m < sh.nargs; ++m)
        if (sh.conflict[m]) work[sh.dat_slot[m]][mapv(sh, m, e)] |= 1u << color;
    }
    ncolor += 32;
  }
  return ncolors;
}

int greedy_block_colours(const LoopShape &sh, int nblocks, const int *offset,
                         const int *nelems,
                         std::vector<std::vector<unsigned>> &work,
                         int *blkcol) {
  for (int b = 0; b < nblocks; ++b) blkcol[b] = -1;
  int repeat = 1, ncolor = 0, ncolors = 0;
  while (repeat) {
    repeat = 0;
    for (auto &w : work) std::fill(w.begin(), w.end(), 0u);
    for (int b = 0; b < nblocks; ++b) {
      if (blkcol[b] != -1) continue;
      int lo = offset[b], hi = lo + nelems[b];
      unsigned mask = 0;
      for (int m = 0; m < sh.nargs; ++m)
        if (sh.conflict[m])
          for (int e = lo; e < hi; ++e) mask |= work[sh.dat_slot[m]][mapv(sh, m, e)];
      int color = ffs((int)~mask) - 1;
      if (color == -1) {
        repeat = 1;
        continue;
      }
      blkcol[b] = ncolor + color;
      ncolors = std::max(ncolors, ncolor + color + 1);
      for (int m = 0; m < sh.nargs; ++m)
        if (sh.conflict[m])
          for (int e = lo; e < hi; ++e) work[sh.dat_slot[m]][mapv(sh, m, e)] |= 1u << color;
    }
    ncolor += 32;
  }
  return ncolors;
}

std::unique_ptr<StagedPlanRec> build_staged_plan(op_set set, int part_size,
                                                 int nargs, const op_arg *args,
                                                 int on_device,
                                                 hipStream_t stream) {
  auto rec = std::make_unique<StagedPlanRec>();
  op2h_staged_plan &P = rec->plan;
  LoopShape sh = loop_shape(set, nargs, args);
  const int n = set->size;
  const int bs = part_size > 0 ? part_size : 256;
  const int nblocks = n > 0 ? (n + bs - 1) / bs : 0;
  P.set_size = n;
  P.part_size = bs;
  P.nblocks = nblocks;
  P.nmaps = (int)sh.maps.size();
  for (int m = 0; m < OP2H_MAX_ARGS; ++m) {
    P.map_of_arg[m] = m < nargs ? sh.map_slot[m] : -1;
    P.loc_of_arg[m] = m < nargs ? sh.loc_slot[m] : -1;
  }
  std::vector<int> offset(nblocks), nelems(nblocks);
  for (int b = 0; b < nblocks; ++b) {
    offset[b] = b * bs;
    nelems[b] = std::min(bs, n - b * bs);
  }
  // staging lists: unique sorted targets per (block, map), local ids per loc slot
  const int nm = (int)sh.maps.size(), nl = (int)sh.locs.size();
  std::vector<std::vector<int>> ind_offs(nm, std::vector<int>(nblocks)),
      ind_sizes(nm, std::vector<int>(nblocks)), ind_map(nm);
  std::vector<std::vector<unsigned short>> loc_map(nl, std::vector<unsigned short>(n));
  std::vector<std::vector<int>> inv(nm);
  for (int s = 0; s < nm; ++s) inv[s].assign(std::max(sh.maps[s]->to->size, 1), 0);
  std::vector<int> tmp;
  for (int s = 0; s < nm; ++s) P.max_ind_h[s] = 0;
  for (int b = 0; b < nblocks; ++b) {
    int lo = offset[b], hi = lo + nelems[b];
    for (int s = 0; s < nm; ++s) {
      op_map mp = sh.maps[s];
      tmp.clear();
      for (int l = 0; l < nl; ++l)
        if (sh.locs[l].first == mp)
          for (int e = lo; e < hi; ++e)
            tmp.push_back(host_map(mp)[(size_t)e * mp->dim + sh.locs[l].second]);
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      if (tmp.size() > 65535) {
        printf("op2hip: plan block has more than 65535 targets\n");
        exit(-1);
      }
      ind_offs[s][b] = (int)ind_map[s].size();
      ind_sizes[s][b] = (int)tmp.size();
      P.max_ind_h[s] = std::max(P.max_ind_h[s], (int)tmp.size());
      for (size_t k = 0; k < tmp.size(); ++k) {
        ind_map[s].push_back(tmp[k]);
        inv[s][tmp[k]] = (int)k;
      }
      for (int l = 0; l < nl; ++l)
        if (sh.locs[l].first == mp)
          for (int e = lo; e < hi; ++e)
            loc_map[l][e] = (unsigned short)inv[s][host_map(mp)[(size_t)e * mp->dim +
                                                           sh.locs[l].second]];
    }
  }
  // staging lists and partition go to the device first: the device
  // colouring reads them
  P.offset = rec->mem.upload(offset, stream);
  P.nelems = rec->mem.upload(nelems, stream);
  for (int s = 0; s < nm; ++s) {
    P.ind_offs[s] = rec->mem.upload(ind_offs[s], stream);
    P.ind_sizes[s] = rec->mem.upload(ind_sizes[s], stream);
    P.ind_map[s] = rec->mem.upload(ind_map[s], stream);
  }
  for (int l = 0; l < nl; ++l) P.loc_map[l] = rec->mem.upload(loc_map[l], stream);
  // colouring
  std::vector<int> nthrcol(nblocks), blkcol(nblocks);
  int ncolors = -1;
  unsigned char *thrcol_d = rec->mem.alloc<unsigned char>(std::max(n, 1));
  int *nthrcol_d = rec->mem.alloc<int>(std::max(nblocks, 1));
  if (on_device) {
    int *blkcol_d = rec->mem.alloc<int>(std::max(nblocks, 1));
    ncolors = gpu_colour_blocks_and_threads(sh, nblocks, bs, P, blkcol_d, thrcol_d,
                                            nthrcol_d, stream);
    if (ncolors >= 0) {
      PCHECK(hipMemcpyAsync(blkcol.data(), blkcol_d, nblocks * sizeof(int),
                            hipMemcpyDeviceToHost, stream));
      PCHECK(hipMemcpyAsync(nthrcol.data(), nthrcol_d, nblocks * sizeof(int),
                            hipMemcpyDeviceToHost, stream));
      PCHECK(hipStreamSynchronize(stream));
      for (int b = 0; b < nblocks; ++b)
        if (nthrcol[b] > 255) ncolors = -1;
      P.built_on_device = ncolors >= 0;
    }
  }
  if (ncolors < 0) {
    std::vector<int> thrcol_i(std::max(n, 1));
    std::vector<unsigned char> thrcol(std::max(n, 1));
    std::vector<std::vector<unsigned>> work(sh.dat_to_size.size());
    for (size_t d = 0; d < work.size(); ++d)
      work[d].assign(std::max(sh.dat_to_size[d], 1), 0u);
    for (int b = 0; b < nblocks; ++b) {
      nthrcol[b] = greedy_block_thread_colours(sh, offset[b], offset[b] + nelems[b],
                                               work, &thrcol_i[offset[b]]);
      if (nthrcol[b] > 255) {
        printf("op2hip: more than 255 thread colours in a block\n");
        exit(-1);
      }
    }
    for (int e = 0; e < n; ++e) thrcol[e] = (unsigned char)thrcol_i[e];
    ncolors = greedy_block_colours(sh, nblocks, offset.data(), nelems.data(), work,
                                   blkcol.data());
    PCHECK(hipMemcpyAsync(thrcol_d, thrcol.data(), n, hipMemcpyHostToDevice, stream));
    PCHECK(hipMemcpyAsync(nthrcol_d, nthrcol.data(), nblocks * sizeof(int),
                          hipMemcpyHostToDevice, stream));
    PCHECK(hipStreamSynchronize(stream));
    P.built_on_device = 0;
  }
  P.ncolors = ncolors;
  rec->ncolblk_h.assign(std::max(ncolors, 1), 0);
  for (int b = 0; b < nblocks; ++b) rec->ncolblk_h[blkcol[b]]++;
  rec->colstart_h.assign(ncolors + 1, 0);
  for (int c = 0; c < ncolors; ++c)
    rec->colstart_h[c + 1] = rec->colstart_h[c] + rec->ncolblk_h[c];
  std::vector<int> blkmap(nblocks), fill(rec->colstart_h.begin(), rec->colstart_h.end() - 1);
  for (int b = 0; b < nblocks; ++b) blkmap[fill[blkcol[b]]++] = b;
  int maxthr = 0;
  for (int b = 0; b < nblocks; ++b) maxthr = std::max(maxthr, nthrcol[b]);
  rec->max_nthrcol_h.assign(1, maxthr);
  P.ncolblk_h = rec->ncolblk_h.data();
  P.colstart_h = rec->colstart_h.data();
  P.max_nthrcol_h = rec->max_nthrcol_h.data();
  P.blkmap = rec->mem.upload(blkmap, stream);
  P.nthrcol = nthrcol_d;
  P.thrcol = thrcol_d;
  PCHECK(hipStreamSynchronize(stream));
  return rec;
}

std::unique_ptr<ColorPlanRec> build_color_plan(op_set set, int nargs,
                                               const op_arg *args,
                                               int on_device,
                                               hipStream_t stream) {
  auto rec = std::make_unique<ColorPlanRec>();
  op2h_color_plan &P = rec->plan;
  LoopShape sh = loop_shape(set, nargs, args);
  const int n = set->size;
  P.set_size = n;
  std::vector<int> col(std::max(n, 1));
  int ncolors = 0;
  ncolors = -1;
  if (on_device) {
    int *col_d = rec->mem.alloc<int>(std::max(n, 1));
    ncolors = gpu_colour_elements(sh, n, col_d, stream);
    if (ncolors >= 0) {
      PCHECK(hipMemcpyAsync(col.data(), col_d, n * sizeof(int),
                            hipMemcpyDeviceToHost, stream));
      PCHECK(hipStreamSynchronize(stream));
      P.built_on_device = 1;
    }
  }
  if (ncolors < 0) {
    // OP_COLOR2: one block spanning the set, greedy thread colouring
    std::vector<std::vector<unsigned>> work(sh.dat_to_size.size());
    for (size_t d = 0; d < work.size(); ++d)
      work[d].assign(std::max(sh.dat_to_size[d], 1), 0u);
    ncolors = greedy_block_thread_colours(sh, 0, n, work, col.data());
  }
  P.ncolors = ncolors;
  rec->col_offsets_h.assign(ncolors + 1, 0);
  for (int e = 0; e < n; ++e) rec->col_offsets_h[col[e] + 1]++;
  for (int c = 0; c < ncolors; ++c) rec->col_offsets_h[c + 1] += rec->col_offsets_h[c];
  std::vector<int> reord(std::max(n, 1)),
      fill(rec->col_offsets_h.begin(), rec->col_offsets_h.end() - 1);
  for (int e = 0; e < n; ++e) reord[fill[col[e]]++] = e;
  P.col_offsets_h = rec->col_offsets_h.data();
  P.col_reord = rec->mem.upload(reord, stream);
  PCHECK(hipStreamSynchronize(stream));
  return rec;
}

std::unique_ptr<GatherPlanRec> build_gather_plan(op_map map, hipStream_t stream) {
  auto rec = std::make_unique<GatherPlanRec>();
  op2h_gather_plan &P = rec->plan;
  const int from = map->from->size, to = map->to->size, dim = map->dim;
  P.from_size = from;
  P.to_size = to;
  P.dim = dim;
  std::vector<int> count(to + 1, 0);
  for (size_t i = 0; i < (size_t)from * dim; ++i) count[host_map(map)[i] + 1]++;
  std::vector<int> targets;
  for (int t = 0; t < to; ++t)
    if (count[t + 1]) targets.push_back(t);
  for (int t = 0; t < to; ++t) count[t + 1] += count[t];
  std::vector<int> entries((size_t)from * dim), fill(count.begin(), count.end() - 1);
  // element order, then slot order; under transparent renumbering, the
  // application's element order
  const int *key = old_index_host(map->from);
  std::vector<int> order(from);
  for (int e = 0; e < from; ++e) order[e] = e;
  if (key) std::sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  for (int e : order)
    for (int k = 0; k < dim; ++k) {
      int t = host_map(map)[(size_t)e * dim + k];
      entries[fill[t]++] = e * dim + k;
    }
  std::vector<int> offsets(targets.size() + 1);
  for (size_t i = 0; i < targets.size(); ++i) offsets[i] = count[targets[i]];
  offsets[targets.size()] = count[to];
  P.ntargets = (int)targets.size();
  P.max_count = 0;
  for (size_t i = 0; i < targets.size(); ++i)
    P.max_count = std::max(P.max_count, offsets[i + 1] - offsets[i]);
  P.offsets = rec->mem.upload(offsets, stream);
  P.entries = rec->mem.upload(entries, stream);
  P.targets = rec->mem.upload(targets, stream);
  PCHECK(hipStreamSynchronize(stream));
  return rec;
}

int download_staged(const op2h_staged_plan *p, int *thrcol, int *blkcol,
                    int *offset, int *nelems) {
  std::vector<unsigned char> tc(std::max(p->set_size, 1));
  std::vector<int> bm(std::max(p->nblocks, 1));
  PCHECK(hipDeviceSynchronize());
  if (thrcol) {
    PCHECK(hipMemcpy(tc.data(), p->thrcol, p->set_size, hipMemcpyDeviceToHost));
    for (int e = 0; e < p->set_size; ++e) thrcol[e] = tc[e];
  }
  if (blkcol) {
    PCHECK(hipMemcpy(bm.data(), p->blkmap, p->nblocks * sizeof(int),
                     hipMemcpyDeviceToHost));
    for (int c = 0; c < p->ncolors; ++c)
      for (int k = p->colstart_h[c]; k < p->colstart_h[c + 1]; ++k) blkcol[bm[k]] = c;
  }
  if (offset)
    PCHECK(hipMemcpy(offset, p->offset, p->nblocks * sizeof(int), hipMemcpyDeviceToHost));
  if (nelems)
    PCHECK(hipMemcpy(nelems, p->nelems, p->nblocks * sizeof(int), hipMemcpyDeviceToHost));
  return 0;
}

int download_color(const op2h_color_plan *p, int *color_of_elem) {
  std::vector<int> reord(std::max(p->set_size, 1));
  PCHECK(hipDeviceSynchronize());
  PCHECK(hipMemcpy(reord.data(), p->col_reord, p->set_size * sizeof(int),
                   hipMemcpyDeviceToHost));
  for (int c = 0; c < p->ncolors; ++c)
    for (int k = p->col_offsets_h[c]; k < p->col_offsets_h[c + 1]; ++k)
      color_of_elem[reord[k]] = c;
  return 0;
}

template int *DeviceAllocs::upload<int>(const std::vector<int> &, hipStream_t);
template unsigned short *DeviceAllocs::upload<unsigned short>(
    const std::vector<unsigned short> &, hipStream_t);

}  // namespace op2h
