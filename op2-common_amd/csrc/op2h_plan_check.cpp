// op2h_plan_check.cpp — invariants of every plan the runtime builds.
//
// The reference validates each op_plan_core plan with op_plan_check
// (op2/src/core/op_rt_support.cpp:129-295, called at :1105): element totals,
// block offsets, the block permutation, the staging-list offsets and order,
// and that local ids lead back to the map entries; its thread- and
// block-colouring check is an empty stub (:290-294).  Here the same checks
// run on the staged plan, and the colouring checks are real: no two elements
// of one block with the same thread colour, and no two blocks of the same
// block colour, write one target; the flat colouring and the owner-compute
// and gather plans get the analogous checks (every increment listed once, in
// element order, under the block that owns its target).
//
// The checks work on host images of a plan (exported, so plans from any
// source can be checked); the runtime downloads each plan it builds and runs
// them when OP_diags > 0.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "op2_hip_rt.h"

namespace {

struct Report {
  int errors = 0;
  void check(bool ok, const char *what) {
    if (!ok) {
      std::printf(" *** OP_plan_check: %s error \n", what);
      ++errors;
    } else if (OP_diags > 6) {
      std::printf(" *** OP_plan_check: %s OK \n", what);
    }
    std::fflush(stdout);
  }
};

// The written targets of element e: (dat, target) over the writing accesses.
void written(const op2h_check_access *acc, int nacc, int e,
             std::vector<std::pair<int, int>> &out) {
  out.clear();
  for (int a = 0; a < nacc; ++a)
    if (acc[a].writes)
      out.push_back({acc[a].dat, acc[a].table[(size_t)e * acc[a].dim + acc[a].idx]});
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
}

bool accesses_valid(const op2h_check_access *acc, int nacc, int set_size) {
  for (int a = 0; a < nacc; ++a) {
    if (!acc[a].table || acc[a].dim < 1 || acc[a].idx < 0 || acc[a].idx >= acc[a].dim ||
        acc[a].dat < 0 || acc[a].to_size < 0)
      return false;
    for (size_t i = 0; i < (size_t)set_size * acc[a].dim; ++i)
      if (acc[a].table[i] < 0 || acc[a].table[i] >= acc[a].to_size) return false;
  }
  return true;
}

// Targets of one dat share a stamp array; dats are numbered densely.
struct Stamps {
  std::vector<std::vector<int>> tag, who;
  Stamps(const op2h_check_access *acc, int nacc) {
    int nd = 0;
    for (int a = 0; a < nacc; ++a) nd = std::max(nd, acc[a].dat + 1);
    tag.resize(nd);
    who.resize(nd);
    for (int a = 0; a < nacc; ++a) {
      auto &t = tag[acc[a].dat];
      if ((int)t.size() < acc[a].to_size) {
        t.assign(acc[a].to_size, -1);
        who[acc[a].dat].assign(acc[a].to_size, -1);
      }
    }
  }
};

}  // namespace

extern "C" int op2h_check_staged_plan(const op2h_staged_plan_host *p, int nacc,
                                      const op2h_check_access *acc) {
  Report r;
  const int n = p->set_size, nb = p->nblocks, nc = p->ncolors;
  if (!accesses_valid(acc, nacc, n) || nb < 0 || nc < 0 || n < 0) {
    r.check(false, "arguments");
    return r.errors;
  }
  long long total = 0;
  for (int b = 0; b < nb; ++b) total += p->nelems[b];
  r.check(total == n, "nelems");
  bool ok = true;
  long long run = 0;
  for (int b = 0; b < nb; ++b) {
    ok &= p->offset[b] == run;
    run += p->nelems[b];
  }
  r.check(ok, "offset");
  std::vector<int> perm(p->blkmap, p->blkmap + nb);
  std::sort(perm.begin(), perm.end());
  ok = true;
  for (int b = 0; b < nb; ++b) ok &= perm[b] == b;
  r.check(ok, "blkmap");
  ok = p->colstart[0] == 0 && p->colstart[nc] == nb;
  for (int c = 0; c < nc; ++c) ok &= p->colstart[c] <= p->colstart[c + 1];
  r.check(ok, "colstart");
  if (r.errors) return r.errors;  // the rest indexes through these
  // staging lists: offsets, strictly ascending targets in range
  ok = true;
  bool order = true;
  for (int s = 0; s < p->nmaps; ++s) {
    long long off = 0;
    for (int b = 0; b < nb; ++b) {
      ok &= p->ind_offs[s][b] == off;
      int last = -1;
      for (int k = 0; k < p->ind_sizes[s][b]; ++k) {
        int t = p->ind_map[s][off + k];
        order &= t > last && t < p->map_to_size[s];
        last = t;
      }
      off += p->ind_sizes[s][b];
    }
  }
  r.check(ok, "ind_offs");
  r.check(order, "ind_maps");
  if (r.errors) return r.errors;
  // local ids lead back to the map entries (the "most likely source of errors")
  int bad = 0;
  for (int l = 0; l < p->nlocs; ++l) {
    const int s = p->loc_map_slot[l], idx = p->loc_idx[l];
    if (s < 0 || s >= p->nmaps || idx < 0 || idx >= p->map_dim[s]) {
      ++bad;
      continue;
    }
    for (int b = 0; b < nb; ++b)
      for (int e = p->offset[b]; e < p->offset[b] + p->nelems[b]; ++e) {
        const int loc = p->loc_map[l][e];
        bad += loc >= p->ind_sizes[s][b] ||
               p->ind_map[s][p->ind_offs[s][b] + loc] !=
                   p->map_table[s][(size_t)e * p->map_dim[s] + idx];
      }
  }
  if (bad) std::printf(" *** OP_plan_check: %d maps error(s) \n", bad);
  else if (OP_diags > 6) std::printf(" *** OP_plan_check: maps     OK \n");
  std::fflush(stdout);
  r.errors += bad != 0;
  // thread colours: below the block's count, and no written target shared by
  // two elements of one colour within a block
  std::vector<std::pair<int, int>> w;
  std::vector<long long> keys;
  ok = true;
  bool conflict = false;
  for (int b = 0; b < nb && ok; ++b) {
    keys.clear();
    for (int e = p->offset[b]; e < p->offset[b] + p->nelems[b]; ++e) {
      const int c = p->thrcol[e];
      ok &= c < p->nthrcol[b];
      written(acc, nacc, e, w);
      for (auto &dt : w)  // (dat, target, colour) distinct per element
        keys.push_back((((long long)dt.first << 40) | ((long long)dt.second << 8)) | c);
    }
    std::sort(keys.begin(), keys.end());
    conflict |= std::adjacent_find(keys.begin(), keys.end()) != keys.end();
  }
  r.check(ok, "thread colour range");
  r.check(!conflict, "thread colouring");
  // block colours: no written target shared by two blocks of one colour
  Stamps st(acc, nacc);
  conflict = false;
  for (int c = 0; c < nc; ++c)
    for (int k = p->colstart[c]; k < p->colstart[c + 1]; ++k) {
      const int b = p->blkmap[k];
      for (int e = p->offset[b]; e < p->offset[b] + p->nelems[b]; ++e) {
        written(acc, nacc, e, w);
        for (auto &dt : w) {
          int &tag = st.tag[dt.first][dt.second], &who = st.who[dt.first][dt.second];
          conflict |= tag == c && who != b;
          tag = c;
          who = b;
        }
      }
    }
  r.check(!conflict, "block colouring");
  return r.errors;
}

extern "C" int op2h_check_color_plan(int set_size, int ncolors, const int *col_offsets,
                                     const int *col_reord, int nacc,
                                     const op2h_check_access *acc) {
  Report r;
  if (!accesses_valid(acc, nacc, set_size) || ncolors < 0) {
    r.check(false, "arguments");
    return r.errors;
  }
  bool ok = col_offsets[0] == 0 && col_offsets[ncolors] == set_size;
  for (int c = 0; c < ncolors; ++c) ok &= col_offsets[c] <= col_offsets[c + 1];
  r.check(ok, "col_offsets");
  std::vector<int> perm(col_reord, col_reord + set_size);
  std::sort(perm.begin(), perm.end());
  ok = true;
  for (int e = 0; e < set_size; ++e) ok &= perm[e] == e;
  r.check(ok, "col_reord");
  if (r.errors) return r.errors;
  Stamps st(acc, nacc);
  std::vector<std::pair<int, int>> w;
  bool conflict = false;
  for (int c = 0; c < ncolors; ++c)
    for (int k = col_offsets[c]; k < col_offsets[c + 1]; ++k) {
      const int e = col_reord[k];
      written(acc, nacc, e, w);
      for (auto &dt : w) {
        int &tag = st.tag[dt.first][dt.second], &who = st.who[dt.first][dt.second];
        conflict |= tag == c && who != e;
        tag = c;
        who = e;
      }
    }
  r.check(!conflict, "element colouring");
  return r.errors;
}

extern "C" int op2h_check_owner_plan(const op2h_owner_plan_host *p) {
  Report r;
  const int nb = p->nblocks, nto = p->to_size, nel = p->set_size, ninc = p->ninc;
  bool ok = nb >= 0 && nto >= 0 && nel >= 0 && ninc >= 1 && p->map_dim >= 1;
  for (int k = 0; k < ninc && ok; ++k) ok &= p->inc_idx[k] >= 0 && p->inc_idx[k] < p->map_dim;
  for (size_t i = 0; ok && i < (size_t)nel * p->map_dim; ++i)
    ok &= p->map_table[i] >= 0 && p->map_table[i] < nto;
  if (!ok) {
    r.check(false, "arguments");
    return r.errors;
  }
  ok = p->owned_offs[0] == 0 && p->elem_offs[0] == 0;
  for (int b = 0; b < nb; ++b)
    ok &= p->owned_offs[b] <= p->owned_offs[b + 1] && p->elem_offs[b] <= p->elem_offs[b + 1];
  const int n_owned = p->owned_offs[nb];
  ok &= p->csr_offs[0] == 0;
  for (int i = 0; ok && i < n_owned; ++i) ok &= p->csr_offs[i] <= p->csr_offs[i + 1];
  r.check(ok, "owner offsets");
  if (r.errors) return r.errors;
  // every target owned at most once, ascending within its block; every
  // incremented target owned
  std::vector<int> owner(nto, -1), count(nto, 0);
  for (int e = 0; e < nel; ++e)
    for (int k = 0; k < ninc; ++k) count[p->map_table[(size_t)e * p->map_dim + p->inc_idx[k]]]++;
  ok = true;
  for (int b = 0; b < nb; ++b)
    for (int i = p->owned_offs[b]; i < p->owned_offs[b + 1]; ++i) {
      const int t = p->owned[i];
      if (t < 0 || t >= nto || owner[t] >= 0 ||
          (i > p->owned_offs[b] && t <= p->owned[i - 1])) {
        ok = false;
        continue;
      }
      owner[t] = b;
    }
  for (int t = 0; t < nto; ++t) ok &= count[t] == 0 || owner[t] >= 0;
  r.check(ok, "owned targets");
  // instances ascending and in range; increments decode to their target, in
  // element order, each (element, increment) listed exactly once
  ok = true;
  auto order_of = [&](int e) {
    return p->elem_key && p->instances_by_key ? p->elem_key[e] : e;
  };
  for (int b = 0; b < nb; ++b)
    for (int i = p->elem_offs[b]; i < p->elem_offs[b + 1]; ++i)
      ok &= p->elems[i] >= 0 && p->elems[i] < nel &&
            (i == p->elem_offs[b] || order_of(p->elems[i]) > order_of(p->elems[i - 1]));
  r.check(ok, "instances");
  if (r.errors) return r.errors;
  bool lists = true, complete = true;
  std::vector<char> used;
  for (int b = 0; b < nb; ++b) {
    const int i0 = p->elem_offs[b], ni = p->elem_offs[b + 1] - i0;
    used.assign(ni, 0);
    for (int o = p->owned_offs[b]; o < p->owned_offs[b + 1]; ++o) {
      const int t = p->owned[o];
      long long last = -1;
      for (int j = p->csr_offs[o]; j < p->csr_offs[o + 1]; ++j) {
        const int ent = p->csr[j], inst = ent / ninc, k = ent % ninc;
        // element order: the element's own index, or its key when given
        const long long pos =
            inst < ni ? (p->elem_key ? (long long)p->elem_key[p->elems[i0 + inst]] : inst) * ninc + k
                      : -1;
        if (inst >= ni || pos <= last) {
          lists = false;
          continue;
        }
        last = pos;
        used[inst] = 1;
        lists &= p->map_table[(size_t)p->elems[i0 + inst] * p->map_dim + p->inc_idx[k]] == t;
        if (p->csr4 && j - p->csr_offs[o] < 4)
          lists &= p->csr4[4 * (size_t)o + (j - p->csr_offs[o])] == ent;
      }
      complete &= p->csr_offs[o + 1] - p->csr_offs[o] == count[t];
      if (p->csr4) {
        lists &= p->csr_offs[o + 1] - p->csr_offs[o] <= 4;
        for (int j = p->csr_offs[o + 1] - p->csr_offs[o]; j < 4; ++j)
          lists &= p->csr4[4 * (size_t)o + j] == 0xFFFF;
      }
    }
    for (int i = 0; i < ni; ++i) complete &= used[i] != 0;
  }
  r.check(lists, "increment lists");
  r.check(complete, "increment completeness");
  return r.errors;
}

extern "C" int op2h_check_gather_plan(int from_size, int to_size, int dim, const int *map,
                                      int ntargets, const int *offsets, const int *entries,
                                      const int *targets, const int *elem_key) {
  Report r;
  bool ok = from_size >= 0 && to_size >= 0 && dim >= 1 && ntargets >= 0 &&
            offsets[0] == 0 && offsets[ntargets] == from_size * dim;
  for (int i = 0; ok && i < ntargets; ++i)
    ok &= offsets[i] <= offsets[i + 1] && targets[i] >= 0 && targets[i] < to_size &&
          (i == 0 || targets[i] > targets[i - 1]);
  r.check(ok, "gather offsets");
  if (r.errors) return r.errors;
  ok = true;
  for (int i = 0; i < ntargets; ++i)
    for (int j = offsets[i]; j < offsets[i + 1]; ++j)
      ok &= entries[j] >= 0 && entries[j] < from_size * dim &&
            map[entries[j]] == targets[i] &&
            (j == offsets[i] ||
             (elem_key ? (long long)elem_key[entries[j] / dim] * dim + entries[j] % dim >
                             (long long)elem_key[entries[j - 1] / dim] * dim + entries[j - 1] % dim
                       : entries[j] > entries[j - 1]));
  r.check(ok, "gather entries");
  return r.errors;
}

// ---- the runtime's hooks: download a plan it has just built and check it ----
#include "op2h_plan_build.h"

namespace op2h {
namespace {

template <class T>
std::vector<T> fetch(const T *d, size_t n) {
  std::vector<T> h(std::max(n, (size_t)1));
  if (n && hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) {
    std::printf("op2hip: plan check could not read a plan back from the device\n");
    exit(-1);
  }
  return h;
}

std::vector<op2h_check_access> accesses(op_set set, int nargs, const op_arg *args) {
  LoopShape sh = loop_shape(set, nargs, args);
  std::vector<op2h_check_access> acc;
  for (int m = 0; m < nargs; ++m) {
    if (sh.dat_slot[m] < 0) continue;
    const op_map mp = args[m].map;
    acc.push_back({host_map(mp), mp->dim, args[m].idx, mp->to->size, sh.dat_slot[m],
                   sh.conflict[m] ? 1 : 0});
  }
  return acc;
}

void verdict(int errors, const char *kind, const char *name) {
  if (errors == 0) return;
  std::printf("op2hip: the %s plan of loop %s failed %d check(s)\n", kind, name, errors);
  exit(-1);
}

}  // namespace

void check_staged_or_exit(const StagedPlanRec &rec, const char *name, op_set set, int nargs,
                          const op_arg *args) {
  (void)hipDeviceSynchronize();
  const op2h_staged_plan &P = rec.plan;
  LoopShape sh = loop_shape(set, nargs, args);
  const int n = P.set_size, nb = P.nblocks;
  op2h_staged_plan_host h{};
  h.set_size = n;
  h.nblocks = nb;
  h.ncolors = P.ncolors;
  h.colstart = P.colstart_h;
  auto blkmap = fetch(P.blkmap, nb), offset = fetch(P.offset, nb), nelems = fetch(P.nelems, nb),
       nthrcol = fetch(P.nthrcol, nb);
  auto thrcol = fetch(P.thrcol, n);
  h.blkmap = blkmap.data();
  h.offset = offset.data();
  h.nelems = nelems.data();
  h.nthrcol = nthrcol.data();
  h.thrcol = thrcol.data();
  h.nmaps = P.nmaps;
  std::vector<std::vector<int>> io(P.nmaps), is(P.nmaps), im(P.nmaps);
  for (int s = 0; s < P.nmaps; ++s) {
    io[s] = fetch(P.ind_offs[s], nb);
    is[s] = fetch(P.ind_sizes[s], nb);
    long long tot = 0;
    for (int b = 0; b < nb; ++b) tot += is[s][b];
    im[s] = fetch(P.ind_map[s], (size_t)tot);
    h.map_table[s] = host_map(sh.maps[s]);
    h.map_dim[s] = sh.maps[s]->dim;
    h.map_to_size[s] = sh.maps[s]->to->size;
    h.ind_offs[s] = io[s].data();
    h.ind_sizes[s] = is[s].data();
    h.ind_map[s] = im[s].data();
  }
  h.nlocs = (int)sh.locs.size();
  std::vector<std::vector<unsigned short>> lm(h.nlocs);
  for (int l = 0; l < h.nlocs; ++l) {
    lm[l] = fetch(P.loc_map[l], n);
    h.loc_map[l] = lm[l].data();
    h.loc_idx[l] = sh.locs[l].second;
    h.loc_map_slot[l] =
        (int)(std::find(sh.maps.begin(), sh.maps.end(), sh.locs[l].first) - sh.maps.begin());
  }
  auto acc = accesses(set, nargs, args);
  verdict(op2h_check_staged_plan(&h, (int)acc.size(), acc.data()), "staged", name);
}

void check_color_or_exit(const ColorPlanRec &rec, const char *name, op_set set, int nargs,
                         const op_arg *args) {
  (void)hipDeviceSynchronize();
  const op2h_color_plan &P = rec.plan;
  auto reord = fetch(P.col_reord, P.set_size);
  auto acc = accesses(set, nargs, args);
  verdict(op2h_check_color_plan(P.set_size, P.ncolors, P.col_offsets_h, reord.data(),
                                (int)acc.size(), acc.data()),
          "colour", name);
}

void check_owner_or_exit(const OwnerPlanRec &rec, const char *name, int nargs,
                         const op_arg *args) {
  (void)hipDeviceSynchronize();
  const op2h_owner_plan &P = rec.plan;
  op2h_owner_plan_host h{};
  h.set_size = P.set_size;
  h.to_size = P.to_size;
  h.nblocks = P.nblocks;
  h.ninc = P.ninc;
  for (int m = 0; m < nargs; ++m)
    if (P.inc_of_arg[m] >= 0) {
      h.map_table = host_map(args[m].map);
      h.map_dim = args[m].map->dim;
      h.inc_idx[P.inc_of_arg[m]] = args[m].idx;
    }
  std::vector<int> oo(P.nblocks + 1), ow(std::max(rec.n_owned, 1)), eo(P.nblocks + 1),
      el(std::max(P.total, 1)), co(rec.n_owned + 1);
  std::vector<unsigned short> csr(std::max(rec.n_csr, 1));
  download_owner(rec, oo.data(), ow.data(), eo.data(), el.data(), co.data(), csr.data());
  std::vector<unsigned short> c4;
  if (P.csr4) c4 = fetch(P.csr4, 4 * (size_t)rec.n_owned);
  h.owned_offs = oo.data();
  h.owned = ow.data();
  h.elem_offs = eo.data();
  h.elems = el.data();
  h.csr_offs = co.data();
  h.csr = csr.data();
  h.csr4 = P.csr4 ? c4.data() : nullptr;
  for (int m = 0; m < nargs; ++m)
    if (P.inc_of_arg[m] >= 0) h.elem_key = old_index_host(args[m].map->from);
  h.instances_by_key = rec.key_instances ? 1 : 0;
  verdict(op2h_check_owner_plan(&h), "owner", name);
}

void check_gather_or_exit(const GatherPlanRec &rec, op_map map) {
  (void)hipDeviceSynchronize();
  const op2h_gather_plan &P = rec.plan;
  auto off = fetch(P.offsets, (size_t)P.ntargets + 1);
  auto ent = fetch(P.entries, (size_t)P.from_size * P.dim);
  auto tg = fetch(P.targets, (size_t)P.ntargets);
  verdict(op2h_check_gather_plan(P.from_size, P.to_size, P.dim, host_map(map), P.ntargets,
                                 off.data(), ent.data(), tg.data(), old_index_host(map->from)),
          "gather", map->name);
}

}  // namespace op2h
