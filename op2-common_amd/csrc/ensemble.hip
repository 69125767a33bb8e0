// Batched airfoil ensembles (include/op2amd_ensemble.h).
//
// Members are concatenated into one mesh; every loop of the time step is one
// persistent launch over all members' work items (runs of one member's
// elements, found through per-member offset tables), except in the looped
// baseline, which launches each loop once per member over that member's
// items.  res_calc and bres_calc use the owner-compute plan (each cell sums
// its edges' fluxes in edge order), and the reductions' partials are formed
// per tile, so results do not depend on the launch structure or the
// workgroup sizes.
#include "op2amd_ensemble.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "airfoil_device.h"
#include "op2h_plan_build.h"

#define ECHECK(x)                                                            \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      printf("op2hip ensemble: %s failed at %s:%d: %s\n", #x, __FILE__,      \
             __LINE__, hipGetErrorString(e_));                               \
      exit(-1);                                                              \
    }                                                                        \
  } while (0)

namespace {

using namespace op2_dev;

constexpr int kTile = 1024;          // cells per tile of the update chain
constexpr int kGroups = kTile / 64;  // reduction groups (one wave's cells) per tile
constexpr int kPatch = 256;          // default cells per owner patch of res_calc

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

double now() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// ---- kernels ------------------------------------------------------------

// Device-clock stamps of work items, for per-member step times inside the
// batched step: item i of the launch at offset `off` of the step writes
// {begin, end} (wall_clock64 ticks) at t[2 * (slot * per_step + off + i)],
// slot = steps completed so far (the history counter) mod capacity.  t ==
// nullptr: no stamps.
struct Stamps {
  unsigned long long *t;
  const unsigned long long *count;
  int per_step, capacity, off;
};

__device__ __forceinline__ unsigned long long stamp_now(const Stamps &S) {
  return S.t ? (unsigned long long)wall_clock64() : 0ull;
}

__device__ __forceinline__ void stamp(const Stamps &S, int item, unsigned long long t0) {
  const unsigned long long t1 = (unsigned long long)wall_clock64();
  const size_t k = (size_t)(*S.count % (unsigned long long)S.capacity) * S.per_step + S.off + item;
  S.t[2 * k] = t0;
  S.t[2 * k + 1] = t1;
}

// res_calc, persistent over the owner patches [blo, bhi) of the members'
// owner partitions: each edge evaluated once per patch it touches, each
// cell's increments summed in edge order (op2_dev::owner_res_block), so every
// member gets its sequential executor's bits.  zero: every cell is owned, so
// res holds the update's zeros in every cell (not stored: dead) and each sum
// starts from +0.0 without reading it.  LDS: fluxes [max_e][4].
__global__ void __launch_bounds__(1024)
    k_ens_res_owner(const double *__restrict x, const double *__restrict q,
                    const double *__restrict adt, double *__restrict res, OwnerArgs P,
                    int blo, int bhi, int zero, Stamps S) {
  extern __shared__ double lf[];
  for (int b = blo + xcd_block(kXcds); b < bhi; b += gridDim.x) {
    const unsigned long long t0 = stamp_now(S);
    owner_res_block<false, true>(b, x, q, adt, res, P, 0, 0, lf, zero != 0);
    __syncthreads();  // the next patch rewrites lf
    if (S.t && threadIdx.x == 0) stamp(S, b, t0);
  }
}

// res_calc and bres_calc in one pass, persistent over the owner patches
// [blo, bhi): each owned cell sums its interior edges' fluxes in edge order,
// then its boundary edges' (bres_calc with the member's own free stream) in
// boundary-edge order, the order of res_calc then bres_calc
// (op2_dev::owner_res_bres_block).  zero as for k_ens_res_owner.
__global__ void __launch_bounds__(1024)
    k_ens_res_bres_owner(const double *__restrict x, const double *__restrict q,
                         const double *__restrict adt, double *__restrict res, OwnerArgs P,
                         BoundaryArgs B, int blo, int bhi, int zero, Stamps S) {
  extern __shared__ double lf[];
  for (int b = blo + xcd_block(kXcds); b < bhi; b += gridDim.x) {
    const unsigned long long t0 = stamp_now(S);
    owner_res_bres_block<false, true>(b, x, q, adt, res, P, B, 0, 0, lf, zero != 0);
    __syncthreads();  // the next patch rewrites lf
    if (S.t && threadIdx.x == 0) stamp(S, b, t0);
  }
}

// bres_calc, persistent over items [ilo, ihi) of <= 64 boundary cells of one
// member ({first, end} target), one item per wave: one lane per boundary cell
// adds its boundary edges in edge order.
__global__ void k_ens_bres(const double *__restrict x, const double *__restrict q,
                           const double *__restrict adt, double *res,
                           const int2 *__restrict pbedge, const int *__restrict bound,
                           const int *__restrict bmesh, const double *__restrict qinf,
                           const int *__restrict offsets, const int *__restrict entries,
                           const int *__restrict targets, const int2 *__restrict items,
                           int ilo, int ihi, Stamps S) {
  const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  const int stride = gridDim.x * nw;
  for (int i = ilo + xcd_block(kXcds) * nw + (threadIdx.x >> 6); i < ihi; i += stride) {
    const unsigned long long t0 = stamp_now(S);
    const int2 it = items[i];
    const int t = it.x + lane;
    if (t < it.y) {
      const int cell = targets[t];
      double acc[4], q1[4];
      ld4<false>(res, cell, 0, acc);
      ld4<false>(q, cell, 0, q1);
      const double a1 = adt[cell];
      for (int k = offsets[t]; k < offsets[t + 1]; ++k) {
        const int e = entries[k];
        const int2 bn = pbedge[e];
        double x1[2], x2[2], r[4] = {0.0, 0.0, 0.0, 0.0};
        ld2<false>(x, bn.x, 0, x1);
        ld2<false>(x, bn.y, 0, x2);
        const int bd = bound[e];
        op2_k::bres_calc_q(x1, x2, q1, &a1, r, &bd, qinf + 4 * (size_t)bmesh[e]);
        for (int d = 0; d < 4; ++d) acc[d] += r[d];
      }
      st4<false>(res, cell, 0, acc);
    }
    if (S.t && lane == 0) stamp(S, i, t0);
  }
}

struct MaxLoc {
  double v;
  int loc;
};

__device__ __forceinline__ MaxLoc maxloc_combine(MaxLoc a, MaxLoc b) {
  if (b.v > a.v || (b.v == a.v && b.loc < a.loc)) return b;
  return a;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ MaxLoc wave_maxloc(MaxLoc m) {
  for (int o = 32; o > 0; o >>= 1) {
    MaxLoc t;
    t.v = __shfl_xor(m.v, o, 64);
    t.loc = __shfl_xor(m.loc, o, 64);
    m = maxloc_combine(m, t);
  }
  return m;
}

// update, save_soln and adt_calc of each cell in one pass, persistent over
// tiles [tlo, thi) of <= kTile cells of one member ({first cell, member}):
//  UPD   q = update(qold, res, adt) (update.h), the member's rms/maxerr/errloc
//        partials when RED: one per 64-cell group (a wave's sum), at
//        part_*[tile * kGroups + group];
//  SAVE  qold = q (save_soln.h), the next time step's;
//  then  adt = adt_calc(x, q) (adt_calc.h) for the next half step.
// DEAD: the update's zeros in res are not stored (res_calc, which comes next,
// starts every cell's sum from +0.0).  prow/pbase (or null): the cells' nodes
// as 16-bit offsets from the lowest node of their tile's 64-cell group.  A workgroup of any multiple of 64
// threads covers a tile in kTile / blockDim.x rounds; the groups are fixed by
// the tile, so the bits do not depend on the workgroup size.  qold and adt
// are read before they are written, by the same thread for the same cell: not
// restrict-qualified.
template <bool UPD, bool SAVE, bool RED, bool DEAD>
__global__ void __launch_bounds__(kTile)
    k_ens_chain(const double *__restrict x, double *qold, double *__restrict q,
                double *__restrict res, double *adt, const int4 *__restrict pcell,
                const uint2 *__restrict prow, const int *__restrict pbase,
                const int2 *__restrict tiles, const int *__restrict coff, int tlo, int thi,
                double *__restrict part_sum, double *__restrict part_max,
                int *__restrict part_loc, Stamps S) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = tlo + xcd_block(kXcds); t < thi; t += gridDim.x) {
    const unsigned long long t0 = stamp_now(S);
    const int2 tl = tiles[t];
    const int base = coff[tl.y], end = min(tl.x + kTile, coff[tl.y + 1]);
    for (int r = 0; r < kTile; r += blockDim.x) {
      const int c = tl.x + r + (int)threadIdx.x;
      double rms = 0.0;
      MaxLoc ml{-INFINITY, 0x7fffffff};
      if (c < end) {
        double qn[4];
        if (UPD) {
          double qo[4], rr[4], a = adt[c], mx = ml.v;
          int loc = ml.loc, local = c - base;  // index within the member
          ld4<false>(qold, c, 0, qo);
          ld4<false>(res, c, 0, rr);
          op2_k::update(qo, qn, rr, &a, &rms, &mx, &local, &loc);
          ml.v = mx;
          ml.loc = loc;
          st4<false>(q, c, 0, qn);
          if (!DEAD) st4<false>(res, c, 0, rr);
        } else {
          ld4<false>(q, c, 0, qn);
        }
        if (SAVE) st4<false>(qold, c, 0, qn);
        int4 cn;
        if (prow) {  // compact rows: 16-bit offsets from the 64-cell group's base
          const uint2 u = prow[c];
          const int b = pbase[(size_t)t * kGroups + ((r + (int)threadIdx.x) >> 6)];
          cn = make_int4(b + (int)(u.x & 0xFFFFu), b + (int)(u.x >> 16),
                         b + (int)(u.y & 0xFFFFu), b + (int)(u.y >> 16));
        } else {
          cn = pcell[c];
        }
        double x1[2], x2[2], x3[2], x4[2], a2;
        ld2<false>(x, cn.x, 0, x1);
        ld2<false>(x, cn.y, 0, x2);
        ld2<false>(x, cn.z, 0, x3);
        ld2<false>(x, cn.w, 0, x4);
        op2_k::adt_calc(x1, x2, x3, x4, qn, &a2);
        adt[c] = a2;
      }
      if (RED) {
        rms = wave_sum(rms);
        ml = wave_maxloc(ml);
        if (lane == 0) {
          const size_t g = (size_t)t * kGroups + (r >> 6) + wave;
          part_sum[g] = rms;
          part_max[g] = ml.v;
          part_loc[g] = ml.loc;
        }
      }
    }
    // every wave finishes the tile before any starts the next, which is
    // gridDim.x tiles further on: the workgroup's loads and stores stay on
    // one contiguous run of memory (without this barrier, drifting waves
    // spread over several runs and the loop took 10-20 % longer)
    __syncthreads();
    if (S.t && threadIdx.x == 0) stamp(S, t, t0);
  }
}

// One workgroup per member: its tiles' partials, each summed over its groups
// in order, then over the tiles.
__global__ void __launch_bounds__(256)
    k_ens_finish(const double *__restrict part_sum, const double *__restrict part_max,
                 const int *__restrict part_loc, const int *__restrict mesh_tile_off,
                 int mlo, double *__restrict slots, Stamps S) {
  __shared__ double s_sum[4];
  __shared__ double s_max[4];
  __shared__ int s_loc[4];
  const int m = mlo + blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const unsigned long long t0 = stamp_now(S);
  double tot = 0.0;
  MaxLoc tm{-INFINITY, 0x7fffffff};
  for (int t = mesh_tile_off[m] + tid; t < mesh_tile_off[m + 1]; t += 256) {
    // the tile's partial: its groups in order
    double ts = 0.0;
    MaxLoc tg{-INFINITY, 0x7fffffff};
    for (int g = 0; g < kGroups; ++g) {
      const size_t k = (size_t)t * kGroups + g;
      ts += part_sum[k];
      tg = maxloc_combine(tg, MaxLoc{part_max[k], part_loc[k]});
    }
    tot += ts;
    tm = maxloc_combine(tm, tg);
  }
  tot = wave_sum(tot);
  tm = wave_maxloc(tm);
  if (lane == 0) {
    s_sum[wave] = tot;
    s_max[wave] = tm.v;
    s_loc[wave] = tm.loc;
  }
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    MaxLoc mm{-INFINITY, 0x7fffffff};
    for (int w = 0; w < 4; ++w) {
      t += s_sum[w];
      mm = maxloc_combine(mm, MaxLoc{s_max[w], s_loc[w]});
    }
    slots[3 * (size_t)m] = t;
    slots[3 * (size_t)m + 1] = mm.v;
    slots[3 * (size_t)m + 2] = (double)mm.loc;
    if (S.t) stamp(S, m, t0);
  }
}

// history[record][k] = slots[k]; one workgroup; the counter advances once.
__global__ void k_ens_history(const double *__restrict slots, int nwords,
                              double *__restrict hist, unsigned long long *count,
                              int capacity) {
  unsigned long long c = *count;
  size_t base = (size_t)(c % (unsigned long long)capacity) * nwords;
  for (int k = threadIdx.x; k < nwords; k += blockDim.x) hist[base + k] = slots[k];
  __syncthreads();
  if (threadIdx.x == 0) *count = c + 1;
}

// ---- host -----------------------------------------------------------------

template <class T> T *dev_upload(const std::vector<T> &v, std::vector<void *> &keep) {
  void *p = nullptr;
  ECHECK(hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(T)));
  if (!v.empty()) ECHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  keep.push_back(p);
  return (T *)p;
}

template <class T> T *dev_alloc(size_t n, std::vector<void *> &keep) {
  void *p = nullptr;
  ECHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  ECHECK(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
  ECHECK(hipDeviceSynchronize());  // done before the non-blocking stream uses it
  keep.push_back(p);
  return (T *)p;
}

}  // namespace

extern "C" int op2amd_ensemble_describe(const op2amd_ensemble_spec *spec,
                                        op2amd_ensemble_member *members) {
  if (!spec || !members || spec->nmesh < 1 || spec->min_I < 2 || spec->min_J < 2 ||
      spec->max_I < spec->min_I)
    return 1;
  SplitMix64 rng(spec->seed);
  const double smax = (double)spec->max_I / spec->min_I;
  // constants of airfoil.cpp:149-159
  const double gam = 1.4f, gm1 = gam - 1.0f, p = 1.0f, r = 1.0f;
  for (int m = 0; m < spec->nmesh; ++m) {
    double s = 1.0 + (smax - 1.0) * rng.uniform();
    double um = rng.uniform(), ua = rng.uniform();
    op2amd_ensemble_member &mb = members[m];
    mb.I = std::max(2, (int)std::lround(spec->min_I * s));
    mb.J = std::max(2, (int)std::lround(spec->min_J * s));
    mb.ncell = mb.I * mb.J;
    mb.mach = spec->mach_lo + (spec->mach_hi - spec->mach_lo) * um;
    mb.alpha_deg = spec->alpha_lo_deg + (spec->alpha_hi_deg - spec->alpha_lo_deg) * ua;
    double a = mb.alpha_deg * (M_PI / 180.0);
    double u = sqrt(gam * p / r) * mb.mach;
    double e = p / (r * gm1) + 0.5f * u * u;
    mb.qinf[0] = r;
    mb.qinf[1] = r * u * cos(a);
    mb.qinf[2] = r * u * sin(a);
    mb.qinf[3] = r * e;
  }
  return 0;
}

extern "C" int op2amd_ensemble_run(const op2amd_ensemble_spec *spec,
                                   const op2amd_ensemble_opts *o,
                                   op2amd_ensemble_result *res) {
  if (!spec || !o || !res || o->niter < 1 || o->warmup < 0 || o->warmup >= o->niter ||
      o->max_workgroups < 0 || (o->patch_cells != 0 && (o->patch_cells < 64 || o->patch_cells > 1024)))
    return 1;
  const int patch = o->patch_cells ? o->patch_cells : kPatch;
  double t_setup = now();
  const int M = spec->nmesh;
  std::vector<op2amd_ensemble_member> mem(M);
  if (op2amd_ensemble_describe(spec, mem.data())) return 1;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    printf("op2hip ensemble: no HIP device available\n");
    return 2;
  }
  ECHECK(hipSetDevice(0));

  // ---- concatenate the members ----
  std::vector<int> noff(M + 1, 0), coff(M + 1, 0), eoff(M + 1, 0), boff(M + 1, 0);
  std::vector<double> x;
  std::vector<int4> cell;
  std::vector<int2> edge, ecell, bedge;
  std::vector<int> becell, bound, bmesh;
  for (int m = 0; m < M; ++m) {
    op2amd_mesh mesh;
    if (op2amd_mesh_generate(mem[m].I, mem[m].J, spec->seed + 1 + (uint64_t)m,
                             spec->ordering, &mesh))
      return 3;
    const int n0 = noff[m], c0 = coff[m];
    x.insert(x.end(), mesh.x, mesh.x + 2 * (size_t)mesh.nnode);
    for (int c = 0; c < mesh.ncell; ++c)
      cell.push_back(make_int4(mesh.cell[4 * c] + n0, mesh.cell[4 * c + 1] + n0,
                               mesh.cell[4 * c + 2] + n0, mesh.cell[4 * c + 3] + n0));
    for (int e = 0; e < mesh.nedge; ++e) {
      edge.push_back(make_int2(mesh.edge[2 * e] + n0, mesh.edge[2 * e + 1] + n0));
      ecell.push_back(make_int2(mesh.ecell[2 * e] + c0, mesh.ecell[2 * e + 1] + c0));
    }
    for (int b = 0; b < mesh.nbedge; ++b) {
      bedge.push_back(make_int2(mesh.bedge[2 * b] + n0, mesh.bedge[2 * b + 1] + n0));
      becell.push_back(mesh.becell[b] + c0);
      bound.push_back(mesh.bound[b]);
      bmesh.push_back(m);
    }
    noff[m + 1] = n0 + mesh.nnode;
    coff[m + 1] = c0 + mesh.ncell;
    eoff[m + 1] = eoff[m] + mesh.nedge;
    boff[m + 1] = boff[m] + mesh.nbedge;
    op2amd_mesh_free(&mesh);
  }
  const int Nc = coff[M], Ne = eoff[M], Nb = boff[M];
  res->total_cells = Nc;
  res->total_edges = Ne;

  // res_calc: each member's owner partition (patches of `patch` cells that
  // never straddle two members), concatenated; 8-byte instance records where a
  // patch's cells and nodes span at most 65535 indices
  std::vector<int> mblk(M + 1, 0), owned_offs{0}, owned, elem_offs{0};
  std::vector<int4> inst;
  std::vector<uint2> inst8;
  std::vector<int2> pbase;
  std::vector<ushort4> csr4;
  int max_e = 0;
  {
    std::vector<int> lmap;
    const int inc[2] = {0, 1};
    for (int m = 0; m < M; ++m) {
      const int c0 = coff[m], e0 = eoff[m], nel = eoff[m + 1] - e0;
      lmap.resize(2 * (size_t)nel);
      for (int e = 0; e < nel; ++e) {
        lmap[2 * (size_t)e] = ecell[e0 + e].x - c0;
        lmap[2 * (size_t)e + 1] = ecell[e0 + e].y - c0;
      }
      op2h::OwnerLists L;
      std::string err;
      if (!op2h::owner_lists(coff[m + 1] - c0, nel, 2, lmap.data(), 2, inc, patch, L, err)) {
        printf("op2hip ensemble: member %d: %s\n", m, err.c_str());
        return 4;
      }
      if (L.max_incs > 4) {
        printf("op2hip ensemble: a cell has more than 4 edges\n");
        return 4;
      }
      max_e = std::max(max_e, L.max_elems);
      for (int b = 0; b < L.nblocks; ++b) {
        const size_t first = inst.size();
        int cmin = INT_MAX, cmax = -1, nmin = INT_MAX, nmax = -1;
        for (int j = L.elem_offs[b]; j < L.elem_offs[b + 1]; ++j) {
          const int e = e0 + L.elems[j];
          const int4 r = make_int4(edge[e].x, edge[e].y, ecell[e].x, ecell[e].y);
          inst.push_back(r);
          nmin = std::min(nmin, std::min(r.x, r.y));
          nmax = std::max(nmax, std::max(r.x, r.y));
          cmin = std::min(cmin, std::min(r.z, r.w));
          cmax = std::max(cmax, std::max(r.z, r.w));
        }
        const bool fits = inst.size() > first && cmax - cmin <= 0xFFFF && nmax - nmin <= 0xFFFF;
        pbase.push_back(fits ? make_int2(cmin, nmin) : make_int2(0, -1));
        for (size_t j = first; j < inst.size(); ++j) {
          const int4 r = inst[j];
          inst8.push_back(fits ? make_uint2((unsigned)(r.x - nmin) | ((unsigned)(r.y - nmin) << 16),
                                            (unsigned)(r.z - cmin) | ((unsigned)(r.w - cmin) << 16))
                               : make_uint2(0u, 0u));
        }
        elem_offs.push_back((int)inst.size());
        for (int i = L.owned_offs[b]; i < L.owned_offs[b + 1]; ++i) {
          owned.push_back(c0 + L.owned[i]);
          unsigned short v[4] = {0xFFFF, 0xFFFF, 0xFFFF, 0xFFFF};
          for (int k = L.csr_offs[i]; k < L.csr_offs[i + 1]; ++k) v[k - L.csr_offs[i]] = L.csr[k];
          csr4.push_back(make_ushort4(v[0], v[1], v[2], v[3]));
        }
        owned_offs.push_back((int)owned.size());
      }
      mblk[m + 1] = mblk[m] + L.nblocks;
    }
  }
  // bres: boundary cells in cell order with their bedges in bedge order
  std::vector<int> bcount(Nc, 0);
  for (int b = 0; b < Nb; ++b) bcount[becell[b]]++;
  std::vector<int> btargets, boffsets{0}, bentries(Nb);
  std::vector<int> bpos(Nc, -1);
  for (int c = 0; c < Nc; ++c)
    if (bcount[c]) {
      bpos[c] = (int)btargets.size();
      btargets.push_back(c);
      boffsets.push_back(boffsets.back() + bcount[c]);
    }
  {
    std::vector<int> fill(boffsets.begin(), boffsets.end() - 1);
    for (int b = 0; b < Nb; ++b) bentries[fill[bpos[becell[b]]]++] = b;
  }
  std::vector<int> mesh_btarget_off(M + 1, 0);
  for (int m = 0; m <= M; ++m)
    mesh_btarget_off[m] = (int)(std::lower_bound(btargets.begin(), btargets.end(), coff[m]) -
                                btargets.begin());
  // work items, runs of one member's elements, with per-member offsets into
  // each table: tiles of <= kTile cells for update+save_soln+adt_calc, the
  // owner patches (mblk) for res_calc, <= 64 boundary cells for bres_calc
  std::vector<int2> tiles, bitems;  // {first cell, member}, {first, end target}
  std::vector<int> mesh_tile_off{0}, mesh_bitem_off{0};
  for (int m = 0; m < M; ++m) {
    for (int c = coff[m]; c < coff[m + 1]; c += kTile) tiles.push_back(make_int2(c, m));
    mesh_tile_off.push_back((int)tiles.size());
    for (int t = mesh_btarget_off[m]; t < mesh_btarget_off[m + 1]; t += 64)
      bitems.push_back(make_int2(t, std::min(t + 64, mesh_btarget_off[m + 1])));
    mesh_bitem_off.push_back((int)bitems.size());
  }
  const int ntiles = (int)tiles.size(), npatch = mblk[M];
  // compact cell rows for the chains: per 64-cell group of a tile its lowest
  // node, per cell four 16-bit offsets from it (8 bytes instead of 16), when
  // every group's nodes span at most 65535
  std::vector<uint2> crow(Nc);
  std::vector<int> cbase((size_t)ntiles * kGroups, 0);
  bool compact = true;
  for (int t = 0; compact && t < ntiles; ++t) {
    const int c0 = tiles[t].x, c1 = std::min(c0 + kTile, coff[tiles[t].y + 1]);
    for (int g = 0; compact && c0 + 64 * g < c1; ++g) {
      const int g0 = c0 + 64 * g, g1 = std::min(c1, g0 + 64);
      int lo = INT_MAX, hi = INT_MIN;
      for (int c = g0; c < g1; ++c) {
        const int4 v = cell[c];
        lo = std::min(std::min(lo, std::min(v.x, v.y)), std::min(v.z, v.w));
        hi = std::max(std::max(hi, std::max(v.x, v.y)), std::max(v.z, v.w));
      }
      if ((long long)hi - lo > 0xFFFF) compact = false;
      cbase[(size_t)t * kGroups + g] = lo;
      for (int c = g0; c < g1; ++c) {
        const int4 v = cell[c];
        crow[c] = make_uint2((unsigned)(v.x - lo) | ((unsigned)(v.y - lo) << 16),
                             (unsigned)(v.z - lo) | ((unsigned)(v.w - lo) << 16));
      }
    }
  }
  // every cell owned by a patch: res_calc writes all of res, so the update
  // need not zero it (dead res) and res_calc need not read it
  const bool dead = (int)owned.size() == Nc;
  // bres_calc fused into res_calc: per owned cell its interior increments,
  // then its boundary edges as 0x8000 | j, j indexing the patch's boundary
  // edges (in cell order, each cell's in bedge order); possible when every
  // boundary cell is owned and no cell has more than four entries
  std::vector<ushort4> brec(owned.size());
  std::vector<int> bedge_offs(npatch + 1, 0), bedge_list;
  bool fused = dead && 2 * max_e < 0x8000 && !o->separate_bres;
  for (int b = 0; fused && b < npatch; ++b) {
    bedge_offs[b] = (int)bedge_list.size();
    for (int i = owned_offs[b]; fused && i < owned_offs[b + 1]; ++i) {
      unsigned short ent[4] = {csr4[i].x, csr4[i].y, csr4[i].z, csr4[i].w};
      int k = 0;
      while (k < 4 && ent[k] != 0xFFFF) ++k;
      const int t = bpos[owned[i]];
      if (t >= 0)
        for (int j = boffsets[t]; j < boffsets[t + 1]; ++j) {
          const int local = (int)bedge_list.size() - bedge_offs[b];
          if (k == 4 || local >= 0x7FFF) {
            fused = false;
            break;
          }
          ent[k++] = (unsigned short)(0x8000 | local);
          bedge_list.push_back(bentries[j]);
        }
      brec[i] = make_ushort4(ent[0], ent[1], ent[2], ent[3]);
    }
  }
  bedge_offs[npatch] = (int)bedge_list.size();
  fused = fused && (int)bedge_list.size() == Nb;
  // bres_calc's own items when it is not fused
  const int nbitem = fused ? 0 : (int)bitems.size();

  // ---- device data ----
  std::vector<void *> keep;
  std::vector<double> q0(4 * (size_t)Nc), qinf(4 * (size_t)M);
  for (int m = 0; m < M; ++m)
    for (int k = 0; k < 4; ++k) qinf[4 * m + k] = mem[m].qinf[k];
  for (int m = 0; m < M; ++m)
    for (int c = coff[m]; c < coff[m + 1]; ++c)
      for (int k = 0; k < 4; ++k) q0[4 * (size_t)c + k] = mem[m].qinf[k];
  double *d_x = dev_upload(x, keep);
  double *d_q = dev_upload(q0, keep);
  double *d_qold = dev_alloc<double>(4 * (size_t)Nc, keep);
  double *d_res = dev_alloc<double>(4 * (size_t)Nc, keep);
  double *d_adt = dev_alloc<double>(Nc, keep);
  int4 *d_cell = dev_upload(cell, keep);
  const uint2 *d_crow = compact ? dev_upload(crow, keep) : nullptr;
  const int *d_cbase = compact ? dev_upload(cbase, keep) : nullptr;
  OwnerArgs OA{};
  OA.inst = dev_upload(inst, keep);
  OA.base = dev_upload(pbase, keep);
  OA.inst8 = dev_upload(inst8, keep);
  OA.elem_offs = dev_upload(elem_offs, keep);
  OA.owned_offs = dev_upload(owned_offs, keep);
  OA.owned = dev_upload(owned, keep);
  OA.csr4 = dev_upload(csr4, keep);
  OA.max_e = max_e;
  const size_t res_lds = sizeof(double) * 4 * (size_t)max_e;
  int2 *d_bedge = dev_upload(bedge, keep);
  int *d_bound = dev_upload(bound, keep);
  int *d_bmesh = dev_upload(bmesh, keep);
  double *d_qinf = dev_upload(qinf, keep);
  int *d_btargets = dev_upload(btargets, keep);
  int *d_boffsets = dev_upload(boffsets, keep);
  int *d_bentries = dev_upload(bentries, keep);
  int2 *d_bitems = dev_upload(bitems, keep);
  BoundaryArgs BA{};
  if (fused) {
    BA.rec = dev_upload(brec, keep);
    BA.offs = dev_upload(bedge_offs, keep);
    BA.edges = dev_upload(bedge_list, keep);
    BA.pbedge = d_bedge;
    BA.bound = d_bound;
    BA.bmesh = d_bmesh;
    BA.qinf = d_qinf;
  }
  int2 *d_tiles = dev_upload(tiles, keep);
  int *d_coff = dev_upload(coff, keep);
  int *d_mtoff = dev_upload(mesh_tile_off, keep);
  double *d_psum = dev_alloc<double>((size_t)ntiles * kGroups, keep);
  double *d_pmax = dev_alloc<double>((size_t)ntiles * kGroups, keep);
  int *d_ploc = dev_alloc<int>((size_t)ntiles * kGroups, keep);
  double *d_slots = dev_alloc<double>(3 * (size_t)M, keep);
  const int hist_words = 3 * M;
  const int capacity = o->niter;
  double *d_hist = dev_alloc<double>((size_t)hist_words * capacity, keep);
  unsigned long long *d_count = dev_alloc<unsigned long long>(1, keep);
  // per-step stamp layout, in launch order: res_calc, bres_calc, the chain,
  // twice, then the per-member finish
  const int off_res[2] = {0, npatch + nbitem + ntiles};
  const int off_bres[2] = {npatch, 2 * npatch + nbitem + ntiles};
  const int off_chain[2] = {npatch + nbitem, 2 * (npatch + nbitem) + ntiles};
  const int off_fin = 2 * (npatch + nbitem + ntiles), per_step = off_fin + M;
  const bool stamps = o->batched && o->member_timing;
  unsigned long long *d_stamp =
      stamps ? dev_alloc<unsigned long long>(2 * (size_t)per_step * capacity, keep) : nullptr;
  {
    // constants of airfoil.cpp:149-152 in this translation unit's copies
    double gam = 1.4f, gm1 = gam - 1.0f, cfl = 0.9f, eps = 0.05f;
    ECHECK(hipMemcpyToSymbol(HIP_SYMBOL(gam_d), &gam, sizeof(double)));
    ECHECK(hipMemcpyToSymbol(HIP_SYMBOL(gm1_d), &gm1, sizeof(double)));
    ECHECK(hipMemcpyToSymbol(HIP_SYMBOL(cfl_d), &cfl, sizeof(double)));
    ECHECK(hipMemcpyToSymbol(HIP_SYMBOL(eps_d), &eps, sizeof(double)));
  }
  hipStream_t st;
  ECHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  ECHECK(hipDeviceSynchronize());

  int nsm = 0;
  ECHECK(hipDeviceGetAttribute(&nsm, hipDeviceAttributeMultiprocessorCount, 0));
  using ChainFn = void (*)(const double *, double *, double *, double *, double *,
                           const int4 *, const uint2 *, const int *, const int2 *, const int *,
                           int, int, double *, double *, int *, Stamps);
  const ChainFn k_prologue = k_ens_chain<false, true, false, false>;
  const ChainFn k_chain0 = dead ? k_ens_chain<true, false, false, true>
                                : k_ens_chain<true, false, false, false>;
  const ChainFn k_chain1 = dead ? k_ens_chain<true, true, true, true>
                                : k_ens_chain<true, true, true, false>;
  // persistent grids: as many workgroups as the GPU holds at once for the
  // kernel and workgroup size (occupancy), capped by the launch's items
  auto resident = [&](const void *fn, int block, size_t lds) {
    int nb = 0;
    ECHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, lds));
    if (nb < 1) {
      printf("op2hip ensemble: a kernel does not fit a compute unit (block %d, LDS %zu)\n",
             block, lds);
      exit(-1);
    }
    return o->max_workgroups > 0 ? std::min(nb * nsm, o->max_workgroups) : nb * nsm;
  };
  // workgroup sizes of the first half step's chain (update+adt_calc),
  // res_calc (its patch size, fixed by the partition), bres_calc and the
  // second half step's chain (update+save_soln+adt_calc): the defaults, or
  // the autotuner's picks; no kernel's result depends on them
  int blk[4] = {256, patch, 128, 256};
  int cap[4] = {0, 0, 0, 0};
  auto set_caps = [&]() {
    cap[0] = resident((const void *)k_chain0, blk[0], 0);
    cap[1] = resident(fused ? (const void *)k_ens_res_bres_owner : (const void *)k_ens_res_owner,
                      blk[1], res_lds);
    cap[2] = resident((const void *)k_ens_bres, blk[2], 0) * (blk[2] / 64);  // waves
    cap[3] = resident((const void *)k_chain1, blk[3], 0);
  };
  set_caps();
  const int res_zero = dead ? 1 : 0;
  auto S_at = [&](int off) {
    return Stamps{stamps ? d_stamp : nullptr, d_count, per_step, capacity, off};
  };
  // save_soln and adt_calc of members [mlo, mhi), once before the first step
  // (each step's last chain does both for the next step)
  auto prologue = [&](int mlo, int mhi) {
    const int t0 = mesh_tile_off[mlo], t1 = mesh_tile_off[mhi];
    if (t1 > t0)
      k_prologue<<<std::min(t1 - t0, cap[0]), blk[0], 0, st>>>(
          d_x, d_qold, d_q, d_res, d_adt, d_cell, d_crow, d_cbase, d_tiles, d_coff, t0, t1,
          d_psum, d_pmax, d_ploc, Stamps{nullptr, d_count, per_step, capacity, 0});
  };
  // one time step of members [mlo, mhi): twice res_calc and bres_calc (one
  // launch when fused) and the chain, then the per-member reductions; with
  // `ev`, res_calc, bres_calc and the chain of each half step are bracketed by
  // events ev[2j], ev[2j+1] in that order (bres_calc's pair brackets nothing
  // when fused)
  auto step_range = [&](int mlo, int mhi, hipEvent_t *ev = nullptr) {
    const int p0 = mblk[mlo], p1 = mblk[mhi];
    const int b0 = mesh_bitem_off[mlo], b1 = mesh_bitem_off[mhi];
    const int t0 = mesh_tile_off[mlo], t1 = mesh_tile_off[mhi];
    int e = 0;
    auto mark = [&]() {
      if (ev) ECHECK(hipEventRecord(ev[e++], st));
    };
    for (int k = 0; k < 2; ++k) {
      mark();
      if (p1 > p0 && fused)
        k_ens_res_bres_owner<<<std::min(p1 - p0, cap[1]), blk[1], res_lds, st>>>(
            d_x, d_q, d_adt, d_res, OA, BA, p0, p1, res_zero, S_at(off_res[k]));
      else if (p1 > p0)
        k_ens_res_owner<<<std::min(p1 - p0, cap[1]), blk[1], res_lds, st>>>(
            d_x, d_q, d_adt, d_res, OA, p0, p1, res_zero, S_at(off_res[k]));
      mark();
      mark();
      if (b1 > b0 && !fused) {
        const int wpb = blk[2] / 64, g = (std::min(b1 - b0, cap[2]) + wpb - 1) / wpb;
        k_ens_bres<<<g, blk[2], 0, st>>>(d_x, d_q, d_adt, d_res, d_bedge, d_bound, d_bmesh,
                                         d_qinf, d_boffsets, d_bentries, d_btargets,
                                         d_bitems, b0, b1, S_at(off_bres[k]));
      }
      mark();
      mark();
      if (t1 > t0)
        (k ? k_chain1 : k_chain0)<<<std::min(t1 - t0, cap[3 * k]), blk[3 * k], 0, st>>>(
            d_x, d_qold, d_q, d_res, d_adt, d_cell, d_crow, d_cbase, d_tiles, d_coff, t0, t1,
            d_psum, d_pmax, d_ploc, S_at(off_chain[k]));
      mark();
    }
    k_ens_finish<<<mhi - mlo, 256, 0, st>>>(d_psum, d_pmax, d_ploc, d_mtoff, mlo, d_slots,
                                            S_at(off_fin));
  };
  auto history = [&]() {
    k_ens_history<<<1, 256, 0, st>>>(d_slots, hist_words, d_hist, d_count, capacity);
  };

  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<hipEvent_t> ev_a, ev_b;
  std::vector<double> member_ms(M, 0.0);
  int timed_steps = 0;
  if (!o->batched && o->member_timing) {
    ev_a.resize(M);
    ev_b.resize(M);
    for (int m = 0; m < M; ++m) {
      ECHECK(hipEventCreate(&ev_a[m]));
      ECHECK(hipEventCreate(&ev_b[m]));
    }
  }
  // autotuning: steps 1..10 run both chains and bres_calc with workgroup
  // size candidate (i - 1) mod 5, so each candidate twice; per kernel, a
  // candidate's time is its faster step (its launches in the step summed),
  // and the fastest candidate wins
  const int cand[5] = {64, 128, 256, 512, 1024};
  const int ntune = (o->batched && o->autotune) ? std::min(10, o->niter) : 0;
  double tune_ms[4][5];
  for (auto &row : tune_ms)
    for (double &v : row) v = INFINITY;
  hipEvent_t tev[12];
  if (ntune)
    for (auto &e : tev) ECHECK(hipEventCreate(&e));
  prologue(0, M);
  ECHECK(hipGetLastError());
  double t0 = 0.0;
  for (int it = 1; it <= o->niter; ++it) {
    if (it == o->warmup + 1) {
      ECHECK(hipStreamSynchronize(st));
      res->setup_seconds = now() - t_setup;
      t0 = now();
    }
    if (it <= ntune) {
      const int ci = (it - 1) % 5;
      blk[0] = blk[2] = blk[3] = cand[ci];
      set_caps();
      step_range(0, M, tev);
      history();
      ECHECK(hipStreamSynchronize(st));
      // event pairs: (res, bres, chain) twice
      const int kind[6] = {1, 2, 0, 1, 2, 3};
      double step_ms[4] = {0.0, 0.0, 0.0, 0.0};
      for (int j = 0; j < 6; ++j) {
        float ms = 0.f;
        ECHECK(hipEventElapsedTime(&ms, tev[2 * j], tev[2 * j + 1]));
        step_ms[kind[j]] += ms;
      }
      for (int k = 0; k < 4; ++k) tune_ms[k][ci] = std::min(tune_ms[k][ci], step_ms[k]);
      if (it == ntune) {
        for (int k : {0, 2, 3}) {
          if (k == 2 && fused) continue;
          int best = 0;
          for (int c = 1; c < 5; ++c)
            if (tune_ms[k][c] < tune_ms[k][best]) best = c;
          blk[k] = cand[best];
        }
        set_caps();
      }
    } else if (o->batched) {
      if (o->graphed) {
        if (!gexec) {
          ECHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
          step_range(0, M);
          history();
          ECHECK(hipStreamEndCapture(st, &graph));
          ECHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        }
        ECHECK(hipGraphLaunch(gexec, st));
      } else {
        step_range(0, M);
        history();
      }
    } else {
      for (int m = 0; m < M; ++m) {
        if (o->member_timing) ECHECK(hipEventRecord(ev_a[m], st));
        step_range(m, m + 1);
        if (o->member_timing) ECHECK(hipEventRecord(ev_b[m], st));
      }
      history();
      if (o->member_timing && it > o->warmup) {
        ECHECK(hipStreamSynchronize(st));
        for (int m = 0; m < M; ++m) {
          float ms = 0.f;
          ECHECK(hipEventElapsedTime(&ms, ev_a[m], ev_b[m]));
          member_ms[m] += ms;
        }
        ++timed_steps;
      }
    }
    ECHECK(hipGetLastError());
  }
  ECHECK(hipStreamSynchronize(st));
  res->seconds = now() - t0;

  // ---- results ----
  std::vector<double> hist((size_t)hist_words * capacity);
  unsigned long long count = 0;
  ECHECK(hipMemcpy(&count, d_count, sizeof(count), hipMemcpyDeviceToHost));
  ECHECK(hipMemcpy(hist.data(), d_hist, hist.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (count != (unsigned long long)o->niter) {
    printf("op2hip ensemble: %llu history records for %d iterations\n", count, o->niter);
    return 5;
  }
  for (int it = 0; it < o->niter; ++it)
    for (int m = 0; m < M; ++m) {
      const double *h = &hist[(size_t)it * hist_words + 3 * m];
      const int nc = coff[m + 1] - coff[m];
      if (res->rms) res->rms[(size_t)it * M + m] = std::sqrt((0.0 + h[0]) / (double)nc);
      // maxerr/errloc start at 0 (airfoil.cpp:265-267): replaced when exceeded
      if (res->maxerr) res->maxerr[(size_t)it * M + m] = h[1] > 0.0 ? h[1] : 0.0;
      if (res->errloc) res->errloc[(size_t)it * M + m] = h[1] > 0.0 ? (int)h[2] : 0;
    }
  if (res->q_out)
    ECHECK(hipMemcpy(res->q_out, d_q, 4 * (size_t)Nc * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<double> busy_ms(M, 0.0);
  if (stamps) {
    // per timed step: member m's latency runs from the step's first item
    // start to m's last item end; its busy time sums, over the step's
    // launches, m's last item end minus its first item start
    std::vector<unsigned long long> ts(2 * (size_t)per_step * capacity);
    ECHECK(hipMemcpy(ts.data(), d_stamp, ts.size() * sizeof(unsigned long long),
                     hipMemcpyDeviceToHost));
    int rate_khz = 0;
    ECHECK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, 0));
    const double tick_ms = 1.0 / (double)std::max(rate_khz, 1);
    // launch segments: offset, member offset table, items
    struct Seg {
      int off;
      const std::vector<int> *moff;
    };
    const std::vector<int> no_items(M + 1, 0);  // bres_calc fused: no items of its own
    const std::vector<int> *bmoff = fused ? &no_items : &mesh_bitem_off;
    const Seg segs[7] = {{off_res[0], &mblk},          {off_bres[0], bmoff},
                         {off_chain[0], &mesh_tile_off}, {off_res[1], &mblk},
                         {off_bres[1], bmoff},           {off_chain[1], &mesh_tile_off},
                         {off_fin, nullptr}};
    std::vector<double> lat(M, 0.0);
    for (int it = o->warmup; it < o->niter; ++it) {
      const unsigned long long *T = &ts[2 * (size_t)it * per_step];
      unsigned long long s0 = ~0ull;
      for (int k = 0; k < per_step; ++k) s0 = std::min(s0, T[2 * k]);
      for (int m = 0; m < M; ++m) {
        unsigned long long last = 0;
        for (const Seg &g : segs) {
          const int i0 = g.moff ? (*g.moff)[m] : m, i1 = g.moff ? (*g.moff)[m + 1] : m + 1;
          if (i1 <= i0) continue;
          unsigned long long b = ~0ull, e = 0;
          for (int i = i0; i < i1; ++i) {
            b = std::min(b, T[2 * (size_t)(g.off + i)]);
            e = std::max(e, T[2 * (size_t)(g.off + i) + 1]);
          }
          busy_ms[m] += (double)(e - b) * tick_ms;
          last = std::max(last, e);
        }
        lat[m] += (double)(last - s0) * tick_ms;
      }
      ++timed_steps;
    }
    for (int m = 0; m < M; ++m) {
      member_ms[m] = lat[m];
      if (timed_steps) busy_ms[m] /= timed_steps;
    }
  }
  if (res->member_step_ms) {
    for (int m = 0; m < M; ++m)
      res->member_step_ms[m] =
          o->member_timing ? (timed_steps ? member_ms[m] / timed_steps : 0.0)
                           : (o->batched ? 1e3 * res->seconds / std::max(1, o->niter - o->warmup)
                                         : 0.0);
  }
  if (res->member_busy_ms)
    for (int m = 0; m < M; ++m) res->member_busy_ms[m] = stamps ? busy_ms[m] : 0.0;
  for (int k = 0; k < 4; ++k) {
    res->blocks[k] = blk[k];
    res->grids[k] = cap[k];
  }
  res->grids[2] = (cap[2] + blk[2] / 64 - 1) / (blk[2] / 64);
  if (fused) res->blocks[2] = res->grids[2] = 0;
  res->fused_bres = fused ? 1 : 0;
  res->dead_res = dead ? 1 : 0;
  if (ntune)
    for (auto e : tev) (void)hipEventDestroy(e);
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  for (auto e : ev_a) (void)hipEventDestroy(e);
  for (auto e : ev_b) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(st);
  for (void *p : keep) (void)hipFree(p);
  return 0;
}
